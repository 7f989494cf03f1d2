"""Structured meshes of the benchmark workloads, as arrays ready for the bulk
model commands.

* ``quad_strip``: 2D plane-strain FourNodeQuad strip nx x ny on lx x ly,
  clamped at x = 0 (the 64 x 16 on 20 x 5 pushover workload).
* ``brick_cube``: 3D stdBrick cube n^3 of unit size, base z = 0 fixed, with
  the consistent nodal loads of a unit uniform top pressure (-z): interior
  h^2, edge h^2/2, corner h^2/4 (the 64^3 J2 workload at n = 64).

Node and element tags are 1-based and lexicographic (x fastest).
"""
import numpy as np

J2_STEEL = dict(K=166.7e3, G=76.9e3, sig0=250.0, sigInf=250.0, delta=0.0, H=2000.0)


def quad_strip(nx=64, ny=16, lx=20.0, ly=5.0):
    i, j = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="xy")
    node_tags = (1 + i + (nx + 1) * j).ravel().astype(np.int32)
    crds = np.stack([lx * i / nx, ly * j / ny], axis=-1).reshape(-1, 2).astype(np.float64)
    ei, ej = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ei, ej = ei.ravel(), ej.ravel()
    n0 = 1 + ei + (nx + 1) * ej
    conn = np.stack([n0, n0 + 1, n0 + 1 + (nx + 1), n0 + (nx + 1)], axis=1).astype(np.int32)
    ele_tags = (1 + ei + nx * ej).astype(np.int32)
    fixed = node_tags[(i.ravel() == 0)]
    tip = int(1 + nx + (nx + 1) * ny)  # free-end top node
    return dict(node_tags=node_tags, crds=crds, ele_tags=ele_tags, conn=conn,
                fixed_nodes=fixed.astype(np.int32), tip=tip)


def brick_cube(n=64, size=1.0):
    m = n + 1
    k, j, i = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    h = size / n
    node_tags = (1 + i + m * j + m * m * k).astype(np.int32)
    crds = np.stack([i * h, j * h, k * h], axis=1).astype(np.float64)
    ek, ej, ei = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    ei, ej, ek = ei.ravel(), ej.ravel(), ek.ravel()
    b = 1 + ei + m * ej + m * m * ek
    conn = np.stack([b, b + 1, b + 1 + m, b + m,
                     b + m * m, b + m * m + 1, b + m * m + 1 + m, b + m * m + m], axis=1).astype(np.int32)
    ele_tags = (1 + ei + n * ej + n * n * ek).astype(np.int32)
    fixed = node_tags[k == 0]
    top = k == n
    w = np.ones(len(i))
    w[(i == 0) | (i == n)] *= 0.5
    w[(j == 0) | (j == n)] *= 0.5
    load_tags = node_tags[top]
    Pz = -(h * h) * w[top]
    return dict(node_tags=node_tags, crds=crds, ele_tags=ele_tags, conn=conn,
                fixed_nodes=fixed.astype(np.int32), load_tags=load_tags.astype(np.int32),
                Pz=Pz.astype(np.float64), h=h)


def build_quad_strip(model, mesh, mat_tag=1, thk=1.0, mat=J2_STEEL):
    model.nodes(mesh["node_tags"], mesh["crds"])
    fx = mesh["fixed_nodes"]
    model.fix_many(fx, np.ones((len(fx), 2), np.int32))
    model.nDMaterial("J2Plasticity", mat_tag, mat["K"], mat["G"], mat["sig0"], mat["sigInf"],
                     mat["delta"], mat["H"])
    model.elements("quad", mesh["ele_tags"], mesh["conn"], thk, mat_tag)


def build_brick_cube(model, mesh, mat_tag=1, mat=J2_STEEL, pressure=1.0):
    model.nodes(mesh["node_tags"], mesh["crds"])
    fx = mesh["fixed_nodes"]
    model.fix_many(fx, np.ones((len(fx), 3), np.int32))
    model.nDMaterial("J2Plasticity", mat_tag, mat["K"], mat["G"], mat["sig0"], mat["sigInf"],
                     mat["delta"], mat["H"])
    model.elements("stdBrick", mesh["ele_tags"], mesh["conn"], mat_tag)
    if pressure:
        P = np.zeros((len(mesh["load_tags"]), 3))
        P[:, 2] = pressure * mesh["Pz"]
        model.loads(mesh["load_tags"], P)


# ---------------------------------------------------------------------------
# Reinforced-concrete frames
# ---------------------------------------------------------------------------
STEEL02 = (420.0, 200e3, 0.01, 18.0, 0.925, 0.15)           # fy E b R0 cR1 cR2
CONCRETE02 = (-30.0, -0.002, -6.0, -0.02, 0.1, 3.0, 1500.0)  # fc epsc0 fcu epscu rat ft Ets


def rc_section_fibers(width=500.0, depth=500.0, cover=50.0, nc=10, bar_area=314.0,
                      conc_tag=1, steel_tag=2):
    """500 x 500 RC section: a 10 x 10 concrete patch (100 Concrete02 fibers)
    and 20 Steel02 bars on the perimeter (6 + 6 on two faces, 4 + 4 on the
    others).  Returns rows (y, z, A, matTag)."""
    dy, dz = depth / nc, width / nc
    rows = []
    for i in range(nc):
        for j in range(nc):
            rows.append((-depth / 2 + (i + 0.5) * dy, -width / 2 + (j + 0.5) * dz, dy * dz, conc_tag))
    yb = depth / 2 - cover
    zb = width / 2 - cover
    for z in np.linspace(-zb, zb, 6):
        rows.append((yb, z, bar_area, steel_tag))
        rows.append((-yb, z, bar_area, steel_tag))
    for y in np.linspace(-yb, yb, 6)[1:-1]:
        rows.append((y, zb, bar_area, steel_tag))
        rows.append((y, -zb, bar_area, steel_tag))
    return rows


def rc_frame_3d(nbx=10, nby=10, nst=20, bay=6000.0, story=3500.0):
    """Space frame: columns (i,j,k)-(i,j,k+1), beams along x and y at every
    floor.  Node tag 1 + i + (nbx+1) j + (nbx+1)(nby+1) k; columns first, then
    x beams, then y beams.  Base nodes fixed.  Returns arrays plus the
    inverted-triangle lateral pattern (x) and the roof control node."""
    mx, my = nbx + 1, nby + 1
    k, j, i = np.meshgrid(np.arange(nst + 1), np.arange(my), np.arange(mx), indexing="ij")
    i, j, k = i.ravel(), j.ravel(), k.ravel()
    node_tags = (1 + i + mx * j + mx * my * k).astype(np.int32)
    crds = np.stack([bay * i, bay * j, story * k], axis=1).astype(np.float64)

    def tag(ii, jj, kk):
        return 1 + ii + mx * jj + mx * my * kk

    cols, bx, by = [], [], []
    for kk in range(nst):
        for jj in range(my):
            for ii in range(mx):
                cols.append((tag(ii, jj, kk), tag(ii, jj, kk + 1)))
    for kk in range(1, nst + 1):
        for jj in range(my):
            for ii in range(nbx):
                bx.append((tag(ii, jj, kk), tag(ii + 1, jj, kk)))
        for jj in range(nby):
            for ii in range(mx):
                by.append((tag(ii, jj, kk), tag(ii, jj + 1, kk)))
    cols, bx, by = (np.array(a, np.int32).reshape(-1, 2) for a in (cols, bx, by))
    fixed = node_tags[k == 0]
    floors = k > 0
    load_tags = node_tags[floors]
    Px = (k[floors] / nst).astype(np.float64)  # inverted triangle
    roof = int(tag(nbx // 2, nby // 2, nst))
    return dict(node_tags=node_tags, crds=crds, columns=cols, beams_x=bx, beams_y=by,
                fixed_nodes=fixed.astype(np.int32), load_tags=load_tags.astype(np.int32), Px=Px,
                roof=roof, height=story * nst)


def build_rc_frame_3d(model, fr, GJ=1.1e14, loads=True, element="dispBeamColumn", nip=5):
    model.nodes(fr["node_tags"], fr["crds"])
    fx = fr["fixed_nodes"]
    model.fix_many(fx, np.ones((len(fx), 6), np.int32))
    model.uniaxialMaterial("Concrete02", 1, *CONCRETE02)
    model.uniaxialMaterial("Steel02", 2, *STEEL02)
    model.section("Fiber", 1, rc_section_fibers(), GJ=GJ)
    model.geomTransf("Linear", 1, 1.0, 0.0, 0.0)  # columns (axis z)
    model.geomTransf("Linear", 2, 0.0, 0.0, 1.0)  # beams
    nc, nx_, ny_ = len(fr["columns"]), len(fr["beams_x"]), len(fr["beams_y"])
    model.elements(element, np.arange(1, nc + 1), fr["columns"], 1, 1, nip)
    model.elements(element, np.arange(nc + 1, nc + nx_ + 1), fr["beams_x"], 2, 1, nip)
    model.elements(element, np.arange(nc + nx_ + 1, nc + nx_ + ny_ + 1), fr["beams_y"], 2, 1, nip)
    if loads:
        P = np.zeros((len(fr["load_tags"]), 6))
        P[:, 0] = fr["Px"]
        model.loads(fr["load_tags"], P)


def rc_frame_2d(nst, nbay, bay=6000.0, story=3500.0, tag0=0):
    """Plane frame: (nbay+1) columns lines, nst stories; node tag
    tag0 + 1 + i + (nbay+1) k; columns then beams.  Base fixed."""
    m = nbay + 1
    k, i = np.meshgrid(np.arange(nst + 1), np.arange(m), indexing="ij")
    i, k = i.ravel(), k.ravel()
    node_tags = (tag0 + 1 + i + m * k).astype(np.int32)
    crds = np.stack([bay * i, story * k], axis=1).astype(np.float64)
    cols = [(tag0 + 1 + ii + m * kk, tag0 + 1 + ii + m * (kk + 1)) for kk in range(nst) for ii in range(m)]
    beams = [(tag0 + 1 + ii + m * kk, tag0 + 2 + ii + m * kk) for kk in range(1, nst + 1) for ii in range(nbay)]
    return dict(node_tags=node_tags, crds=crds, columns=np.array(cols, np.int32),
                beams=np.array(beams, np.int32), fixed_nodes=node_tags[k == 0],
                floor_nodes=node_tags[k > 0], floor_level=k[k > 0], roof=int(tag0 + 1 + m * nst))


def build_rc_frame_2d(model, fr, ele0=1):
    model.nodes(fr["node_tags"], fr["crds"])
    fx = fr["fixed_nodes"]
    model.fix_many(fx, np.ones((len(fx), 3), np.int32))
    model.uniaxialMaterial("Concrete02", 1, *CONCRETE02)
    model.uniaxialMaterial("Steel02", 2, *STEEL02)
    model.section("Fiber", 1, rc_section_fibers())
    model.geomTransf("Linear", 1)
    nc, nb = len(fr["columns"]), len(fr["beams"])
    model.elements("dispBeamColumn", np.arange(ele0, ele0 + nc), fr["columns"], 1, 1, 5)
    model.elements("dispBeamColumn", np.arange(ele0 + nc, ele0 + nc + nb), fr["beams"], 1, 1, 5)


GRAVITY = 9810.0  # mm / s^2


def gravity_masses(fr, floor_load=100.0, bay=6000.0):
    """Lumped floor masses from the gravity load on the beams: each floor
    node carries floor_load (N/mm) over half of each adjacent bay, divided
    by g (N, mm, s: masses in tonnes)."""
    x = dict(zip(fr["node_tags"].tolist(), fr["crds"][:, 0].tolist()))
    xs = np.array([x[int(t)] for t in fr["floor_nodes"]])
    edge = (xs == fr["crds"][:, 0].min()) | (xs == fr["crds"][:, 0].max())
    trib = np.where(edge, 0.5 * bay, bay)
    return floor_load * trib / GRAVITY


def define_transient_rc_frame_2d(model, fr, ag, dt, mass=None, alphaM=0.35, tol=1e-8, maxit=25):
    """Plane RC frame under a UniformExcitation record in the x direction:
    lumped floor masses (x and y; from the gravity load on the beams unless
    a uniform mass is given), mass-proportional damping, Newmark average
    acceleration, NormDispIncr test."""
    build_rc_frame_2d(model, fr)
    m = np.zeros((len(fr["floor_nodes"]), 3))
    m[:, 0] = m[:, 1] = gravity_masses(fr) if mass is None else mass
    model.masses(fr["floor_nodes"], m)
    model.uniformExcitation(1, ag, dt)
    model.rayleigh(alphaM)
    model.test("NormDispIncr", tol, maxit)
    model.integrator("Newmark", 0.5, 0.25)


def frame_population(n, seed=2024):
    """Stories uniform in 3..20 and bays uniform in 1..6 for n frames, and
    the PGA of each (uniform 0.1-0.6 g), all seeded."""
    rng = np.random.Generator(np.random.Philox(seed))
    stories = rng.integers(3, 21, n)
    bays = rng.integers(1, 7, n)
    pga = rng.uniform(0.1, 0.6, n)
    return stories, bays, pga


# ---------------------------------------------------------------------------
# Frames with their own materials per element (parameter studies)
# ---------------------------------------------------------------------------
THETA_NAMES = ("fy", "E", "b", "fc")
STEEL_TAG0, CONC_TAG0 = 100000, 200000


def default_theta(nele):
    """Per-element parameters (fy, E, b of Steel02; fc of Concrete02)."""
    return np.tile(np.array([STEEL02[0], STEEL02[1], STEEL02[2], CONCRETE02[0]]), (nele, 1))


def frame_elements(fr):
    """(connectivity, geomTransf tag) of every element in tag order."""
    if "beams" in fr:
        conn = np.concatenate([fr["columns"], fr["beams"]])
        return conn, np.ones(len(conn), np.int32)
    conn = np.concatenate([fr["columns"], fr["beams_x"], fr["beams_y"]])
    transf = np.full(len(conn), 2, np.int32)
    transf[: len(fr["columns"])] = 1
    return conn, transf


def build_rc_frame_per_element(model, fr, theta, GJ=1.1e14, element="dispBeamColumn", nip=5):
    """RC frame (rc_frame_2d or rc_frame_3d layout, no loads) where element
    e (tag e, 1-based) has its own Steel02 (tag STEEL_TAG0 + e), Concrete02
    (tag CONC_TAG0 + e) and fiber section (tag e) with the parameters of row
    e - 1 of theta: (fy, E, b, fc)."""
    three_d = "beams_x" in fr
    ndf = 6 if three_d else 3
    model.nodes(fr["node_tags"], fr["crds"])
    fx = fr["fixed_nodes"]
    model.fix_many(fx, np.ones((len(fx), ndf), np.int32))
    if three_d:
        model.geomTransf("Linear", 1, 1.0, 0.0, 0.0)
        model.geomTransf("Linear", 2, 0.0, 0.0, 1.0)
    else:
        model.geomTransf("Linear", 1)
    conn, transf = frame_elements(fr)
    theta = np.asarray(theta, np.float64).reshape(len(conn), 4)
    for e in range(len(conn)):
        tag = e + 1
        fy, E, b, fc = theta[e]
        model.uniaxialMaterial("Steel02", STEEL_TAG0 + tag, fy, E, b, *STEEL02[3:])
        model.uniaxialMaterial("Concrete02", CONC_TAG0 + tag, fc, *CONCRETE02[1:])
        model.section("Fiber", tag, rc_section_fibers(conc_tag=CONC_TAG0 + tag, steel_tag=STEEL_TAG0 + tag),
                      GJ=GJ if three_d else None)
        model.elements(element, [tag], conn[e:e + 1], int(transf[e]), tag, nip)


def define_transient_frame(model, fr, theta, ag, dt, mass=60.0, alphaM=0.35, tol=1e-8, maxit=25,
                           element="dispBeamColumn"):
    """Per-element RC frame (2D or 3D) under a UniformExcitation record in x:
    lumped translational floor masses, mass-proportional damping, Newmark
    average acceleration, NormDispIncr test."""
    three_d = "beams_x" in fr
    build_rc_frame_per_element(model, fr, theta, element=element)
    floor = fr["floor_nodes"] if "floor_nodes" in fr else fr["load_tags"]
    ndf = 6 if three_d else 3
    m = np.zeros((len(floor), ndf))
    m[:, : (3 if three_d else 2)] = mass
    model.masses(floor, m)
    model.uniformExcitation(1, ag, dt)
    model.rayleigh(alphaM)
    model.test("NormDispIncr", tol, maxit)
    model.integrator("Newmark", 0.5, 0.25)


def roof_nodes(fr):
    """Tags of the nodes on the top floor."""
    z = fr["crds"][:, -1]
    return fr["node_tags"][z == z.max()].astype(np.int32)
