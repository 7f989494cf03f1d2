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
