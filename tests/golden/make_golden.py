"""Write tests/golden/reference_vectors.npz from the reference's own sources.

Runs in the build container only: it loads oracle/_ref/libxararef.so (built
by `make -C oracle ref` from the unmodified reference material, section and
graph translation units) and records, for seeded strain paths and graphs,
the stresses, tangents, section resultants, RCM orders and whole-element
tangents and resisting forces (Brick, FourNodeQuad, DispBeamColumn2d/3d)
the reference computes.  The fixtures are data only (inputs and expected outputs).
Usage: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from tests import _oracle  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_vectors.npz")

J2 = np.array([166.7e3, 76.9e3, 250.0, 250.0, 0.0, 2000.0, 0.0])
J2_SAT = np.array([166.7e3, 76.9e3, 250.0, 400.0, 12.0, 500.0, 0.0])
STEEL = np.array([420.0, 200e3, 0.01, 18.0, 0.925, 0.15, 0.0, 1.0, 0.0, 1.0, 0.0, 0.0])
STEEL_ISO = np.array([420.0, 200e3, 0.02, 15.0, 0.925, 0.15, 0.05, 1.0, 0.05, 1.0, 0.0, 0.0])
CONC = np.array([-30.0, -0.002, -6.0, -0.02, 0.1, 3.0, 1500.0, 0, 0, 0, 0, 0])
# Saturation softening so steep (delta 3000, sigma_inf 50 < sigma_0 250) that
# the local Newton iteration on the plastic multiplier needs more than 25
# iterations once the trial deviatoric stress norm passes about 2,500
# (J2Plasticity.cpp:309-313): from about 3,000 on it is still far from its
# root at the 26th iteration, while norms up to about 2,000 converge in at
# most 22 iterations.
J2_FAIL = np.array([166.7e3, 76.9e3, 250.0, 50.0, 3000.0, 2000.0, 0.0])


def _p(a, t):
    import ctypes as C
    return a.ctypes.data_as(C.POINTER(t))


def dp(a):
    import ctypes as C
    return _p(a, C.c_double)


def ip(a):
    import ctypes as C
    return _p(a, C.c_int)


def j2_case(ref, prm, n, w, seed, scale):
    rng = np.random.default_rng(seed)
    eps = np.ascontiguousarray(np.cumsum(rng.normal(0, scale, (n, w)), axis=0))
    act = (rng.random(n) < 0.6).astype(np.int32)
    s = np.zeros((n, w))
    t = np.zeros((n, w * w))
    fn = ref.j2_3d_path if w == 6 else ref.j2_ps_path
    fn(dp(prm), n, dp(eps), ip(act), dp(s), dp(t))
    return eps, act, s, t


def j2_fail_path(w, seed):
    """Strain path for J2_FAIL that keeps the committed hardening variable
    at zero (the failure needs the steep start of the saturation curve):
    steps cycle through an elastic state (trial deviatoric stress norm
    50-150, committed), a converging plastic one (300-1,500) and a failing
    one (3,500-4,500), each a random direction with a small volumetric part.
    Failures are decided far from the 25-iteration limit."""
    rng = np.random.default_rng(seed)
    n = 24
    lo = np.select([np.arange(n) % 3 == 0, np.arange(n) % 3 == 1], [50.0, 300.0], 3500.0)
    hi = np.select([np.arange(n) % 3 == 0, np.arange(n) % 3 == 1], [150.0, 1500.0], 4500.0)
    targets = rng.uniform(lo, hi)
    G = J2_FAIL[1]
    eps = np.zeros((n, w))
    for i in range(n):
        d = rng.normal(size=w)
        if w == 6:
            d[:3] -= d[:3].mean()
            # tensor norm of the engineering-shear vector: shears counted twice, halved
            nrm = np.sqrt(np.sum(d[:3] ** 2) + 2 * np.sum((0.5 * d[3:]) ** 2))
        else:
            # plane strain: eps22 = 0; deviator of (e00, e11, 0) plus the shear
            t = np.array([d[0], d[1], 0.0])
            t -= t.mean()
            nrm = np.sqrt(np.sum(t ** 2) + 2 * (0.5 * d[2]) ** 2)
        eps[i] = d * targets[i] / (2 * G * nrm)
    act = (np.arange(n) % 3 == 0).astype(np.int32)
    return np.ascontiguousarray(eps), act


def uniaxial_case(ref, kind, prm, n, seed, amp, shift=0.0):
    rng = np.random.default_rng(seed)
    x = np.linspace(0, 6 * np.pi, n)
    eps = np.ascontiguousarray(amp * np.sin(x) * (1 + x / (6 * np.pi)) + rng.normal(0, amp * 0.02, n) + shift)
    eps[5] = eps[4]  # exercise the zero-increment branches
    act = rng.integers(0, 3, n).astype(np.int32)
    s = np.zeros(n)
    t = np.zeros(n)
    ref.uniaxial_path(kind, dp(prm), n, dp(eps), ip(act), dp(s), dp(t))
    return eps, act, s, t


def section_case(ref, is3d, seed):
    rng = np.random.default_rng(seed)
    # 500 x 500 RC section: 10 x 10 concrete patch + 20 bars
    yy, zz = np.meshgrid(np.linspace(-225, 225, 10), np.linspace(-225, 225, 10))
    y = list(yy.ravel())
    z = list(zz.ravel())
    A = [2500.0] * 100
    kinds = [2] * 100
    bars = np.linspace(-200, 200, 10)
    for b in bars:
        y += [b, b]
        z += [-200.0, 200.0]
        A += [314.0, 314.0]
        kinds += [1, 1]
    y, z, A = np.array(y), np.array(z), np.array(A)
    kinds = np.array(kinds, np.int32)
    prm = np.zeros((len(kinds), 12))
    prm[kinds == 1] = STEEL
    prm[kinds == 2] = CONC
    order = 4 if is3d else 2
    n = 80
    sc = np.array([6e-4, 8e-6, 8e-6, 1e-5])[:order]
    dfm = np.ascontiguousarray(np.cumsum(rng.normal(0, 1, (n, order)) * sc, axis=0) - np.r_[3e-4, 0, 0, 0][:order])
    act = rng.integers(0, 3, n).astype(np.int32)
    so = np.zeros((n, order))
    ko = np.zeros((n, order * order))
    ref.fiber_section_path(1 if is3d else 0, len(kinds), ip(kinds), dp(np.ascontiguousarray(prm)), dp(y), dp(z),
                           dp(A), 1e12, 1, n, dp(dfm), ip(act), dp(so), dp(ko))
    return dict(kinds=kinds, prm=prm, y=y, z=z, A=A, dfm=dfm, act=act, s=so, ks=ko)


def graph_cases():
    cases = []
    # 2D quad grid with an isolated vertex and a second component
    for (nx, ny) in ((5, 4), (12, 7)):
        nv = (nx + 1) * (ny + 1)
        adj = [set() for _ in range(nv + 5)]
        for i in range(nx):
            for j in range(ny):
                q = [i + (nx + 1) * j, i + 1 + (nx + 1) * j, i + 1 + (nx + 1) * (j + 1), i + (nx + 1) * (j + 1)]
                for a in q:
                    for b in q:
                        if a != b:
                            adj[a].add(b)
        # component {nv+1..nv+4} chain, vertex nv isolated
        for a in range(nv + 1, nv + 4):
            adj[a].add(a + 1)
            adj[a + 1].add(a)
        cases.append(adj)
    return cases


def rcm_case(ref, adj):
    nv = len(adj)
    ptr = np.zeros(nv + 1, np.int32)
    flat = []
    for v in range(nv):
        flat += sorted(adj[v])
        ptr[v + 1] = len(flat)
    flat = np.array(flat, np.int32)
    order = np.zeros(nv, np.int32)
    ref.rcm(nv, ip(ptr), ip(flat), ip(order))
    return ptr, flat, order


# ---- whole elements (Brick, FourNodeQuad, DispBeamColumn2d/3d) ----------
BRICK_CRD = np.array([[0, 0, 0], [1.1, 0.05, 0], [1.0, 1.2, 0.1], [-0.1, 0.9, 0], [0.05, 0, 1.0],
                      [1.0, -0.1, 1.1], [1.2, 1.0, 0.9], [0, 1.1, 1.05]], float)
QUAD_CRD = np.array([[0.0, 0.0], [1.3, 0.1], [1.1, 0.9], [-0.2, 1.2]])
BEAM2D_CRD = np.array([[0.0, 0.0], [3000.0, 400.0]])
BEAM3D_CRD = np.array([[0.0, 0.0, 0.0], [500.0, 300.0, 3000.0]])
BEAM3D_VECXZ = np.array([1.0, 0.2, 0.0])
# nodes, DOFs per node, per-DOF step scale of the random-walk paths
ELEMENT_CASES = {"brick": (8, 3), "quad": (4, 2), "beam2d": (2, 3), "beam3d": (2, 6)}


def element_fibers():
    """A small asymmetric RC section: 6 x 4 concrete fibers, 3 bars on the
    +z face and 2 on the -z face (kind 1 Steel02, 2 Concrete02)."""
    yy, zz = np.meshgrid(np.linspace(-175, 175, 6), np.linspace(-150, 150, 4))
    y, z = list(yy.ravel()), list(zz.ravel())
    A, kinds = [4800.0] * 24, [2] * 24
    for yb, zb in ((-160, 170), (0, 170), (160, 170), (-160, -170), (160, -170)):
        y.append(float(yb))
        z.append(float(zb))
        A.append(510.0)
        kinds.append(1)
    kinds = np.array(kinds, np.int32)
    prm = np.zeros((len(kinds), 12))
    prm[kinds == 1] = STEEL
    prm[kinds == 2] = CONC
    return kinds, prm, np.array(y), np.array(z), np.array(A)


def element_inputs(case, seed, n=10):
    """Seeded random-walk nodal displacement path (n states) and commits."""
    nen, ndf = ELEMENT_CASES[case]
    rng = np.random.default_rng(1000 + 17 * seed + len(case))
    if case in ("brick", "quad"):
        step = np.full(nen * ndf, 1.2e-3)
    else:
        per = [1.0, 1.0, 2.5e-4] if ndf == 3 else [1.0, 1.0, 1.0, 2.5e-4, 2.5e-4, 2.5e-4]
        step = np.tile(per, nen)
    U = np.ascontiguousarray(np.cumsum(rng.normal(size=(n, nen * ndf)) * step, axis=0))
    act = (rng.random(n) < 0.6).astype(np.int32)
    return U, act


def element_fail_inputs(case, seed, n=12):
    """Nodal displacement path for the J2_FAIL elements, cycling through a
    small elastic state (committed), a moderate plastic one and a large one
    (strains about 0.02: trial stress norms past the failure threshold at
    most Gauss points), so the committed hardening variable stays zero."""
    nen, ndf = ELEMENT_CASES[case]
    rng = np.random.default_rng(2000 + seed + len(case))
    U = np.zeros((n, nen * ndf))
    for k in range(n):
        scale = (3e-4, 4e-3, 2.5e-2)[k % 3]
        U[k] = rng.normal(size=nen * ndf) * scale
    act = (np.arange(n) % 3 == 0).astype(np.int32)
    return np.ascontiguousarray(U), act


DAMP = (0.0, 2.0e-3, 3.0e-3)  # alphaM, betaK, betaKcomm of the damping paths


def run_ref_damp(ref, case, U, act, alphaM=DAMP[0], betaK=DAMP[1], betaKc=DAMP[2]):
    """Element::getDamp of the reference's element along the path, the
    Rayleigh factors set right after the element joins its domain."""
    nen, ndf = ELEMENT_CASES[case]
    n, nd = len(U), nen * ndf
    D = np.zeros((n, nd, nd))
    kinds, prm, y, z, A = element_fibers()
    kind = {"brick": 1, "quad": 2, "beam2d": 3, "beam3d": 4}[case]
    crd = {"brick": BRICK_CRD, "quad": QUAD_CRD, "beam2d": BEAM2D_CRD, "beam3d": BEAM3D_CRD}[case]
    rc = ref.element_damp_path(kind, dp(np.ascontiguousarray(crd.ravel())), dp(BEAM3D_VECXZ), dp(J2), len(kinds),
                               ip(kinds), dp(np.ascontiguousarray(prm)), dp(y), dp(z), dp(A), alphaM, betaK,
                               betaKc, n, dp(np.ascontiguousarray(U)), ip(np.ascontiguousarray(act, np.int32)),
                               dp(D))
    if rc != 0:
        raise RuntimeError(f"reference {case} damping path: {rc} failed updates")
    return D


def element_model(factory, case, j2=J2):
    """The single-element model of `case` through the Model command API
    (factory(ndm, ndf) -> Model): nodes 1..nen, element 1."""
    nen, ndf = ELEMENT_CASES[case]
    if case == "brick":
        m = factory(3, 3)
        for a in range(8):
            m.node(a + 1, *BRICK_CRD[a])
        m.nDMaterial("J2Plasticity", 1, *j2[:6])
        m.element("stdBrick", 1, *range(1, 9), 1)
    elif case == "quad":
        m = factory(2, 2)
        for a in range(4):
            m.node(a + 1, *QUAD_CRD[a])
        m.nDMaterial("J2Plasticity", 1, *j2[:6])
        m.element("quad", 1, 1, 2, 3, 4, 1.3, "PlaneStrain", 1)
    else:
        is3d = case == "beam3d"
        m = factory(3, 6) if is3d else factory(2, 3)
        crd = BEAM3D_CRD if is3d else BEAM2D_CRD
        for a in range(2):
            m.node(a + 1, *crd[a])
        m.uniaxialMaterial("Steel02", 1, *STEEL[:11])
        m.uniaxialMaterial("Concrete02", 2, *CONC[:7])
        kinds, _, y, z, A = element_fibers()
        m.section("Fiber", 1, [(y[f], z[f], A[f], int(kinds[f])) for f in range(len(kinds))],
                  GJ=1.0e12 if is3d else None)
        if is3d:
            m.geomTransf("Linear", 1, *BEAM3D_VECXZ)
        else:
            m.geomTransf("Linear", 1)
        m.element("dispBeamColumn", 1, 1, 2, 1, "Legendre", 1, 5)
    return m, nen, ndf


def run_ref_element(ref, case, U, act, j2=J2):
    """The same path through the reference's own element classes."""
    nen, ndf = ELEMENT_CASES[case]
    n, nd = len(U), nen * ndf
    U = np.ascontiguousarray(U, np.float64)
    act = np.ascontiguousarray(act, np.int32)
    K, R = np.zeros((n, nd, nd)), np.zeros((n, nd))
    if case == "brick":
        rc = ref.brick_path(dp(np.ascontiguousarray(BRICK_CRD.ravel())), dp(j2), dp(np.zeros(3)), n, dp(U),
                            ip(act), dp(K), dp(R))
    elif case == "quad":
        rc = ref.quad_path(dp(np.ascontiguousarray(QUAD_CRD.ravel())), dp(j2), 1.3, dp(np.zeros(2)), n, dp(U),
                           ip(act), dp(K), dp(R))
    else:
        is3d = case == "beam3d"
        kinds, prm, y, z, A = element_fibers()
        crd = np.ascontiguousarray((BEAM3D_CRD if is3d else BEAM2D_CRD).ravel())
        rc = ref.dispbeam_path(int(is3d), dp(crd), dp(BEAM3D_VECXZ), 5, len(kinds), ip(kinds),
                               dp(np.ascontiguousarray(prm)), dp(y), dp(z), dp(A), 1.0e12 if is3d else -1.0, 1, n,
                               dp(U), ip(act), dp(K), dp(R))
    if rc != 0:
        raise RuntimeError(f"reference {case} path: {rc} failed updates")
    return K, R


def main():
    ref = _oracle.load_ref()
    if ref is None:
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    d = {}
    d["j2_prm"], d["j2sat_prm"] = J2, J2_SAT
    for name, prm in (("j2", J2), ("j2sat", J2_SAT)):
        for w in (6, 3):
            e, a, s, t = j2_case(ref, prm, 60, w, 11 + w, 1.2e-3)
            key = f"{name}_{'3d' if w == 6 else 'ps'}"
            d[key + "_eps"], d[key + "_act"], d[key + "_stress"], d[key + "_tangent"] = e, a, s, t
    for name, kind, prm, amp, shift in (("steel", 1, STEEL, 0.012, 0.0), ("steeliso", 1, STEEL_ISO, 0.02, 0.0),
                                        ("conc", 2, CONC, 0.003, -0.002)):
        e, a, s, t = uniaxial_case(ref, kind, prm, 240, 5, amp, shift)
        d[name + "_prm"], d[name + "_eps"], d[name + "_act"], d[name + "_stress"], d[name + "_tangent"] = prm, e, a, s, t
    for is3d in (0, 1):
        c = section_case(ref, bool(is3d), 21 + is3d)
        for k, v in c.items():
            d[f"sec{2 + is3d}d_{k}"] = v
    for i, adj in enumerate(graph_cases()):
        ptr, flat, order = rcm_case(ref, adj)
        d[f"rcm{i}_ptr"], d[f"rcm{i}_adj"], d[f"rcm{i}_order"] = ptr, flat, order
    for case in ELEMENT_CASES:
        U, act = element_inputs(case, 7)
        K, R = run_ref_element(ref, case, U, act)
        d[f"ele_{case}_U"], d[f"ele_{case}_act"], d[f"ele_{case}_K"], d[f"ele_{case}_R"] = U, act, K, R
    # Rayleigh damping matrices (betaK on the current, betaKcomm on the
    # committed tangent) along the element paths
    for case in ELEMENT_CASES:
        d[f"damp_{case}_D"] = run_ref_damp(ref, case, d[f"ele_{case}_U"], d[f"ele_{case}_act"])
    # the return-map failure path: material points and whole elements
    d["j2fail_prm"] = J2_FAIL
    for w, key in ((6, "j2fail_3d"), (3, "j2fail_ps")):
        eps, act = j2_fail_path(w, 40 + w)
        s, t = np.zeros((len(eps), w)), np.zeros((len(eps), w * w))
        fn = ref.j2_3d_path if w == 6 else ref.j2_ps_path
        bad = fn(dp(J2_FAIL), len(eps), dp(eps), ip(act), dp(s), dp(t))
        d[key + "_eps"], d[key + "_act"], d[key + "_stress"], d[key + "_tangent"] = eps, act, s, t
        d[key + "_bad"] = np.array([bad], np.int32)
    for case in ("brick", "quad"):
        U, act = element_fail_inputs(case, 9)
        K, R = run_ref_element(ref, case, U, act, J2_FAIL)
        d[f"elefail_{case}_U"], d[f"elefail_{case}_act"] = U, act
        d[f"elefail_{case}_K"], d[f"elefail_{case}_R"] = K, R
    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(d)} arrays)")


if __name__ == "__main__":
    main()
