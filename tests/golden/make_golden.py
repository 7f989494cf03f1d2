"""Write tests/golden/reference_vectors.npz from the reference's own sources.

Runs in the build container only: it loads oracle/_ref/libxararef.so (built
by `make -C oracle ref` from the unmodified reference material, section and
graph translation units) and records, for seeded strain paths and graphs,
the stresses, tangents, section resultants and RCM orders the reference
computes.  The fixtures are data only (inputs and expected outputs).
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
    np.savez_compressed(OUT, **d)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes, {len(d)} arrays)")


if __name__ == "__main__":
    main()
