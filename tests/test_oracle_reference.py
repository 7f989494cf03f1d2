"""The CPU oracle against the reference's own code: the committed golden
vectors (tests/golden/reference_vectors.npz, written by make_golden.py from
the reference sources) and, where oracle/_ref is built, fresh seeded paths
through the reference library itself."""
import os

import numpy as np
import pytest

from tests.golden.make_golden import J2, J2_SAT, J2_FAIL, STEEL, STEEL_ISO, CONC, dp, ip, graph_cases, j2_fail_path

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD, allow_pickle=False)


def close(a, b, tol=1e-13):
    a, b = np.asarray(a), np.asarray(b)
    s = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) / s <= tol


@pytest.mark.parametrize("name,prm", [("j2", J2), ("j2sat", J2_SAT)])
@pytest.mark.parametrize("kind", ["3d", "ps"])
def test_j2_golden(orc, gold, name, prm, kind):
    key = f"{name}_{kind}"
    eps = np.ascontiguousarray(gold[key + "_eps"])
    act = np.ascontiguousarray(gold[key + "_act"])
    n, w = eps.shape
    s, t = np.zeros((n, w)), np.zeros((n, w * w))
    fn = orc.j2_3d_path if kind == "3d" else orc.j2_ps_path
    fn(dp(prm), n, dp(eps), ip(act), dp(s), dp(t))
    assert close(s, gold[key + "_stress"]) and close(t, gold[key + "_tangent"])
    # the path must actually exercise the plastic branch
    assert not np.allclose(t[-1], t[0])


@pytest.mark.parametrize("kind", ["3d", "ps"])
def test_j2_return_map_failure_golden(orc, gold, kind):
    """The local Newton iteration on the plastic multiplier exceeding 25
    iterations (J2Plasticity.cpp:309-313): the integrator returns before
    touching the trial state, so stress and tangent stay those of the
    previous trial strain; J2ThreeDimensional::setTrialStrain passes the -1
    on (:141), J2PlaneStrain::setTrialStrain returns 0 regardless (:133-135).
    Every third step of the recorded path fails (far from the iteration
    limit), the others converge, elastic and plastic."""
    key = f"j2fail_{kind}"
    eps = np.ascontiguousarray(gold[key + "_eps"])
    act = np.ascontiguousarray(gold[key + "_act"])
    n, w = eps.shape
    s, t = np.zeros((n, w)), np.zeros((n, w * w))
    fn = orc.j2_3d_path if kind == "3d" else orc.j2_ps_path
    bad = fn(dp(gold["j2fail_prm"]), n, dp(eps), ip(act), dp(s), dp(t))
    assert close(s, gold[key + "_stress"]) and close(t, gold[key + "_tangent"])
    assert bad == int(gold[key + "_bad"][0]) == (n // 3 if kind == "3d" else 0)
    fail = np.arange(2, n, 3)
    assert np.array_equal(s[fail], s[fail - 1]) and np.array_equal(t[fail], t[fail - 1])
    ok = np.arange(1, n, 3)
    assert not np.allclose(t[ok], t[ok - 1])  # the converged steps are plastic


@pytest.mark.parametrize("name,kind,prm", [("steel", 1, STEEL), ("steeliso", 1, STEEL_ISO), ("conc", 2, CONC)])
def test_uniaxial_golden(orc, gold, name, kind, prm):
    eps = np.ascontiguousarray(gold[name + "_eps"])
    act = np.ascontiguousarray(gold[name + "_act"])
    n = len(eps)
    s, t = np.zeros(n), np.zeros(n)
    orc.uniaxial_path(kind, dp(np.ascontiguousarray(gold[name + "_prm"])), n, dp(eps), ip(act), dp(s), dp(t))
    assert close(s, gold[name + "_stress"]) and close(t, gold[name + "_tangent"])


@pytest.mark.parametrize("dim", [2, 3])
def test_fiber_section_golden(orc, gold, dim):
    k = f"sec{dim}d_"
    g = {n[len(k):]: gold[n] for n in gold.files if n.startswith(k)}
    order = 4 if dim == 3 else 2
    n = len(g["dfm"])
    so, ko = np.zeros((n, order)), np.zeros((n, order * order))
    orc.fiber_section_path(dim == 3, len(g["kinds"]), ip(np.ascontiguousarray(g["kinds"])),
                           dp(np.ascontiguousarray(g["prm"])), dp(g["y"]), dp(g["z"]), dp(g["A"]), 1e12, 1, n,
                           dp(np.ascontiguousarray(g["dfm"])), ip(np.ascontiguousarray(g["act"])), dp(so), dp(ko))
    assert close(so, g["s"]) and close(ko, g["ks"])


@pytest.mark.parametrize("case", [0, 1])
def test_rcm_golden(orc, gold, case):
    ptr, adj, order = gold[f"rcm{case}_ptr"], gold[f"rcm{case}_adj"], gold[f"rcm{case}_order"]
    out = np.zeros(len(order), np.int32)
    orc.rcm(len(order), ip(np.ascontiguousarray(ptr)), ip(np.ascontiguousarray(adj)), ip(out))
    assert np.array_equal(out, order)


# ---- live comparison with the reference library (build container only) ----

@pytest.mark.parametrize("seed", [0, 1, 2])
def test_j2_live_reference(orc, ref, seed):
    rng = np.random.default_rng(100 + seed)
    prm = np.array([166.7e3, 76.9e3, 250.0, 300.0 + 50 * seed, 5.0 * seed, 1500.0, 0.0])
    for w, fo, fr in ((6, orc.j2_3d_path, ref.j2_3d_path), (3, orc.j2_ps_path, ref.j2_ps_path)):
        n = 150
        eps = np.ascontiguousarray(np.cumsum(rng.normal(0, 1.5e-3, (n, w)), axis=0))
        act = (rng.random(n) < 0.5).astype(np.int32)
        out = []
        for f in (fo, fr):
            s, t = np.zeros((n, w)), np.zeros((n, w * w))
            f(dp(prm), n, dp(eps), ip(act), dp(s), dp(t))
            out.append((s, t))
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("seed", [3, 4])
def test_j2_return_map_failure_live_reference(orc, ref, seed):
    for w, fo, fr in ((6, orc.j2_3d_path, ref.j2_3d_path), (3, orc.j2_ps_path, ref.j2_ps_path)):
        eps, act = j2_fail_path(w, seed)
        n = len(eps)
        out = []
        for f in (fo, fr):
            s, t = np.zeros((n, w)), np.zeros((n, w * w))
            bad = f(dp(J2_FAIL), n, dp(eps), ip(act), dp(s), dp(t))
            out.append((s, t, bad))
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
        assert out[0][2] == out[1][2] == (n // 3 if w == 6 else 0)


@pytest.mark.parametrize("kind,prm,amp,shift", [(1, STEEL, 0.015, 0.0), (1, STEEL_ISO, 0.02, 0.001),
                                                (2, CONC, 0.004, -0.0015)])
def test_uniaxial_live_reference(orc, ref, kind, prm, amp, shift):
    rng = np.random.default_rng(7)
    n = 500
    eps = np.ascontiguousarray(np.cumsum(rng.normal(0, amp / 8, n)) + shift)
    eps[10] = eps[9]
    act = rng.integers(0, 3, n).astype(np.int32)
    out = []
    for f in (orc.uniaxial_path, ref.uniaxial_path):
        s, t = np.zeros(n), np.zeros(n)
        f(kind, dp(prm), n, dp(eps), ip(act), dp(s), dp(t))
        out.append((s, t))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_rcm_live_reference(orc, ref):
    rng = np.random.default_rng(3)
    for trial in range(5):
        nv = 40 + 30 * trial
        adj = [set() for _ in range(nv)]
        for _ in range(nv * 2):
            a, b = rng.integers(0, nv, 2)
            if a != b:
                adj[a].add(b)
                adj[b].add(a)
        ptr = np.zeros(nv + 1, np.int32)
        flat = []
        for v in range(nv):
            flat += sorted(adj[v])
            ptr[v + 1] = len(flat)
        flat = np.array(flat, np.int32)
        o1, o2 = np.zeros(nv, np.int32), np.zeros(nv, np.int32)
        orc.rcm(nv, ip(ptr), ip(flat), ip(o1))
        ref.rcm(nv, ip(ptr), ip(flat), ip(o2))
        assert np.array_equal(o1, o2)


def test_golden_graphs_cover_components():
    # the fixture graphs contain an isolated vertex and a second component,
    # the restart branch of RCM::number (RCM.cpp:225-237)
    for adj in graph_cases():
        assert any(len(a) == 0 for a in adj)
