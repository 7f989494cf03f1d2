"""GPU parity of the wide-band Cholesky preconditioner (xg_wband.hip: the
band of a system of more than 1,024 equations factored in 64 x 64 tiles,
the PCG preconditioned by the exact factor) against the CPU oracle's direct
solve: single solves on a 3D RC frame (half bandwidth above 63, several
tiles per column) and on a quad strip of more than 16,384 equations (one
tile below the diagonal), a DisplacementControl pushover whose two solves
per Newton iteration share the sweeps, graph replay of the factor and
sweep launches against the launches issued one by one, and the fall-back
to block Jacobi where a softened tangent is no longer positive definite."""
import numpy as np
import pytest

from tests import _oracle
from xara_amd import meshes
from xara_amd.model import Model

pytestmark = pytest.mark.gpu
RTOL = 1e-9


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b))
    return 0.0 if s == 0 else float(np.max(np.abs(a - b)) / s)


def lateral_field(fr, amp, seed):
    rng = np.random.default_rng(seed)
    z = fr["crds"][:, 2]
    U = np.zeros((len(z), 6))
    U[:, 0] = amp * z / z.max()
    U[:, 1] = 0.3 * amp * (z / z.max()) ** 2
    U += rng.normal(0, abs(amp) * 0.05, U.shape) * (z[:, None] > 0)
    U[:, 3:] *= 1e-4
    return U


def build(case, m):
    if case == "frame3d":
        fr = meshes.rc_frame_3d(5, 5, 8)
        meshes.build_rc_frame_3d(m, fr)
        return fr
    mesh = meshes.quad_strip(400, 20, lx=80.0, ly=4.0)
    meshes.build_quad_strip(m, mesh)
    return mesh


@pytest.mark.parametrize("case", ["frame3d", "quad"])
def test_wide_band_solve_matches_oracle(case):
    print(f"case: wide band solve {case}", flush=True)
    ndm, ndf = (3, 6) if case == "frame3d" else (2, 2)
    g, o = Model(ndm, ndf), _oracle.oracle_model(ndm, ndf)
    for m in (g, o):
        mesh = build(case, m)
        m.system("GpuBlockPCG", 1e-13, 10000)
        if m is g:
            m.cg_config(16, True, True, wide=True)
        m.setup()
    assert g.neq > 1024 and np.array_equal(g.node_eqns(), o.node_eqns())
    rng = np.random.default_rng(21)
    if case == "frame3d":
        U = lateral_field(mesh, 40.0, 22)
    else:
        U = 1e-3 * rng.normal(size=(g.num_nodes, 2)) * (mesh["crds"][:, :1] > 0)
    for m in (g, o):
        m.set_trial_disp(U)
        m.update()
        m.form_tangent(values=False)
    B = rng.normal(size=g.neq)
    c0 = g.counters()
    Xg = g.solve(B)
    c = g.counters()
    Xo = o.solve(B)
    # backward error through the oracle's own product (the frame's stiffness
    # range puts its rounding floor near 1e-12), and the difference from its
    # direct solve (two backward-stable solves of the slender strip differ by
    # its condition number times the rounding: about 1e-9)
    assert np.linalg.norm(o.spmv(Xg) - B) <= 1e-10 * np.linalg.norm(B)
    assert relerr(Xg, Xo) <= (RTOL if case == "frame3d" else 1e-8)
    bw = c["band_bw"]
    assert bw > (63 if case == "frame3d" else 0), bw
    assert g.neq > 16384 or bw > 63
    # the exact preconditioner: one or two iterations
    assert c["cg_iters"] - c0["cg_iters"] <= 3 * (c["solves"] - c0["solves"])
    for m in (g, o):
        m.close()


def test_wide_band_pushover_matches_oracle_and_graph_replay():
    """DisplacementControl pushover of the 5x5x8 frame: the unbalance and
    reference-load solves of each Newton iteration in one set of sweeps;
    trajectory and load factor against the oracle to 1e-9, and bit for bit
    between graph replay and launches issued one by one."""
    print("case: wide band pushover frame3d 5x5x8", flush=True)
    fr = meshes.rc_frame_3d(5, 5, 8)
    models = []
    for kind in ("graph", "eager", "oracle"):
        m = _oracle.oracle_model(3, 6) if kind == "oracle" else Model(3, 6)
        meshes.build_rc_frame_3d(m, fr)
        m.system("GpuBlockPCG", 1e-12, 10000)
        m.test("NormUnbalance", 1e-3, 40)  # load factors of 1e4..1e5: tighter norms stall at rounding
        m.integrator("DisplacementControl", fr["roof"], 1, 0.5 * fr["height"] / 300.0)
        if kind != "oracle":
            m.cg_config(16, kind == "graph", True, wide=True)
        m.setup()
        models.append(m)
    g, e, o = models
    for step in range(8):
        for m in models:
            assert m.analyze(1) == 0, step
        assert relerr(g.trial_disp(), o.trial_disp()) <= RTOL, step
        assert abs(g.load_factor - o.load_factor) <= RTOL * abs(o.load_factor)
        assert np.array_equal(g.trial_disp(), e.trial_disp()), step
    c = g.counters()
    assert c["band_bw"] > 63
    assert c["cg_iters"] <= 3 * c["solves"]
    for m in models:
        m.close()


def test_wide_band_falls_back_past_the_peak():
    """A DisplacementControl pushover of the 5x5x8 frame through its peak
    (0.25% roof drift per step): where the softened tangent is no longer
    positive definite the band factor meets a pivot that is not positive and
    that matrix is solved with block Jacobi instead (counted in
    band_fallbacks); the trajectory stays with the oracle's direct solve to
    1e-9 through and past the peak."""
    print("case: wide band fallback frame3d 5x5x8", flush=True)
    fr = meshes.rc_frame_3d(5, 5, 8)
    models = []
    for kind in ("gpu", "oracle"):
        m = _oracle.oracle_model(3, 6) if kind == "oracle" else Model(3, 6)
        meshes.build_rc_frame_3d(m, fr)
        m.system("GpuBlockPCG", 1e-12, 20000)
        m.test("NormUnbalance", 1e-3, 40)
        m.integrator("DisplacementControl", fr["roof"], 1, 0.25 * fr["height"] / 100.0)
        m.setup()
        models.append(m)
    g, o = models
    lam = []
    for step in range(20):
        for m in models:
            assert m.analyze(1) == 0, step
        assert relerr(g.trial_disp(), o.trial_disp()) <= RTOL, step
        assert abs(g.load_factor - o.load_factor) <= RTOL * abs(o.load_factor)
        lam.append(o.load_factor)
    assert max(lam) > lam[-1]  # past the peak
    assert g.counters()["band_fallbacks"] > 0
    for m in models:
        m.close()


def test_narrow_band_falls_back_past_the_peak():
    """The same fall-back for the one-workgroup band factor (half bandwidth
    11): a 4-story, 2-bay 2D RC frame pushed by DisplacementControl through
    its peak; the trajectory stays with the oracle to 1e-9."""
    print("case: narrow band fallback frame2d 4x2", flush=True)
    fr = meshes.rc_frame_2d(4, 2)
    models = []
    for kind in ("gpu", "oracle"):
        m = _oracle.oracle_model(2, 3) if kind == "oracle" else Model(2, 3)
        meshes.build_rc_frame_2d(m, fr)
        P = np.zeros((len(fr["floor_nodes"]), 3))
        P[:, 0] = fr["floor_level"] / fr["floor_level"].max()
        m.loads(fr["floor_nodes"], P)
        m.system("GpuBlockPCG", 1e-12, 100000)
        m.test("NormUnbalance", 1e-5, 40)
        m.integrator("DisplacementControl", fr["roof"], 1, 25.0)
        m.setup()
        models.append(m)
    g, o = models
    lam = []
    for step in range(30):
        for m in models:
            assert m.analyze(1) == 0, step
        assert relerr(g.trial_disp(), o.trial_disp()) <= RTOL, step
        lam.append(o.load_factor)
    c = g.counters()
    assert max(lam) > lam[-1] and c["band_fallbacks"] > 0, (max(lam), lam[-1], c["band_fallbacks"])
    for m in models:
        m.close()
