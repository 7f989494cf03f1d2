"""GPU parity of the frame path (DispBeamColumn2d/3d over Steel02 /
Concrete02 fiber sections) against the CPU oracle through the C-ABI."""
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


def frame3d_pair(nbx=2, nby=2, nst=3):
    fr = meshes.rc_frame_3d(nbx, nby, nst)
    ms = []
    for m in (Model(3, 6), _oracle.oracle_model(3, 6)):
        meshes.build_rc_frame_3d(m, fr)
        m.system("GpuBlockPCG", 1e-14, 100000)
        m.setup()
        ms.append(m)
    return fr, ms


def lateral_field(fr, amp, seed):
    rng = np.random.default_rng(seed)
    z = fr["crds"][:, 2]
    U = np.zeros((len(z), 6))
    U[:, 0] = amp * z / z.max()
    U[:, 1] = 0.3 * amp * (z / z.max()) ** 2
    U[:, 2] = -1e-3 * amp * z / z.max()
    U += rng.normal(0, abs(amp) * 0.05, U.shape) * (z[:, None] > 0)
    U[:, 3:] *= 1e-4
    return U


def test_frame3d_numbering_bit_exact():
    print("case: frame3d numbering", flush=True)
    fr, (g, o) = frame3d_pair()
    assert np.array_equal(g.node_eqns(), o.node_eqns())
    for a, b in zip(g.csr_pattern(), o.csr_pattern()):
        assert np.array_equal(a, b)


def test_frame3d_update_sections_and_assembly():
    print("case: frame3d update/assembly", flush=True)
    fr, (g, o) = frame3d_pair()
    nsec = 5 * (len(fr["columns"]) + len(fr["beams_x"]) + len(fr["beams_y"]))
    for step, amp in enumerate((20.0, 60.0, -35.0)):
        U = lateral_field(fr, amp, step)
        for m in (g, o):
            m.set_trial_disp(U)
            m.update()
        for what in ("strain", "stress", "tangent"):
            assert relerr(g.point_data(what, nsec, 16), o.point_data(what, nsec, 16)) <= RTOL, (what, step)
        for m in (g, o):
            m.set_load_factor(1000.0 * (step + 1))
        assert relerr(g.form_unbalance(), o.form_unbalance()) <= RTOL
        assert relerr(g.form_tangent(), o.form_tangent()) <= RTOL
        for m in (g, o):
            m.commit()
    tag = len(fr["columns"]) + 3
    assert relerr(g.element_tangent(tag), o.element_tangent(tag)) <= RTOL
    assert relerr(g.element_resisting_force(tag), o.element_resisting_force(tag)) <= RTOL


@pytest.mark.parametrize("dim", [2, 3])
def test_frame_update_row_widths(dim, monkeypatch):
    """The displacement-based update in its three layouts: one wavefront per
    section, four sections per wavefront (16-lane rows, section totals by
    butterfly) and three (20-lane rows, totals through LDS), on frames whose
    section counts (130 and 50) leave rows empty in both: section strains,
    resultants and tangents against the oracle to 1e-9 over three loading
    steps, and the committed Steel02 / Concrete02 fiber states across the
    layouts to 1e-12 (the same fiber formulas; the compiler may fuse their
    multiply-adds differently in each kernel, and the section sums' order
    differs)."""
    print(f"case: frame{dim}d row widths", flush=True)
    ndm, ndf = (3, 6) if dim == 3 else (2, 3)
    fr = meshes.rc_frame_3d(2, 1, 2) if dim == 3 else meshes.rc_frame_2d(2, 2)
    build = meshes.build_rc_frame_3d if dim == 3 else meshes.build_rc_frame_2d
    amps = (20.0, 60.0, -35.0)

    def field(amp, seed):
        if dim == 3:
            return lateral_field(fr, amp, seed)
        rng = np.random.default_rng(seed)
        y = fr["crds"][:, 1]
        U = np.zeros((len(y), 3))
        U[:, 0] = amp * y / y.max()
        U[:, 1] = -1e-3 * amp * y / y.max()
        U += rng.normal(0, abs(amp) * 0.05, U.shape) * (y[:, None] > 0)
        U[:, 2] *= 1e-4
        return U

    nsec = 5 * len(meshes.frame_elements(fr)[0])
    assert nsec % 4 and nsec % 3
    states = []
    for width in ("0", "16", "20"):
        monkeypatch.setenv("XG_FRAME_ROWS", width)
        g, o = Model(ndm, ndf), _oracle.oracle_model(ndm, ndf)
        for m in (g, o):
            build(m, fr)
            m.system("GpuBlockPCG", 1e-14, 100000)
            m.setup()
        for step, amp in enumerate(amps):
            for m in (g, o):
                m.set_trial_disp(field(amp, step))
                m.update()
            for what in ("strain", "stress", "tangent"):
                pg, po = g.point_data(what, nsec, 16), o.point_data(what, nsec, 16)
                assert relerr(pg, po) <= RTOL, (width, what, step)
            for m in (g, o):
                m.commit()
        states.append((g.get_state("Steel02"), g.get_state("Concrete02")))
        for m in (g, o):
            m.close()
    for width, s in zip(("16", "20"), states[1:]):
        for k in range(2):
            e = relerr(s[k], states[0][k])
            print(f"width {width} state {k}: max rel. difference {e:.3g}", flush=True)
            assert e <= 1e-12, (width, k, e)


def build_mixed_frame_3d(m, fr):
    """The frame of rc_frame_3d with two section definitions of mixed fiber
    kinds: columns of core (Concrete02, tag 3) and cover (Concrete02, tag 1)
    concrete, 20 Steel02 bars and four elastic corner plates; beams of an
    8 x 8 concrete patch, 20 bars and two elastic plates."""
    m.nodes(fr["node_tags"], fr["crds"])
    fx = fr["fixed_nodes"]
    m.fix_many(fx, np.ones((len(fx), 6), np.int32))
    m.uniaxialMaterial("Concrete02", 1, *meshes.CONCRETE02)
    m.uniaxialMaterial("Steel02", 2, *meshes.STEEL02)
    m.uniaxialMaterial("Concrete02", 3, -36.0, -0.0025, -25.0, -0.03, 0.1, 3.0, 1500.0)
    m.uniaxialMaterial("Elastic", 4, 2.0e5)
    col = []
    for y, z, A, mat in meshes.rc_section_fibers():
        core = mat == 1 and abs(y) < 150.0 and abs(z) < 150.0
        col.append((y, z, A, 3 if core else mat))
    col += [(sy * 240.0, sz * 240.0, 40.0, 4) for sy in (-1, 1) for sz in (-1, 1)]
    beam = list(meshes.rc_section_fibers(width=400.0, depth=600.0, nc=8))
    beam += [(-290.0, 0.0, 120.0, 4), (290.0, 0.0, 120.0, 4)]
    m.section("Fiber", 1, col, GJ=1.1e14)
    m.section("Fiber", 2, beam, GJ=1.1e14)
    m.geomTransf("Linear", 1, 1.0, 0.0, 0.0)
    m.geomTransf("Linear", 2, 0.0, 0.0, 1.0)
    nc, nx_, ny_ = len(fr["columns"]), len(fr["beams_x"]), len(fr["beams_y"])
    m.elements("dispBeamColumn", np.arange(1, nc + 1), fr["columns"], 1, 1, 5)
    m.elements("dispBeamColumn", np.arange(nc + 1, nc + nx_ + 1), fr["beams_x"], 2, 2, 5)
    m.elements("dispBeamColumn", np.arange(nc + nx_ + 1, nc + nx_ + ny_ + 1), fr["beams_y"], 2, 2, 5)


@pytest.mark.parametrize("width", ["0", "16", "20"])
def test_frame_rows_mixed_sections(width, monkeypatch):
    """Two section definitions (60 and 70 sections: rows left empty at the
    end of each definition's group) with runs of two Concrete02 materials,
    Steel02 and elastic fibers, in each layout of the update: section
    strains, resultants and tangents, the unbalance and the assembled
    tangent against the oracle to 1e-9 over three loading steps, with a
    commit after each."""
    print(f"case: frame3d mixed sections width {width}", flush=True)
    monkeypatch.setenv("XG_FRAME_ROWS", width)
    fr = meshes.rc_frame_3d(2, 1, 2)
    nsec = 5 * (len(fr["columns"]) + len(fr["beams_x"]) + len(fr["beams_y"]))
    g, o = Model(3, 6), _oracle.oracle_model(3, 6)
    for m in (g, o):
        build_mixed_frame_3d(m, fr)
        m.system("GpuBlockPCG", 1e-14, 100000)
        m.setup()
    for step, amp in enumerate((25.0, 70.0, -40.0)):
        U = lateral_field(fr, amp, 10 + step)
        for m in (g, o):
            m.set_trial_disp(U)
            m.update()
        for what in ("strain", "stress", "tangent"):
            assert relerr(g.point_data(what, nsec, 16), o.point_data(what, nsec, 16)) <= RTOL, (what, step)
        for m in (g, o):
            m.set_load_factor(0.0)
        assert relerr(g.form_unbalance(), o.form_unbalance()) <= RTOL, step
        assert relerr(g.form_tangent(), o.form_tangent()) <= RTOL, step
        for m in (g, o):
            m.commit()
    for m in (g, o):
        m.close()


def test_frame3d_revert_matches_reference_semantics():
    print("case: frame3d revert", flush=True)
    fr, (g, o) = frame3d_pair(1, 1, 2)
    nsec = 5 * (len(fr["columns"]) + len(fr["beams_x"]) + len(fr["beams_y"]))
    for m in (g, o):
        m.set_trial_disp(lateral_field(fr, 40.0, 1))
        m.update()
        m.commit()
        m.set_trial_disp(lateral_field(fr, 90.0, 2))
        m.update()
        m.revert()
    for what in ("stress", "tangent"):
        assert relerr(g.point_data(what, nsec, 16), o.point_data(what, nsec, 16)) <= RTOL


def test_frame3d_pushover_trajectory():
    print("case: frame3d pushover", flush=True)
    fr, (g, o) = frame3d_pair(2, 1, 3)
    for m in (g, o):
        m.test("NormUnbalance", 1e-5, 40)
        m.integrator("DisplacementControl", fr["roof"], 1, 6.0)
    for step in range(12):
        for m in (g, o):
            assert m.analyze(1) == 0
        assert relerr(g.trial_disp(), o.trial_disp()) <= RTOL, step
        assert abs(g.load_factor - o.load_factor) <= RTOL * abs(o.load_factor)


def test_frame2d_pushover_trajectory():
    print("case: frame2d pushover", flush=True)
    fr = meshes.rc_frame_2d(4, 2)
    res = []
    for m in (Model(2, 3), _oracle.oracle_model(2, 3)):
        meshes.build_rc_frame_2d(m, fr)
        P = np.zeros((len(fr["floor_nodes"]), 3))
        P[:, 0] = fr["floor_level"] / fr["floor_level"].max()
        m.loads(fr["floor_nodes"], P)
        m.system("GpuBlockPCG", 1e-14, 100000)
        m.test("NormUnbalance", 1e-5, 40)
        m.integrator("DisplacementControl", fr["roof"], 1, 8.0)
        m.setup()
        traj = []
        for step in range(15):
            assert m.analyze(1) == 0
            traj.append((m.trial_disp().copy(), m.load_factor))
        res.append(traj)
    for (ug, lg), (uo, lo) in zip(*res):
        assert relerr(ug, uo) <= RTOL
        assert abs(lg - lo) <= RTOL * abs(lo)


def test_frame3d_fused_pcg_launch_patterns():
    """The node-blocked product that forms the new search direction on the
    fly for six DOFs per node (k_spmv_nb_dot<L,6,true>) with the fused
    alpha update (k_cg_update_fused<6>), taken by 3D frames when the
    one-workgroup solve is off: fused and graph-replayed iterations equal the
    separate launches bit for bit, and the pushover equals the oracle's to
    1e-9."""
    print("case: frame3d fused pcg launch patterns", flush=True)
    fr = meshes.rc_frame_3d(2, 1, 3)
    runs = []
    for graph, fuse in ((True, True), (False, False), (True, False), (False, True)):
        m = Model(3, 6)
        meshes.build_rc_frame_3d(m, fr)
        m.system("GpuBlockPCG", 1e-14, 100000)
        m.cg_config(16, graph, fuse, small=False)
        m.test("NormUnbalance", 1e-5, 40)
        m.integrator("DisplacementControl", fr["roof"], 1, 6.0)
        m.setup()
        t = []
        for step in range(10):
            assert m.analyze(1) == 0
            t.append(m.trial_disp().copy())
        t.append(np.array([m.counters()["cg_iters"]], dtype=float))
        runs.append(t)
        m.close()
    for other in runs[1:]:
        for a, c in zip(runs[0], other):
            assert np.array_equal(a, c)
    o = _oracle.oracle_model(3, 6)
    meshes.build_rc_frame_3d(o, fr)
    o.test("NormUnbalance", 1e-5, 40)
    o.integrator("DisplacementControl", fr["roof"], 1, 6.0)
    o.setup()
    for step in range(10):
        assert o.analyze(1) == 0
        assert relerr(runs[0][step], o.trial_disp()) <= RTOL, step
