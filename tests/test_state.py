"""Committed state in the reference's sendSelf layouts: a model that loads
the state another model saved after n steps continues exactly as that model
(CPU oracle here; the GPU <-> oracle hand-over is in test_gpu_state.py)."""
import numpy as np

from tests import _oracle
from xara_amd import ground_motion, meshes


def frame_model(ag, dt):
    fr = meshes.rc_frame_2d(2, 2)
    m = _oracle.oracle_model(2, 3)
    meshes.define_transient_frame(m, fr, meshes.default_theta(len(meshes.frame_elements(fr)[0])), ag, dt,
                                  tol=1e-12, maxit=40)
    m.setup()
    return m


def test_frame_transient_state_handover():
    dt = 0.01
    ag = ground_motion.kanai_tajimi(2, 200, dt, 1.0)
    a, b = frame_model(ag, dt), frame_model(ag, dt)
    for _ in range(60):
        assert a.analyze(1, dt) == 0
    state = a.save_state()
    assert state["Steel02"].size == 11 * 20 * 5 * 10 and state["Concrete02"].size == 5 * 100 * 5 * 10
    b.load_state(state)
    for k in ("nodes", "Steel02", "Concrete02", "domain"):
        assert np.array_equal(b.get_state(k), state[k])
    for _ in range(40):
        assert a.analyze(1, dt) == 0 and b.analyze(1, dt) == 0
        assert np.array_equal(a.trial_disp(), b.trial_disp())


def test_brick_pushover_state_handover():
    mesh = meshes.brick_cube(3)

    def make():
        m = _oracle.oracle_model(3, 3)
        meshes.build_brick_cube(m, mesh, pressure=1.0)
        m.test("NormUnbalance", 1e-10, 30)
        m.integrator("LoadControl", 60.0)
        m.setup()
        return m

    a, b = make(), make()
    for _ in range(6):  # first yield between 240 and 300
        assert a.analyze(1) == 0
    st = a.save_state()
    assert (st["J2"].reshape(-1, 10)[:, 0] > 0).any()  # plastic points
    eps = st["J2"].reshape(-1, 10)[:, 1:].reshape(-1, 3, 3)
    assert np.array_equal(eps, np.transpose(eps, (0, 2, 1)))  # symmetric tensors
    b.load_state(st)
    # LoadControl starts Newton from the stored element stresses, which the
    # loaded model recomputes from the committed state at its first update:
    # the iterates differ by rounding, the converged steps agree to 1e-12
    for _ in range(2):
        assert a.analyze(1) == 0 and b.analyze(1) == 0
    Ua, Ub = a.trial_disp(), b.trial_disp()
    assert np.max(np.abs(Ua - Ub)) <= 1e-12 * np.max(np.abs(Ua))


def test_revert_to_start_after_plastic_trial_is_a_fresh_model():
    """J2Plasticity::revertToStart zeroes only the plastic state, stress and
    strain (J2Plasticity.cpp:213-224) -- the material keeps its last tangent
    -- but Domain::revertToStart ends with an update at zero displacement
    (Domain.cpp:1941-1958), which forms the elastic tangent again: the model
    is then the fresh model, and a pushover from there is the fresh one."""
    mesh = meshes.brick_cube(3)
    npts = 8 * 27
    fresh, m = (_oracle.oracle_model(3, 3) for _ in range(2))
    for x in (fresh, m):
        meshes.build_brick_cube(x, mesh, pressure=1.0)
        x.setup()
    U = np.zeros((len(mesh["crds"]), 3))
    U[:, 2] = -4e-3 * mesh["crds"][:, 2]  # compression well past first yield
    m.set_trial_disp(U)
    m.update()
    assert (m.point_data("plastic", npts)[:, 6] > 0).all()
    m.revert_to_start()
    for what in ("stress", "tangent", "plastic", "strain"):
        assert np.array_equal(m.point_data(what, npts), fresh.point_data(what, npts)), what
    assert np.array_equal(m.form_tangent(), fresh.form_tangent())
    for x in (fresh, m):
        x.test("NormUnbalance", 1e-8, 30)
        x.integrator("LoadControl", 40.0)
        assert x.analyze(3) == 0
    assert np.array_equal(m.trial_disp(), fresh.trial_disp())


def test_revert_to_last_commit_updates_at_the_committed_state():
    """Domain::revertToLastCommit ends with an update at the committed
    displacements (Domain.cpp:1925-1939): the stresses are those of the
    committed state again, not of the abandoned trial."""
    mesh = meshes.brick_cube(3)
    npts = 8 * 27
    m = _oracle.oracle_model(3, 3)
    meshes.build_brick_cube(m, mesh, pressure=1.0)
    m.setup()
    U = np.zeros((len(mesh["crds"]), 3))
    U[:, 2] = -1.0e-3 * mesh["crds"][:, 2]
    m.set_trial_disp(U)
    m.update()
    m.commit()
    s_commit = m.point_data("stress", npts).copy()
    m.set_trial_disp(3 * U)
    m.update()
    assert not np.allclose(m.point_data("stress", npts), s_commit)
    m.revert()
    assert np.array_equal(m.point_data("stress", npts), s_commit)
    assert np.array_equal(m.trial_disp(), U)
