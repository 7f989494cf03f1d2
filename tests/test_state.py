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


def test_j2_revert_to_start_keeps_last_tangent():
    """J2Plasticity::revertToStart only zeroes the plastic state, stress and
    strain (J2Plasticity.cpp:213-224, :532-543): the tangent of the last
    setTrialStrain stays.  After a plastic trial and revertToStart the
    material tangent is that plastic tangent (not elastic, not zero) while
    stress and plastic state are zero; the next formTangent uses it."""
    mesh = meshes.brick_cube(3)
    npts = 8 * 27
    m = _oracle.oracle_model(3, 3)
    meshes.build_brick_cube(m, mesh, pressure=1.0)
    m.setup()
    D0 = m.point_data("tangent", npts).copy()
    crd = mesh["crds"]
    U = np.zeros((len(crd), 3))
    U[:, 2] = -4e-3 * crd[:, 2]  # compression well past first yield
    m.set_trial_disp(U)
    m.update()
    assert (m.point_data("plastic", npts)[:, 6] > 0).all()
    Dp = m.point_data("tangent", npts).copy()
    Kp = m.form_tangent()
    assert np.abs(Dp - D0).max() > 1e-3 * np.abs(D0).max()
    m.revert_to_start()
    assert np.array_equal(m.point_data("tangent", npts), Dp)
    assert not m.point_data("stress", npts).any() and not m.point_data("plastic", npts).any()
    assert not m.trial_disp().any()
    # the system tangent of the next Newton step is assembled from it
    assert np.array_equal(m.form_tangent(), Kp)
