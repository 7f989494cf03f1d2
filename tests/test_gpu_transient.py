"""GPU parity of the transient path (Newmark average acceleration,
UniformExcitation with a PathSeries record, Rayleigh damping) on nonlinear
RC frames, against the CPU oracle."""
import numpy as np
import pytest

from tests import _oracle
from xara_amd import ground_motion, meshes
from xara_amd.model import Model

pytestmark = pytest.mark.gpu
RTOL = 1e-9


def relerr(a, b):
    s = np.max(np.abs(b))
    return 0.0 if s == 0 else float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / s)


def build(m, fr, ag, dt, rayleigh=(0.35, 0.0, 0.0)):
    meshes.build_rc_frame_2d(m, fr)
    mass = np.zeros((len(fr["floor_nodes"]), 3))
    mass[:, 0] = mass[:, 1] = 60.0
    m.masses(fr["floor_nodes"], mass)
    m.uniformExcitation(1, ag, dt)
    m.rayleigh(rayleigh[0], rayleigh[1], 0.0, rayleigh[2])
    m.system("GpuBlockPCG", 1e-14, 100000)
    m.test("NormDispIncr", 1e-9, 30)
    m.integrator("Newmark", 0.5, 0.25)
    m.setup()


@pytest.mark.parametrize("nst,nbay,seed,pga", [(3, 2, 0, 0.5), (5, 1, 1, 0.6)])
def test_frame2d_newmark_trajectory(nst, nbay, seed, pga):
    print(f"case: frame2d newmark {nst}x{nbay} seed {seed}", flush=True)
    fr = meshes.rc_frame_2d(nst, nbay)
    dt = 0.01
    ag = ground_motion.kanai_tajimi(seed, 400, dt, pga)
    g, o = Model(2, 3), _oracle.oracle_model(2, 3)
    for m in (g, o):
        build(m, fr, ag, dt)
    for step in range(120):
        for m in (g, o):
            assert m.analyze(1, dt) == 0, step
        assert relerr(g.trial_disp(), o.trial_disp()) <= RTOL, step
    # the record must have driven the frame past first yield somewhere
    nsec = 5 * (len(fr["columns"]) + len(fr["beams"]))
    ks = o.point_data("tangent", nsec, 4)
    assert ks[:, 0].min() < 0.95 * ks[:, 0].max()


@pytest.mark.parametrize("alphaM,betaK,betaKc", [(0.3, 5e-4, 0.0), (0.2, 0.0, 2e-3), (0.0, 3e-4, 1e-3)])
def test_frame2d_newmark_rayleigh_stiffness_damping(alphaM, betaK, betaKc):
    """Rayleigh damping on the current tangent (betaK) and on the tangent
    of the last commit (betaKcomm): Element::getRayleighDampingForces
    (Element.cpp:283-323) in the resisting force and c2 (betaK K + betaKc
    Kc) in the Newmark tangent (Newmark.cpp:262-289), on a frame driven past
    yield, against the oracle to 1e-9 over 120 steps.  (The damping force on
    the current tangent is not differentiated in the Newton tangent, so
    larger betaK can stall the iteration on both sides; these factors stay
    clear of that.)"""
    print(f"case: frame2d newmark rayleigh {alphaM} {betaK} {betaKc}", flush=True)
    fr = meshes.rc_frame_2d(3, 2)
    dt = 0.01
    ag = ground_motion.kanai_tajimi(0, 400, dt, 0.5)
    g, o = Model(2, 3), _oracle.oracle_model(2, 3)
    for m in (g, o):
        build(m, fr, ag, dt, (alphaM, betaK, betaKc))
    for step in range(120):
        for m in (g, o):
            assert m.analyze(1, dt) == 0, step
        assert relerr(g.trial_disp(), o.trial_disp()) <= RTOL, step
    nsec = 5 * (len(fr["columns"]) + len(fr["beams"]))
    ks = o.point_data("tangent", nsec, 4)
    assert ks[:, 0].min() < 0.95 * ks[:, 0].max()


def test_rayleigh_initial_stiffness_is_refused():
    for m in (Model(2, 3), _oracle.oracle_model(2, 3)):
        with pytest.raises(Exception, match="betaKinit"):
            m.rayleigh(0.1, 0.0, 1e-3, 0.0)
