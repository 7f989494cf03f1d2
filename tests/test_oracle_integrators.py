"""Pins of the oracle's analysis pieces that the reference build does not
cover (its analysis classes are not compiled): DisplacementControl
(DisplacementControl.cpp:121-415) by its invariants and an elastic closed
form, and CTestEnergyIncr (CTestEnergyIncr.cpp:85-130) by recomputing its
norm from the solve and unbalance vectors by hand.  Each check fails for a
plausible slip: a wrong sign of dLambda, phat formed at the wrong load
factor, the energy formed from the unbalance before the update, or a missing
factor 1/2."""
import numpy as np
import pytest

from tests import _oracle
from xara_amd import meshes


def quad_strip_model(nx=8, ny=2, dc_incr=0.004, tol=1e-8):
    mesh = meshes.quad_strip(nx, ny)
    m = _oracle.oracle_model(2, 2)
    meshes.build_quad_strip(m, mesh)
    m.load(mesh["tip"], 0.0, 1.0)
    m.system("dense")
    m.test("NormUnbalance", tol, 30)
    m.integrator("DisplacementControl", mesh["tip"], 2, dc_incr)
    return m, mesh


def test_displacement_control_holds_control_dof_and_equilibrium():
    """After step k the control DOF has moved k * incr (newStep imposes
    incr, every update keeps the control component of dU at zero), and the
    converged state is in equilibrium with lambda P: the unbalance
    lambda P - R(U) formed afresh is below the test tolerance.  The strip
    goes plastic, so lambda is not proportional to the displacement."""
    incr, tol = 0.02, 1e-8
    m, mesh = quad_strip_model(dc_incr=incr, tol=tol)
    tip = mesh["tip"]
    lams = []
    for k in range(1, 13):
        assert m.analyze(1) == 0
        u = m.nodeDisp(tip, 2)
        assert abs(u - k * incr) <= 1e-12 * k * incr, (k, u)
        B = m.form_unbalance()
        assert np.linalg.norm(B) <= tol
        lams.append(m.load_factor)
    npts = 4 * len(mesh["ele_tags"])
    assert (m.point_data("plastic", npts)[:, 6] > 0).any()
    d = np.diff(lams)
    # softening of the secant: later increments of lambda are smaller
    assert d[-1] < 0.9 * d[0]
    m.close()


def test_displacement_control_elastic_cantilever_closed_form():
    """Elastic cantilever of one dispBeamColumn2d (cubic field, exact for a
    tip load) under DisplacementControl of the tip deflection: after k steps
    lambda P L^3 / (3 E I) = k incr, so lambda = 3 E I k incr / (P L^3)."""
    L, P, E = 3000.0, 5.0, 2.0e5
    y = np.array([-100.0, 100.0])
    A = np.array([200.0, 200.0])
    EI = E * np.sum(A * y * y)
    incr = 0.5
    m = _oracle.oracle_model(2, 3)
    m.node(1, 0.0, 0.0)
    m.node(2, L, 0.0)
    m.fix(1, 1, 1, 1)
    m.uniaxialMaterial("Elastic", 1, E)
    m.section("Fiber", 1, [(yy, 0.0, aa, 1) for yy, aa in zip(y, A)])
    m.geomTransf("Linear", 1)
    m.element("dispBeamColumn", 1, 1, 2, 1, "Legendre", 1, 5)
    m.load(2, 0.0, P, 0.0)
    m.system("dense")
    m.test("NormDispIncr", 1e-12, 10)
    m.integrator("DisplacementControl", 2, 2, incr)
    for k in range(1, 6):
        assert m.analyze(1) == 0
        want = 3.0 * EI * k * incr / (P * L ** 3)
        assert abs(m.load_factor - want) <= 1e-11 * want, (k, m.load_factor, want)
        assert abs(m.nodeDisp(2, 2) - k * incr) <= 1e-12 * k * incr
    m.close()


def test_displacement_control_zero_reference_load_is_refused():
    """DisplacementControl::domainChanged refuses a zero reference load
    (DisplacementControl.cpp:386-398)."""
    mesh = meshes.quad_strip(4, 1)
    m = _oracle.oracle_model(2, 2)
    meshes.build_quad_strip(m, mesh)
    m.system("dense")
    m.test("NormUnbalance", 1e-8, 10)
    m.integrator("DisplacementControl", mesh["tip"], 2, 0.001)
    with pytest.raises(Exception, match="zero reference load"):
        m.analyze(1)
    m.close()


def test_energy_incr_norm_by_hand():
    """NewtonRaphson + CTestEnergyIncr on a plastic load step, replayed with
    the state-determination primitives: per iteration X = K^-1 B, U += X,
    then B = unbalance after the update, and the test's norm is
    0.5 |X . B| of that pair (CTestEnergyIncr.cpp:101-108; the test runs
    after formUnbalance, NewtonRaphson.cpp:196-205).  The tolerance is put
    between the second and third hand-computed norms, so analyze must stop
    after exactly three iterations with the third norm."""
    mesh = meshes.quad_strip(8, 2)

    def build():
        m = _oracle.oracle_model(2, 2)
        meshes.build_quad_strip(m, mesh)
        m.load(mesh["tip"], 0.0, 1.0)
        m.system("dense")
        m.setup()
        return m

    lam = 180.0  # plastic at the root of the strip
    h = build()
    h.set_load_factor(lam)
    norms = []
    B = h.form_unbalance()
    for _ in range(4):
        h.form_tangent(values=False)
        X = h.solve(B)
        h.incr_disp(X)
        h.update()
        B = h.form_unbalance()
        norms.append(0.5 * abs(float(X @ B)))
    npts = 4 * len(mesh["ele_tags"])
    assert (h.point_data("plastic", npts)[:, 6] > 0).any()
    assert norms[2] < norms[1]
    h.close()
    tol = np.sqrt(norms[1] * norms[2])
    m = build()
    m.test("EnergyIncr", tol, 20)
    m.integrator("LoadControl", lam)
    c0 = m.counters()
    assert m.analyze(1) == 0
    c = m.counters()
    assert c["newton_iters"] - c0["newton_iters"] == 3
    assert abs(c["last_norm"] - norms[2]) <= 1e-9 * norms[2]
    m.close()
