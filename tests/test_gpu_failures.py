"""Failure paths of the GPU analysis, forced from the host (no kernel is made
to fault or hang): a convergence test that cannot pass returns the
reference's TestFailed code and leaves the last committed state, as
BasicAnalysisBuilder::analyzeStatic reverts it (as the oracle does); a
model without supports cannot be solved: the bounded PCG reports
BadLinearSolve, and the analysis fails (the oracle's direct solve meets a
round-off pivot instead of a zero one and fails in the convergence test)."""
import numpy as np
import pytest

from tests import _oracle
from xara_amd import meshes
from xara_amd.model import Model, XaraError

pytestmark = pytest.mark.gpu
XG_BAD_LINEAR_SOLVE, XG_TEST_FAILED = -5, -6


def analyze_code(m, n=1):
    try:
        return m.analyze(n)
    except XaraError as e:
        return e.code


def test_test_failure_reverts_to_last_commit():
    print("case: NormUnbalance that cannot converge", flush=True)
    mesh = meshes.quad_strip(16, 4)
    for m in (Model(2, 2), _oracle.oracle_model(2, 2)):
        meshes.build_quad_strip(m, mesh)
        m.load(mesh["tip"], 0.0, 1.0)
        m.system("GpuPCG", 1e-13, 20000)
        m.test("NormUnbalance", 1e-8, 25)
        m.integrator("DisplacementControl", mesh["tip"], 2, 0.01)
        m.setup()
        assert analyze_code(m, 3) == 0
        U0, lam0 = m.trial_disp().copy(), m.load_factor
        m.test("NormUnbalance", 1e-300, 1)  # one iteration, unreachable tolerance
        assert analyze_code(m) == XG_TEST_FAILED
        assert np.array_equal(m.trial_disp(), U0)
        assert m.load_factor == lam0


def test_unsupported_model_is_a_bad_linear_solve():
    print("case: model without supports", flush=True)
    mesh = meshes.brick_cube(3)
    mesh = dict(mesh, fixed_nodes=mesh["fixed_nodes"][:0])
    codes = []
    for m, system in ((Model(3, 3), "GpuBlockPCG"), (_oracle.oracle_model(3, 3), None)):
        meshes.build_brick_cube(m, mesh, pressure=1.0)
        if system:
            m.system(system, 1e-12, 300)  # bounded: a singular system cannot reach the tolerance
        m.test("NormUnbalance", 1e-8, 10)
        m.integrator("LoadControl", 1.0)
        m.setup()
        codes.append(analyze_code(m))
        assert np.all(m.trial_disp() == 0.0)  # reverted to the (zero) committed state
    assert codes[0] == XG_BAD_LINEAR_SOLVE
    assert codes[1] in (XG_BAD_LINEAR_SOLVE, XG_TEST_FAILED)
