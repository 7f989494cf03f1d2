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


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b))
    return 0.0 if s == 0 else float(np.max(np.abs(a - b)) / s)


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


def _fail_model(m, kind, mesh):
    """Alternate elements of J2 and of the failing J2 (J2_FAIL: the return
    map exceeds its 25 Newton iterations at large trial stress norms)."""
    from tests.golden.make_golden import J2, J2_FAIL
    m.nodes(mesh["node_tags"], mesh["crds"])
    ndf = 3 if kind == "brick" else 2
    fx = mesh["fixed_nodes"]
    m.fix_many(fx, np.ones((len(fx), ndf), np.int32))
    m.nDMaterial("J2Plasticity", 1, *J2[:6])
    m.nDMaterial("J2Plasticity", 2, *J2_FAIL[:6])
    even = mesh["ele_tags"] % 2 == 0
    for mat, sel in ((1, ~even), (2, even)):
        if kind == "brick":
            m.elements("stdBrick", mesh["ele_tags"][sel], mesh["conn"][sel], mat)
        else:
            m.elements("quad", mesh["ele_tags"][sel], mesh["conn"][sel], 1.0, mat)
    m.system("GpuPCG", 1e-13, 5000)
    m.setup()


@pytest.mark.parametrize("kind", ["brick", "quad"])
def test_return_map_failure_leaves_trial_state(kind):
    """J2Plasticity.cpp:309-313 on the device: points whose local Newton
    iteration fails keep their previous trial stress, plastic strain and
    tangent (the update reports no failure: Brick.cpp:839 ignores the code,
    J2PlaneStrain returns 0); the other points update.  Points, element
    forces and the assembled tangent equal the oracle's to 1e-9, and the
    same points fail on both sides."""
    print(f"case: J2 return-map failure, {kind}", flush=True)
    mesh = meshes.brick_cube(3) if kind == "brick" else meshes.quad_strip(8, 3, 2.0, 0.75)
    nen, ndf = (8, 3) if kind == "brick" else (4, 2)
    npts = nen * len(mesh["ele_tags"])
    rng = np.random.default_rng(11)
    nn = len(mesh["node_tags"])
    free = np.ones((nn, 1))
    free[np.isin(mesh["node_tags"], mesh["fixed_nodes"])] = 0.0
    # small elastic, moderate plastic, large (most J2_FAIL points fail)
    fields = [free * rng.normal(size=(nn, ndf)) * s for s in (3e-5, 8e-4, 6e-3)]
    res = []
    for m in (Model(ndf, ndf), _oracle.oracle_model(ndf, ndf)):
        _fail_model(m, kind, mesh)
        steps = []
        for k, U in enumerate(fields):
            m.set_trial_disp(U)
            assert m.update() == 0
            steps.append((m.point_data("stress", npts), m.point_data("tangent", npts),
                          m.point_data("plastic", npts), m.form_unbalance(), m.form_tangent()))
            if k == 0:
                m.commit()
        res.append(steps)
    for sg, so in zip(*res):
        for a, b in zip(sg, so):
            assert relerr(a, b) <= 1e-9
    # failed points: stress unchanged from the previous trial state, same set on both sides
    failed = [np.all(st[2][0] == st[1][0], axis=1) for st in res]
    assert np.array_equal(failed[0], failed[1])
    on_fail = np.repeat(mesh["ele_tags"] % 2 == 0, nen)
    assert failed[0].sum() >= npts // 4 and not failed[0][~on_fail].any()
    # moved points did move
    assert not np.all(res[0][2][0] == res[0][1][0])
