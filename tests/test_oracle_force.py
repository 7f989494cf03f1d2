"""CPU checks of the oracle's forceBeamColumn2d/3d (the checker for the GPU
force-based element).  The reference element is not in the reference
build (it needs the Node/Domain machinery), so it is pinned by what the
mathematics fixes: the exact flexibility of an elastic prismatic member,
equilibrium and compatibility at every converged state (the two conditions
the element iteration enforces), the consistent tangent against central
differences of the resisting force, and revert/commit bookkeeping."""
import numpy as np
import pytest

from tests import _oracle
from xara_amd import meshes

E, A, IZ, IY, GJ = 200e3, 1.2e4, 3.6e7, 2.5e7, 1.1e14
L = 3000.0


def elastic_fibers(three_d):
    """Four elastic fibers with the section's A, Iz (and Iy): +/- a at +/- b."""
    a = A / 4
    yb = np.sqrt(IZ / A)
    zb = np.sqrt(IY / A) if three_d else 0.0
    return [(yb, zb, a, 1), (-yb, zb, a, 1), (yb, -zb, a, 1), (-yb, -zb, a, 1)]


def cantilever(three_d, nip, section_rows, materials, gj=GJ):
    m = _oracle.oracle_model(3 if three_d else 2, 6 if three_d else 3)
    if three_d:
        m.nodes([1, 2], [[0.0, 0.0, 0.0], [L, 0.0, 0.0]])
        m.fix_many([1], [[1] * 6])
        m.geomTransf("Linear", 1, 0.0, 0.0, 1.0)
    else:
        m.nodes([1, 2], [[0.0, 0.0], [L, 0.0]])
        m.fix_many([1], [[1] * 3])
        m.geomTransf("Linear", 1)
    materials(m)
    m.section("Fiber", 1, section_rows, GJ=gj if three_d else None)
    m.elements("forceBeamColumn", [1], [(1, 2)], 1, 1, nip)
    m.setup()
    return m


def elastic(m):
    m.uniaxialMaterial("Elastic", 1, E)


@pytest.mark.parametrize("three_d", [False, True])
@pytest.mark.parametrize("nip", [3, 5, 7])
def test_elastic_cantilever_closed_form(three_d, nip):
    """Lobatto with n >= 3 points integrates the cubic-free flexibility of a
    prismatic member exactly: tip deflection P L^3 / (3 E I), rotation
    P L^2 / (2 E I), axial N L / (E A), twist T L / GJ.  The tabulated
    weights carry 10 digits, so agreement is to ~1e-9."""
    m = cantilever(three_d, nip, elastic_fibers(three_d), elastic)
    ndf = 6 if three_d else 3
    P = np.zeros(ndf)
    P[0], P[1] = 50e3, 20e3
    if three_d:
        P[2], P[3] = -15e3, 4e8
    m.loads([2], [P])
    m.test("NormDispIncr", 1e-12, 10)
    m.integrator("LoadControl", 1.0)
    assert m.analyze(1) == 0
    u = m.trial_disp()[1]
    assert u[0] == pytest.approx(P[0] * L / (E * A), rel=1e-9)
    assert u[1] == pytest.approx(P[1] * L ** 3 / (3 * E * IZ), rel=1e-9)
    if three_d:
        assert u[2] == pytest.approx(P[2] * L ** 3 / (3 * E * IY), rel=1e-9)
        assert u[3] == pytest.approx(P[3] * L / GJ, rel=1e-9)
        assert u[5] == pytest.approx(P[1] * L ** 2 / (2 * E * IZ), rel=1e-9)
    else:
        assert u[2] == pytest.approx(P[1] * L ** 2 / (2 * E * IZ), rel=1e-9)


def rc_cantilever(three_d, nip=5):
    def mats(m):
        m.uniaxialMaterial("Steel02", 2, *meshes.STEEL02)
        m.uniaxialMaterial("Concrete02", 3, *meshes.CONCRETE02)
    return cantilever(three_d, nip, meshes.rc_section_fibers(conc_tag=3, steel_tag=2), mats)


def lobatto(nip):
    from numpy.polynomial.legendre import Legendre
    x = np.concatenate(([-1.0], np.sort(Legendre.basis(nip - 1).deriv().roots()), [1.0]))
    return 0.5 * (x + 1.0)


@pytest.mark.parametrize("three_d", [False, True])
def test_equilibrium_and_compatibility_when_converged(three_d):
    """After a nonlinear update the sections carry b(x) q (equilibrium) and
    the basic deformations equal the weighted sum of b(x)^T e(x)
    (compatibility), to the element tolerance; the steel has yielded."""
    nip = 5
    m = rc_cantilever(three_d, nip)
    ndf = 6 if three_d else 3
    U = np.zeros((2, ndf))
    U[1, 0], U[1, 1] = -0.6, 45.0  # axial shortening, tip drift 1.5%
    if three_d:
        U[1, 2], U[1, 3] = 12.0, 1e-4
    m.set_trial_disp(U)
    assert m.update() == 0
    s = m.point_data("stress", nip)
    e = m.point_data("strain", nip)
    P = m.element_resisting_force(1)
    xi = lobatto(nip)
    # basic forces from the end forces of a member along x
    if three_d:
        q = np.array([P[6], P[5], P[11], P[4], P[10], P[9]])
        bs = np.stack([np.full(nip, q[0]), (xi - 1) * q[1] + xi * q[2], (xi - 1) * q[3] + xi * q[4],
                       np.full(nip, q[5])], axis=1)
    else:
        q = np.array([P[3], P[2], P[5]])
        bs = np.stack([np.full(nip, q[0]), (xi - 1) * q[1] + xi * q[2]], axis=1)
    scale = np.abs(s).max(axis=0)
    assert np.all(np.abs(s - bs) <= 1e-6 * scale)
    # compatibility with the Lobatto weights of the reference table
    w = 0.5 * np.array([0.1, 0.5444444444, 0.7111111111, 0.5444444444, 0.1]) * L
    v = np.zeros(6 if three_d else 3)
    v[0] = np.sum(w * e[:, 0])
    v[1], v[2] = np.sum(w * (xi - 1) * e[:, 1]), np.sum(w * xi * e[:, 1])
    if three_d:
        v[3], v[4] = np.sum(w * (xi - 1) * e[:, 2]), np.sum(w * xi * e[:, 2])
        v[5] = np.sum(w * e[:, 3])
        # LinearCrdTransf3d::getBasicTrialDisp with node i fixed and R = I
        vb = np.array([U[1, 0], -U[1, 1] / L, U[1, 5] - U[1, 1] / L,
                       U[1, 2] / L, U[1, 4] + U[1, 2] / L, U[1, 3]])
    else:
        vb = np.array([U[1, 0], -U[1, 1] / L, U[1, 2] - U[1, 1] / L])
    assert np.allclose(v, vb, rtol=1e-7, atol=1e-12 * np.abs(vb).max())
    # the base section has softened (yielded steel), the tip section not
    ord_ = 4 if three_d else 2
    kz = m.point_data("tangent", nip)[:, ord_ + 1]
    assert kz[0] < 0.8 * kz[-1]


@pytest.mark.parametrize("three_d", [False, True])
def test_tangent_is_the_derivative_of_the_resisting_force(three_d):
    """The consistent element tangent (kv = f^{-1} transformed) equals the
    central difference of the resisting force, each perturbed update
    starting from the same committed state."""
    m = rc_cantilever(three_d)
    ndf = 6 if three_d else 3
    U0 = np.zeros((2, ndf))
    U0[1, 0], U0[1, 1] = -0.3, 25.0
    if three_d:
        U0[1, 2] = 8.0
    m.set_trial_disp(U0)
    assert m.update() == 0
    m.commit()
    U1 = U0.copy()
    U1[1, 1] += 4.0
    m.set_trial_disp(U1)
    assert m.update() == 0
    K = m.element_tangent(1)
    Kfd = np.zeros_like(K)
    for k in range(ndf):
        h = 1e-4 if k < (3 if three_d else 2) else 1e-7
        R = []
        for sgn in (1, -1):
            m.revert()
            U = U1.copy()
            U[1, k] += sgn * h
            m.set_trial_disp(U)
            assert m.update() == 0
            R.append(m.element_resisting_force(1))
        Kfd[:, ndf + k] = (R[0] - R[1]) / (2 * h)
    cols = slice(ndf, 2 * ndf)
    scale = np.abs(K[:, cols]).max()
    assert np.abs(K[:, cols] - Kfd[:, cols]).max() <= 1e-4 * scale


def test_revert_restores_the_committed_element():
    """revertToLastCommit restores q, kv and the section deformations; the
    domain's closing update reproduces the committed resisting force."""
    m = rc_cantilever(False)
    U = np.zeros((2, 3))
    U[1, 1] = 30.0
    m.set_trial_disp(U)
    m.update()
    m.commit()
    R0 = m.element_resisting_force(1).copy()
    K0 = m.element_tangent(1).copy()
    U[1, 1] = 60.0
    m.set_trial_disp(U)
    m.update()
    assert not np.allclose(m.element_resisting_force(1), R0)
    m.revert()
    assert np.allclose(m.element_resisting_force(1), R0, rtol=1e-9, atol=1e-6)
    assert np.allclose(m.element_tangent(1), K0, rtol=1e-9, atol=1e-6)


def test_pushover_close_to_displacement_based_refinement():
    """One force-based element against six displacement-based ones on the
    same RC cantilever pushover: the lateral load at 40 mm of tip drift
    agrees within 10% (the refined displacement-based model being
    slightly stiffer)."""
    lam = []
    for kind, nel in (("forceBeamColumn", 1), ("dispBeamColumn", 6)):
        m = _oracle.oracle_model(2, 3)
        xs = np.linspace(0.0, L, nel + 1)
        m.nodes(np.arange(1, nel + 2), np.stack([xs, np.zeros(nel + 1)], axis=1))
        m.fix_many([1], [[1, 1, 1]])
        m.geomTransf("Linear", 1)
        m.uniaxialMaterial("Steel02", 2, *meshes.STEEL02)
        m.uniaxialMaterial("Concrete02", 3, *meshes.CONCRETE02)
        m.section("Fiber", 1, meshes.rc_section_fibers(conc_tag=3, steel_tag=2))
        m.elements(kind, np.arange(1, nel + 1), np.stack([np.arange(1, nel + 1), np.arange(2, nel + 2)], 1),
                   1, 1, 5)
        m.load(nel + 1, 0.0, 1e3, 0.0)
        m.test("NormDispIncr", 1e-8, 40)
        m.integrator("DisplacementControl", nel + 1, 2, 1.0)
        m.setup()
        assert m.analyze(40) == 0
        lam.append(m.load_factor)
    assert lam[0] == pytest.approx(lam[1], rel=0.10)
    assert lam[0] < lam[1]
