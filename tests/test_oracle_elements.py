"""Element-level pins of the CPU oracle against the reference's own element
classes (Brick.cpp, FourNodeQuad.cpp, DispBeamColumn2d/3d.cpp over
LinearCrdTransf2d/3d.cpp, Node.cpp, Domain.cpp), compiled unmodified into
oracle/_ref by `make -C oracle ref`:

* live: seeded nodal displacement paths with commits through one element of
  each kind, the oracle's Element::getTangentStiff / getResistingForce
  against the reference's (build container only);
* golden: the same comparison against fixtures make_golden.py recorded from
  the reference (runs wherever the fixtures are);
* closed forms that need no reference: FourNodeQuad K = dR/dU by central
  differences on a plastic state and R = K u in the elastic range, and a
  3D cantilever whose section is stiffer about one axis, so the local y / z
  assignment from vecxz (LinearCrdTransf3d.cpp:203) decides which flexural
  rigidity each global deflection sees.
"""
import os

import numpy as np
import pytest

from tests import _oracle
from tests.golden.make_golden import (DAMP, ELEMENT_CASES, J2, J2_FAIL, element_fail_inputs, element_inputs,
                                      element_model, run_ref_damp, run_ref_element)

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz")


def run_oracle_element(case, U, act, j2=J2):
    """Walk the displacement path U (n x element DOFs) with commits act
    through the oracle's single-element model; returns K (n, nd, nd), R."""
    m, nen, ndf = element_model(_oracle.oracle_model, case, j2)
    m.setup()
    n = len(U)
    nd = nen * ndf
    K, R = np.zeros((n, nd, nd)), np.zeros((n, nd))
    for k in range(n):
        m.set_trial_disp(U[k].reshape(nen, ndf))
        m.update()
        K[k] = m.element_tangent(1)
        R[k] = m.element_resisting_force(1)
        if act[k]:
            m.commit()
    m.close()
    return K, R


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("case", sorted(ELEMENT_CASES))
@pytest.mark.parametrize("seed", [0, 1])
def test_element_live_reference(ref, case, seed):
    U, act = element_inputs(case, seed)
    Kr, Rr = run_ref_element(ref, case, U, act)
    Ko, Ro = run_oracle_element(case, U, act)
    # the oracle follows the reference's arithmetic order: same bits, or a
    # few ulps where the compiler contracts a product differently
    assert rel(Ko, Kr) <= 1e-13 and rel(Ro, Rr) <= 1e-13


@pytest.mark.parametrize("case", sorted(ELEMENT_CASES))
def test_element_golden(case):
    g = np.load(GOLD, allow_pickle=False)
    U, act = g[f"ele_{case}_U"], g[f"ele_{case}_act"]
    Ko, Ro = run_oracle_element(case, U, act)
    assert rel(Ko, g[f"ele_{case}_K"]) <= 1e-13 and rel(Ro, g[f"ele_{case}_R"]) <= 1e-13
    # the recorded path leaves the elastic range (tangent changes along it)
    assert rel(g[f"ele_{case}_K"][-1], g[f"ele_{case}_K"][0]) > 1e-3


def run_oracle_damp(case, U, act, alphaM=DAMP[0], betaK=DAMP[1], betaKc=DAMP[2]):
    """The oracle's Rayleigh damping matrix (alphaM M + betaK K + betaKcomm
    Kc) of the single element along the path: the factors set before
    set-up, so Kc starts from the tangent at set-up."""
    import ctypes as C
    m, nen, ndf = element_model(_oracle.oracle_model, case)
    m.rayleigh(alphaM, betaK, 0.0, betaKc)
    m.setup()
    orc = _oracle.load()
    nd = nen * ndf
    D = np.zeros((len(U), nd, nd))
    k = C.c_int(0)
    for i in range(len(U)):
        m.set_trial_disp(U[i].reshape(nen, ndf))
        m.update()
        Di = np.zeros(nd * nd)
        assert orc.element_damp(m._h, 1, Di.ctypes.data_as(C.POINTER(C.c_double)), C.byref(k)) == 0
        D[i] = Di.reshape(nd, nd)
        if act[i]:
            m.commit()
    m.close()
    return D


@pytest.mark.parametrize("case", sorted(ELEMENT_CASES))
def test_rayleigh_damping_matrix_golden(case):
    """Element::getDamp (Element.cpp:181-200) with betaK on the current
    tangent and betaKcomm on the tangent stored at the last commit
    (Element.cpp:74-80, 167-176), against the reference's along a path
    with commits; the committed part must lag the current one."""
    g = np.load(GOLD, allow_pickle=False)
    U, act = g[f"ele_{case}_U"], g[f"ele_{case}_act"]
    D = run_oracle_damp(case, U, act)
    assert rel(D, g[f"damp_{case}_D"]) <= 1e-13
    # a slip that used the current tangent for the committed part fails here
    Kt = g[f"ele_{case}_K"]
    assert rel(D, (DAMP[1] + DAMP[2]) * Kt) > 1e-6


@pytest.mark.parametrize("case", sorted(ELEMENT_CASES))
def test_rayleigh_damping_matrix_live_reference(ref, case):
    U, act = element_inputs(case, 4)
    assert rel(run_oracle_damp(case, U, act), run_ref_damp(ref, case, U, act)) <= 1e-13
    assert rel(run_oracle_damp(case, U, act, 0.0, 0.0, 5e-3), run_ref_damp(ref, case, U, act, 0.0, 0.0, 5e-3)) <= 1e-13


@pytest.mark.parametrize("case", ["brick", "quad"])
def test_element_return_map_failure_golden(case):
    """Brick and FourNodeQuad over J2 whose return map fails at most Gauss
    points of every third state (J2Plasticity.cpp:309-313): Brick::update
    ignores the material's return code (Brick.cpp:839) and J2PlaneStrain
    returns 0, so the elements go on with the failed points' previous trial
    stress and tangent; K and R against the reference's."""
    g = np.load(GOLD, allow_pickle=False)
    U, act = g[f"elefail_{case}_U"], g[f"elefail_{case}_act"]
    Ko, Ro = run_oracle_element(case, U, act, J2_FAIL)
    assert rel(Ko, g[f"elefail_{case}_K"]) <= 1e-13 and rel(Ro, g[f"elefail_{case}_R"]) <= 1e-13


@pytest.mark.parametrize("case", ["brick", "quad"])
def test_element_return_map_failure_live_reference(ref, case):
    U, act = element_fail_inputs(case, 3)
    Kr, Rr = run_ref_element(ref, case, U, act, J2_FAIL)
    Ko, Ro = run_oracle_element(case, U, act, J2_FAIL)
    assert rel(Ko, Kr) <= 1e-13 and rel(Ro, Rr) <= 1e-13


# ---- closed forms that pin FourNodeQuad without the reference ----

def _quad(orc_model, crd, thk=1.3):
    m = orc_model(2, 2)
    for a in range(4):
        m.node(a + 1, *crd[a])
    m.nDMaterial("J2Plasticity", 1, *J2[:6])
    m.element("quad", 1, 1, 2, 3, 4, thk, "PlaneStrain", 1)
    m.setup()
    return m


QUAD_CRD = np.array([[0.0, 0.0], [1.3, 0.1], [1.1, 0.9], [-0.2, 1.2]])


def test_quad_tangent_is_derivative_of_resisting_force():
    """K = dR/dU by central differences on a plastic state, from committed
    plastic history (FourNodeQuad.cpp:226-281 against :507-553): a wrong
    thickness, dvol or B-row in either side breaks the identity."""
    rng = np.random.default_rng(5)
    m = _quad(_oracle.oracle_model, QUAD_CRD)
    U0 = 4e-3 * rng.normal(size=(4, 2))
    m.set_trial_disp(0.5 * U0)
    m.update()
    m.commit()
    m.set_trial_disp(U0)
    m.update()
    K = m.element_tangent(1)
    plastic = m.point_data("plastic", 4)[:, 6]
    assert np.all(plastic > 0), "state must be plastic at every Gauss point"
    h = 1e-7
    Kfd = np.zeros((8, 8))
    for j in range(8):
        cols = []
        for s in (1.0, -1.0):
            Up = U0.copy().ravel()
            Up[j] += s * h
            m.set_trial_disp(Up.reshape(4, 2))
            m.update()
            cols.append(m.element_resisting_force(1))
        Kfd[:, j] = (cols[0] - cols[1]) / (2 * h)
    assert rel(K, Kfd) < 1e-6
    assert rel(K, K.T) < 1e-12  # the consistent J2 tangent is symmetric
    m.close()


def test_quad_elastic_force_equals_tangent_times_displacement():
    """In the elastic range R = K u exactly (linear), with K the textbook
    plane-strain stiffness thk * sum_gp B^T D B detJ, D from K and G."""
    rng = np.random.default_rng(6)
    thk = 1.3
    m = _quad(_oracle.oracle_model, QUAD_CRD, thk)
    u = 2e-5 * rng.normal(size=(4, 2))
    m.set_trial_disp(u)
    m.update()
    K = m.element_tangent(1)
    R = m.element_resisting_force(1)
    assert rel(R, K @ u.ravel()) < 1e-12
    # independent textbook stiffness (2x2 Gauss, bilinear shape functions)
    Kb, G = J2[0], J2[1]
    lam = Kb - 2.0 * G / 3.0
    D = np.array([[lam + 2 * G, lam, 0], [lam, lam + 2 * G, 0], [0, 0, G]])
    g = 1 / np.sqrt(3.0)
    Kt = np.zeros((8, 8))
    for xi, eta in ((-g, -g), (-g, g), (g, -g), (g, g)):
        dN = 0.25 * np.array([[-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)],
                              [-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)]])
        J = dN @ QUAD_CRD
        dNx = np.linalg.solve(J, dN)
        B = np.zeros((3, 8))
        B[0, 0::2] = dNx[0]
        B[1, 1::2] = dNx[1]
        B[2, 0::2] = dNx[1]
        B[2, 1::2] = dNx[0]
        Kt += thk * np.linalg.det(J) * B.T @ D @ B
    assert rel(K, Kt) < 1e-12
    m.close()


# ---- LinearCrdTransf3d local axes (vecxz) pinned by a closed form ----

@pytest.mark.parametrize("vecxz,weak_global", [((1.0, 0.0, 0.0), 0), ((0.0, 1.0, 0.0), 1)])
def test_crdtransf3d_local_axes_select_rigidity(vecxz, weak_global):
    """Vertical elastic cantilever (local x = global Z) whose fibers make the
    rigidity about local z four times that about local y.
    LinearCrdTransf3d::getLocalAxes (LinearCrdTransf3d.cpp:274-320) takes
    local y = vecxz x x_local (normalised) and local z = x_local x y:
    vecxz = X gives y = -Y, z = X; vecxz = Y gives y = X, z = Y.  A tip load
    along local z bends about local y (EIy = sum E A z^2, the weak axis
    here), one along local y about local z (EIz = sum E A y^2): tip
    deflection P L^3 / (3 EI), exact for the cubic displacement field.
    Swapping the local y / z assignment swaps the two deflections."""
    L, P, E = 2000.0, 10.0, 2.0e5
    y = np.array([100.0, -100.0, 0.0, 0.0])
    z = np.array([0.0, 0.0, 50.0, -50.0])
    A = np.array([10.0, 10.0, 10.0, 10.0])
    EIz, EIy = E * np.sum(A * y * y), E * np.sum(A * z * z)
    assert EIz == 4 * EIy
    tip = []
    for d in (0, 1):
        m = _oracle.oracle_model(3, 6)
        m.node(1, 0.0, 0.0, 0.0)
        m.node(2, 0.0, 0.0, L)
        m.fix(1, 1, 1, 1, 1, 1, 1)
        m.uniaxialMaterial("Elastic", 1, E)
        m.section("Fiber", 1, [(yy, zz, aa, 1) for yy, zz, aa in zip(y, z, A)], GJ=1e12)
        m.geomTransf("Linear", 1, *vecxz)
        m.element("dispBeamColumn", 1, 1, 2, 1, "Legendre", 1, 5)
        P6 = [0.0] * 6
        P6[d] = P
        m.load(2, *P6)
        m.system("dense")
        m.test("NormDispIncr", 1e-12, 10)
        m.integrator("LoadControl", 1.0)
        assert m.analyze(1) == 0
        tip.append(m.nodeDisp(2, d + 1))
        m.close()
    for d in (0, 1):
        want = P * L ** 3 / (3 * (EIy if d == weak_global else EIz))
        assert abs(tip[d] - want) <= 1e-9 * want, (d, tip, want)
