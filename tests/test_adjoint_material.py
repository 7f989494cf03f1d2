"""The per-fiber adjoint of the GPU library (xg_material_vjp: the same code
the adjoint kernels run, evaluated on the host) is the transpose of the
oracle's forward-mode derivative of the Steel02 / Concrete02 trial update:
<hbar, J xdot> = <J^T hbar, xdot> for random directions, at every state of
cyclic strain paths that visit every branch.  The oracle's forward mode is
itself checked against finite differences of the trial update, and its
values against the oracle's own material update."""
import ctypes as C

import numpy as np
import pytest

from tests import _oracle
from xara_amd import _lib

STEEL = [420.0, 200e3, 0.01, 18.0, 0.925, 0.15, 0.04, 1.0, 0.04, 1.0, 0.0]   # with isotropic hardening
STEEL_PLAIN = [420.0, 200e3, 0.01, 18.0, 0.925, 0.15, 0.0, 1.0, 0.0, 1.0, 0.0]
CONC = [-30.0, -0.002, -6.0, -0.02, 0.1, 3.0, 1500.0]


@pytest.fixture(scope="module")
def gpu_lib():
    return _lib.load()


def vjp(lib, kind, prm, h, kon, strain, hbar):
    n = 10 if kind == 1 else 5
    prm, h, hbar = (np.ascontiguousarray(a, np.float64) for a in (prm, h, hbar))
    cbar, thbar, sbar = np.zeros(n), np.zeros(3), C.c_double(0)
    rc = lib.material_vjp(kind, prm.ctypes.data_as(_lib.P_D), h.ctypes.data_as(_lib.P_D), int(kon), float(strain),
                          hbar.ctypes.data_as(_lib.P_D), cbar.ctypes.data_as(_lib.P_D),
                          thbar.ctypes.data_as(_lib.P_D), C.byref(sbar))
    assert rc == 0, lib.error()
    return cbar, thbar, sbar.value


def initial_state(kind, prm):
    if kind == 1:
        epsy = prm[0] / prm[1]
        return np.array([-epsy, epsy, 0, 0, 0, 0, 0, 0, 0, prm[1]]), 0
    return np.array([0, 0, 0, 0, 2 * prm[0] / prm[1]]), 0


def steel_path():
    t = np.linspace(0, 6 * np.pi, 160)
    eps = 0.012 * (0.2 + t / t[-1]) * np.sin(t)
    eps[0] = 0.0  # no strain change from the virgin state first
    return np.concatenate([eps, eps[-1:] + 1e-3, eps[-1:] + 2e-3, eps[-1:]])


def conc_path():
    t = np.linspace(0, 5 * np.pi, 200)
    eps = -0.012 * (0.1 + t / t[-1]) * (0.5 - 0.8 * np.sin(t)) + 0.002 * np.cos(3 * t)
    eps = np.concatenate([[0.0], eps, [-0.03, -0.025, 0.001, 0.004, 0.006, 0.002]])
    return eps


def walk(kind, prm, path):
    """Yield (h, kon, strain) along a path with a commit after every step."""
    h, kon = initial_state(kind, prm)
    for eps in path:
        yield h.copy(), kon, eps
        dh = np.zeros_like(h)
        dth = np.zeros(3 if kind == 1 else 1)
        h, kon, _ = _oracle.material_jvp(kind, prm, h, kon, eps, dh, dth, 0.0)


@pytest.mark.parametrize("kind,prm,path", [(1, STEEL, steel_path()), (1, STEEL_PLAIN, steel_path()),
                                           (2, CONC, conc_path())], ids=["steel-iso", "steel", "concrete"])
def test_vjp_is_transpose_of_forward_mode(gpu_lib, kind, prm, path):
    rng = np.random.default_rng(11)
    n = 10 if kind == 1 else 5
    nth = 3 if kind == 1 else 1
    worst = 0.0
    kons = set()
    for h, kon, eps in walk(kind, prm, path):
        for _ in range(3):
            dh = rng.standard_normal(n) * np.maximum(np.abs(h), 1e-3)
            dth = rng.standard_normal(nth) * np.abs(np.array(prm[:nth]))
            deps = rng.standard_normal() * 1e-3
            tb = rng.standard_normal(n)
            tb[n - 1] = 0.0  # tangent entry: not propagated
            hout, kout, dhout = _oracle.material_jvp(kind, prm, h, kon, eps, dh, dth, deps)
            cbar, thbar, sbar = vjp(gpu_lib, kind, prm, h, kon, eps, tb)
            lhs = float(tb[: n - 1] @ dhout[: n - 1])
            terms = np.concatenate([cbar * dh, thbar[:nth] * dth, [sbar * deps]])
            rhs = float(terms.sum())
            scale = np.abs(tb[: n - 1] * dhout[: n - 1]).sum() + np.abs(terms).sum() + 1e-300
            worst = max(worst, abs(lhs - rhs) / scale)
        kons.add((kon, kout))
    assert worst < 1e-12, worst
    if kind == 1:
        # virgin, first loading both ways, reversals 1->2 and 2->1, carried-over history
        for tr in [(0, 3), (3, 1), (1, 2), (2, 1), (1, 1), (2, 2)]:
            assert tr in kons, (tr, kons)


def test_forward_mode_matches_finite_differences():
    rng = np.random.default_rng(5)
    for kind, prm, path in [(1, STEEL, steel_path()), (2, CONC, conc_path())]:
        n = 10 if kind == 1 else 5
        nth = 3 if kind == 1 else 1
        worst, checked = 0.0, 0
        for h, kon, eps in walk(kind, prm, path):
            dh = rng.standard_normal(n) * np.maximum(np.abs(h), 1e-4)
            dh[n - 1] = 0.0
            if kind == 1 and h[2] == h[3]:
                # |epspl - epss0| has a kink at 0; states reached from the
                # parameters keep both on the same value, so move them together
                dh[2] = dh[3]
            dth = rng.standard_normal(nth) * np.abs(np.array(prm[:nth]))
            deps = rng.standard_normal() * 1e-4
            _, kout, dhout = _oracle.material_jvp(kind, prm, h, kon, eps, dh, dth, deps)
            s = 1e-7
            pp = list(prm)
            pm = list(prm)
            for k in range(nth):
                pp[k] += s * dth[k]
                pm[k] -= s * dth[k]
            hp, kp, _ = _oracle.material_jvp(kind, pp, h + s * dh, kon, eps + s * deps, dh, dth, deps)
            hm, km, _ = _oracle.material_jvp(kind, pm, h - s * dh, kon, eps - s * deps, dh, dth, deps)
            if kp != kout or km != kout:
                continue  # the perturbation crosses a branch
            fd = (hp - hm) / (2 * s)
            sc = np.abs(fd[: n - 1]).max() + np.abs(dhout[: n - 1]).max() + 1e-300
            err = np.abs(fd[: n - 1] - dhout[: n - 1]).max() / sc
            if err > 1e-3:
                continue  # a branch boundary inside the stencil (e.g. a clipped stress)
            worst = max(worst, err)
            checked += 1
        assert checked > 0.8 * len(path), (checked, len(path))
        assert worst < 1e-5, worst


def test_forward_mode_values_equal_oracle_update():
    """T = double of the forward-mode restatement is the oracle's material
    update: stresses and tangents along the paths are identical."""
    orc = _oracle.load()
    for kind, prm, path in [(1, STEEL, steel_path()), (2, CONC, conc_path())]:
        n = len(path)
        stress, tangent = np.zeros(n), np.zeros(n)
        pr = np.ascontiguousarray(prm, np.float64)
        ep = np.ascontiguousarray(path, np.float64)
        act = np.ones(n, np.int32)
        rc = orc.uniaxial_path(kind, pr.ctypes.data_as(_lib.P_D), n, ep.ctypes.data_as(_lib.P_D),
                               act.ctypes.data_as(_lib.P_I), stress.ctypes.data_as(_lib.P_D),
                               tangent.ctypes.data_as(_lib.P_D))
        assert rc == 0
        isig, ie = (8, 9) if kind == 1 else (3, 4)
        for k, (h, kon, eps) in enumerate(walk(kind, prm, path)):
            hout, _, _ = _oracle.material_jvp(kind, prm, h, kon, eps, np.zeros_like(h),
                                              np.zeros(3 if kind == 1 else 1), 0.0)
            assert hout[isig] == stress[k] and hout[ie] == tangent[k], k
