"""Speed of the CPU oracle against the reference's own code on the same CPU
and inputs (build container only: needs oracle/_ref/libxararef.so, built by
`make -C oracle ref` from the reference sources; both libraries -O2, one
thread).  Times the per-point J2ThreeDimensional update (setTrialStrain +
getStress + getTangent, J2ThreeDimensional.cpp:122-234) and the Steel02 and
Concrete02 fiber updates (UniaxialMaterial::setTrial) along long seeded
strain paths, so the oracle timed as the CPU baseline is shown to be as fast
as, or faster than, the reference it restates -- not a straw man.

Usage: python oracle/calibrate.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from tests import _oracle  # noqa: E402
from tests.golden.make_golden import CONC, J2, STEEL, dp, ip  # noqa: E402


def best_of(fn, reps=5):
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return min(t)


def main():
    orc, ref = _oracle.load(), _oracle.load_ref()
    if ref is None:
        raise SystemExit("build oracle/_ref first: make -C oracle ref")
    rng = np.random.default_rng(0)
    rows = []
    # J2, 3D: a plastic random walk (most steps on the yield surface)
    n = 200_000
    eps = np.ascontiguousarray(np.cumsum(rng.normal(0, 1.5e-3, (n, 6)), axis=0) * 0.05)
    act = np.ascontiguousarray((rng.random(n) < 0.5).astype(np.int32))
    s, t = np.zeros((n, 6)), np.zeros((n, 36))
    tr = best_of(lambda: ref.j2_3d_path(dp(J2), n, dp(eps), ip(act), dp(s), dp(t)))
    to = best_of(lambda: orc.j2_3d_path(dp(J2), n, dp(eps), ip(act), dp(s), dp(t)))
    rows.append(("J2ThreeDimensional point update", n, tr, to))
    # Steel02 / Concrete02 fibers
    m = 1_000_000
    x = np.linspace(0, 60 * np.pi, m)
    for name, kind, prm, amp, shift in (("Steel02 fiber update", 1, STEEL, 0.012, 0.0),
                                        ("Concrete02 fiber update", 2, CONC, 0.003, -0.002)):
        e = np.ascontiguousarray(amp * np.sin(x) * (1 + 0.3 * np.sin(0.01 * x)) + shift)
        a = np.ascontiguousarray((rng.random(m) < 0.5).astype(np.int32))
        sv, tv = np.zeros(m), np.zeros(m)
        tr = best_of(lambda: ref.uniaxial_path(kind, dp(prm), m, dp(e), ip(a), dp(sv), dp(tv)))
        to = best_of(lambda: orc.uniaxial_path(kind, dp(prm), m, dp(e), ip(a), dp(sv), dp(tv)))
        rows.append((name, m, tr, to))
    print(f"{'update':34s} {'count':>9s} {'reference /s':>14s} {'oracle /s':>14s} {'oracle/reference':>17s}")
    for name, cnt, tr, to in rows:
        print(f"{name:34s} {cnt:9d} {cnt / tr:14.4g} {cnt / to:14.4g} {tr / to:17.2f}")


if __name__ == "__main__":
    main()
