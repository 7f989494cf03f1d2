"""The drop-in adapter inside the reference's own analysis loop (CPU, build
container only: it needs the reference tree).

oracle/_ref/adapter_check (built by `make -C oracle adapter` from
oracle/adapter_check.cpp, the adapter and the reference's unmodified
analysis sources) runs each pushover twice through the reference's
AnalysisModel / PlainHandler / RCM numberer / NewtonRaphson /
CTestNormUnbalance and its LoadControl or DisplacementControl, in the order
of BasicAnalysisBuilder::domainChanged / analyzeStatic
(BasicAnalysisBuilder.cpp:225-300,340-420):

* with the reference's own Brick / FourNodeQuad / DispBeamColumn2d and its
  ProfileSPD direct solver;
* with the adapter's GpuElement proxies, GpuLinSOE and GpuIntegrator<...>,
  their xg_* calls routed to the oracle's orc_* (the same C-ABI).

Every step's load factor and every nodal displacement must agree to 1e-9,
and the pushovers must be nonlinear (secant stiffness well below the first
step's), so the J2 / Steel02 / Concrete02 state really moves through the
adapter's element loops."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/SRC"
BIN = os.path.join(ROOT, "oracle", "_ref", "adapter_check")
RTOL = 1e-9

pytestmark = pytest.mark.skipif(not os.path.isdir(REF), reason="needs the reference tree")


@pytest.fixture(scope="module")
def adapter_check():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "adapter"], check=True,
                   capture_output=True)
    return BIN


def run(binary, kind, nsteps):
    r = subprocess.run([binary, kind, str(nsteps)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    runs = {"reference": [], "adapter": []}
    for line in r.stdout.splitlines():
        f = line.split()
        runs[f[0]].append((int(f[1]), float(f[2]), np.array(f[3:], dtype=float)))
    return runs["reference"], runs["adapter"]


# model, steps, and the largest last-step secant stiffness (load factor over
# the largest nodal displacement) allowed, as a fraction of the first step's
CASES = [("brick", 20, 0.4), ("quad", 20, 0.7), ("frame", 20, 0.7)]


@pytest.mark.parametrize("kind,nsteps,soften", CASES)
def test_adapter_matches_reference_elements_in_reference_loop(adapter_check, kind, nsteps, soften):
    ref, ada = run(adapter_check, kind, nsteps)
    assert len(ref) == len(ada) == nsteps
    for (sr, lr, ur), (sa, la, ua) in zip(ref, ada):
        assert sr == sa
        assert abs(la - lr) <= RTOL * abs(lr), (kind, sr, la, lr)
        assert np.max(np.abs(ua - ur)) <= RTOL * np.max(np.abs(ur)), (kind, sr)
    # nonlinear: the secant stiffness of the last step against the first
    sec = [lam / np.max(np.abs(u)) for _, lam, u in ref]
    assert sec[-1] < soften * sec[0], (kind, sec[0], sec[-1])
