"""Register budgets of the fiber kernels (CPU: compiles the device code for
gfx950 with the library's flags and reads the assembler's NumVgprs).

The section update runs latency-bound, so its occupancy matters: the 3D
kernel sits just under 168 VGPRs (three waves per SIMD; at 169 it drops to
two and runs about 20% slower), the 2D kernel under 128 (four waves), and
the ensemble's step kernels are compiled for four waves."""
import os
import re
import shutil
import subprocess

import pytest

from xara_amd import build as B

CSRC = os.path.join(os.path.dirname(B.__file__), "csrc")
pytestmark = pytest.mark.skipif(shutil.which(B.HIPCC) is None and not os.path.exists(B.HIPCC),
                                reason="hipcc not available")


def vgprs(tmp_path, source):
    out = str(tmp_path / (source + ".s"))
    flags = [f for f in B.FLAGS if f not in ("-fPIC",)]
    subprocess.run([B.HIPCC, "-x", "hip", *flags, "--cuda-device-only", "-S", "-o", out,
                    os.path.join(CSRC, source)], check=True, capture_output=True)
    text = open(out).read()
    res = {}
    for m in re.finditer(r"^(_Z\w+):", text, re.M):
        seg = text[m.end():]
        nv = re.search(r"; NumVgprs: (\d+)", seg)
        if nv:
            res[m.group(1)] = int(nv.group(1))
    return res


def pick(res, *parts):
    names = [k for k in res if all(p in k for p in parts)]
    assert len(names) == 1, (parts, names)
    return res[names[0]]


def test_section_update_register_budget(tmp_path):
    res = vgprs(tmp_path, "xg_frame.hip")
    assert pick(res, "k_frame_updateILb1E") <= 168
    assert pick(res, "k_frame_updateILb0E") <= 128


def test_ensemble_step_kernels_four_waves(tmp_path):
    res = vgprs(tmp_path, "xg_ensemble.hip")
    for k in ("k_ens_predictILi256E", "k_ens_iterateILi256E"):
        assert pick(res, k) <= 128
