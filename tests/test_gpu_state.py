"""Committed-state hand-over between the GPU model and the CPU oracle in the
reference's sendSelf layouts: a transient run on one, its committed state
loaded into the other, both continue on the same trajectory (1e-9)."""
import numpy as np
import pytest

from tests import _oracle
from xara_amd import ground_motion, meshes
from xara_amd.model import Model

pytestmark = pytest.mark.gpu
DT = 0.01


def frame(m, ag):
    fr = meshes.rc_frame_2d(3, 2)
    meshes.define_transient_frame(m, fr, meshes.default_theta(len(meshes.frame_elements(fr)[0])), ag, DT,
                                  tol=1e-12, maxit=40)
    if isinstance(m, Model) and m._b is not _oracle.load():
        m.system("GpuBlockPCG", 1e-14, 100000)
    m.setup()
    return m


@pytest.mark.parametrize("direction", ["oracle_to_gpu", "gpu_to_oracle"])
def test_transient_state_handover(direction):
    print(f"case: state hand-over {direction}", flush=True)
    ag = ground_motion.kanai_tajimi(3, 300, DT, 0.9)
    g, o = frame(Model(2, 3), ag), frame(_oracle.oracle_model(2, 3), ag)
    src, dst = (o, g) if direction == "oracle_to_gpu" else (g, o)
    for _ in range(80):
        assert src.analyze(1, DT) == 0
    st = src.save_state()
    dst.load_state(st)
    for k in ("nodes", "Steel02", "Concrete02", "domain"):
        assert np.array_equal(dst.get_state(k), st[k]), k
    for step in range(60):
        assert src.analyze(1, DT) == 0 and dst.analyze(1, DT) == 0
        a, b = src.trial_disp(), dst.trial_disp()
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(a)), step
