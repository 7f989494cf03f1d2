"""NodeRecorder writes the reference's `recorder Node -time ... disp` lines:
space-separated values in C++ default stream format (6 significant digits)."""
import numpy as np

from tests import _oracle
from xara_amd import meshes
from xara_amd.recorder import NodeRecorder


def test_node_recorder_lines(tmp_path):
    mesh = meshes.quad_strip(8, 2)
    m = _oracle.oracle_model(2, 2)
    meshes.build_quad_strip(m, mesh)
    m.load(mesh["tip"], 0.0, 1.0)
    m.test("NormUnbalance", 1e-8, 25)
    m.integrator("DisplacementControl", mesh["tip"], 2, 0.001)
    m.setup()
    path = tmp_path / "tip.out"
    with NodeRecorder(str(path), [mesh["tip"], 5], [1, 2]) as rec:
        for _ in range(3):
            assert m.analyze(1) == 0
            rec.record(m, m.load_factor)
    lines = path.read_text().splitlines()
    assert len(lines) == 3
    U = m.trial_disp()
    i_tip = int(np.searchsorted(np.sort(mesh["node_tags"]), mesh["tip"]))
    last = lines[-1].split(" ")
    assert len(last) == 5
    assert last[0] == f"{m.load_factor:.6g}"
    assert last[2] == f"{U[i_tip, 1]:.6g}"
    assert abs(float(last[2]) - 0.003) < 1e-9  # the controlled displacement, 3 x 0.001
    assert f"{1e-05:.6g}" == "1e-05" and f"{100000.0:.6g}" == "100000" and f"{1234567.0:.6g}" == "1.23457e+06"
