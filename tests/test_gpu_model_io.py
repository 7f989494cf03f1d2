"""Models read from ``print -json`` documents (the reference's Print field
names, xara_amd/model_io.py) run on the device: the same document loaded
into the GPU library and into the oracle gives the same pushover to 1e-9,
and the numbering and pattern equal those of the model built by commands."""
import json

import numpy as np
import pytest

from tests import _oracle
from tests.test_model_io import conc_json, fiber_section_json, node_json, sp_json, steel_json
from xara_amd import meshes, model_io
from xara_amd.model import Model

pytestmark = pytest.mark.gpu
RTOL = 1e-9


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b))
    return 0.0 if s == 0 else float(np.max(np.abs(a - b)) / s)


def frame_doc(fr):
    conn, transf = meshes.frame_elements(fr)
    theta = meshes.default_theta(len(conn))
    theta[:, 0] *= np.linspace(0.9, 1.1, len(conn))
    uni, secs = [], []
    for e in range(len(conn)):
        tag = e + 1
        fy, E, b, fc = theta[e]
        uni.append(steel_json(meshes.STEEL_TAG0 + tag, fy, E, b, *meshes.STEEL02[3:]))
        uni.append(conc_json(meshes.CONC_TAG0 + tag, fc, *meshes.CONCRETE02[1:]))
        rows = meshes.rc_section_fibers(conc_tag=meshes.CONC_TAG0 + tag, steel_tag=meshes.STEEL_TAG0 + tag)
        secs.append(fiber_section_json(tag, rows, False))
    eles = [{"name": e + 1, "type": "DispBeamColumn2d", "nodes": [int(conn[e, 0]), int(conn[e, 1])],
             "sections": [str(e + 1)] * 5, "integration": {"type": "Legendre"}, "massperlength": 0,
             "crdTransformation": str(int(transf[e]))} for e in range(len(conn))]
    props = {"sections": secs, "nDMaterials": [], "uniaxialMaterials": uni,
             "crdTransformations": [{"name": "1", "type": "LinearCrdTransf2d"}]}
    geom = {"nodes": node_json(fr["node_tags"], fr["crds"], 3), "elements": eles,
            "constraints": sp_json(fr["fixed_nodes"], 3)}
    return json.dumps({"StructuralAnalysisModel": {"properties": props, "geometry": geom}})


def test_frame_pushover_from_json_document():
    print("case: 2D RC frame from a print -json document", flush=True)
    fr = meshes.rc_frame_2d(3, 2)
    doc = frame_doc(fr)
    res = []
    for m in (Model(2, 3), _oracle.oracle_model(2, 3)):
        model_io.load_json(m, doc)
        P = np.zeros((len(fr["floor_nodes"]), 3))
        P[:, 0] = fr["floor_level"] / fr["floor_level"].max()
        m.loads(fr["floor_nodes"], P)
        m.system("GpuBlockPCG", 1e-14, 100000)
        m.test("NormUnbalance", 1e-5, 40)
        m.integrator("DisplacementControl", fr["roof"], 1, 8.0)
        m.setup()
        traj = [(m.node_eqns(),) + m.csr_pattern()]
        for step in range(12):
            assert m.analyze(1) == 0, step
            traj.append((m.trial_disp().copy(), m.load_factor))
        res.append(traj)
    (g0, *gt), (o0, *ot) = res
    for a, b in zip(g0, o0):
        assert np.array_equal(a, b)
    for (ug, lg), (uo, lo) in zip(gt, ot):
        assert relerr(ug, uo) <= RTOL
        assert abs(lg - lo) <= RTOL * abs(lo)
    # the same numbering as the model built by commands
    direct = _oracle.oracle_model(2, 3)
    meshes.build_rc_frame_per_element(direct, fr, meshes.default_theta(len(meshes.frame_elements(fr)[0])))
    direct.setup()
    assert np.array_equal(direct.node_eqns(), g0[0])
