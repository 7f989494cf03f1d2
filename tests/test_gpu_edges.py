"""Edge cases of the GPU path against the CPU oracle through the C-ABI: a
model whose nodes are all fixed (no equations), the smallest systems (one
brick with 12 equations, one row of quads, one frame element), elements of
one group with different materials, and an unknown material tag."""
import numpy as np
import pytest

from tests import _oracle
from xara_amd import meshes
from xara_amd.model import Model, XaraError

pytestmark = pytest.mark.gpu
RTOL = 1e-9


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    s = np.max(np.abs(b))
    return 0.0 if s == 0 else float(np.max(np.abs(a - b)) / s)


def j2(m, tag, sig0_scale=1.0):
    mat = meshes.J2_STEEL
    m.nDMaterial("J2Plasticity", tag, mat["K"], mat["G"], sig0_scale * mat["sig0"], sig0_scale * mat["sigInf"],
                 mat["delta"], mat["H"])


def test_all_nodes_fixed():
    """No equations: setup, a load step and the state queries still work and
    agree (nothing moves, stresses stay zero)."""
    print("case: all nodes fixed", flush=True)
    mesh = meshes.brick_cube(1)
    out = []
    for m in (Model(3, 3), _oracle.oracle_model(3, 3)):
        m.nodes(mesh["node_tags"], mesh["crds"])
        m.fix_many(mesh["node_tags"], np.ones((len(mesh["node_tags"]), 3), np.int32))
        j2(m, 1)
        m.elements("stdBrick", mesh["ele_tags"], mesh["conn"], 1)
        m.system("GpuPCG", 1e-12, 1000)
        m.test("NormUnbalance", 1e-8, 10)
        m.integrator("LoadControl", 1.0)
        m.setup()
        assert m.neq == 0
        assert m.analyze(2) == 0
        out.append((m.trial_disp().copy(), m.point_data("stress", 8)))
    (ug, sg), (uo, so) = out
    assert np.all(ug == 0) and np.all(uo == 0)
    assert np.all(sg == 0) and np.all(so == 0)


def test_single_brick_twelve_equations():
    """One stdBrick, base fixed: 12 equations, into the plastic range."""
    print("case: single brick", flush=True)
    mesh = meshes.brick_cube(1)
    traj = []
    for m in (Model(3, 3), _oracle.oracle_model(3, 3)):
        meshes.build_brick_cube(m, mesh, pressure=1.0)
        m.system("GpuBlockPCG", 1e-13, 1000)
        m.test("NormUnbalance", 1e-9, 20)
        m.integrator("LoadControl", 60.0)
        m.setup()
        assert m.neq == 12
        t = []
        for step in range(4):
            assert m.analyze(1) == 0
            t.append((m.trial_disp().copy(), m.point_data("stress", 8)))
        traj.append(t)
    for (ug, sg), (uo, so) in zip(*traj):
        assert relerr(ug, uo) <= RTOL
        assert relerr(sg, so) <= RTOL


def test_one_row_quad_strip():
    """A 16 x 1 strip of FourNodeQuads under DisplacementControl."""
    print("case: quad strip 16x1", flush=True)
    mesh = meshes.quad_strip(16, 1)
    res = []
    for m in (Model(2, 2), _oracle.oracle_model(2, 2)):
        meshes.build_quad_strip(m, mesh)
        m.load(mesh["tip"], 0.0, 1.0)
        m.system("GpuPCG", 1e-13, 20000)
        m.test("NormUnbalance", 1e-8, 25)
        m.integrator("DisplacementControl", mesh["tip"], 2, 0.002)
        m.setup()
        t = []
        for step in range(6):
            assert m.analyze(1) == 0
            t.append((m.trial_disp().copy(), m.load_factor))
        res.append(t)
    for (ug, lg), (uo, lo) in zip(*res):
        assert relerr(ug, uo) <= RTOL
        assert abs(lg - lo) <= RTOL * abs(lo)


def test_single_frame_element_cantilever():
    """One dispBeamColumn3d over the RC fiber section, fixed at its base,
    pushed past steel yield at the tip."""
    print("case: single 3D frame element", flush=True)
    res = []
    for m in (Model(3, 6), _oracle.oracle_model(3, 6)):
        m.nodes([1, 2], [[0.0, 0.0, 0.0], [0.0, 0.0, 3000.0]])
        m.fix_many([1], [[1] * 6])
        m.geomTransf("Linear", 1, 1.0, 0.0, 0.0)
        m.uniaxialMaterial("Steel02", 2, *meshes.STEEL02)
        m.uniaxialMaterial("Concrete02", 3, *meshes.CONCRETE02)
        m.section("Fiber", 1, meshes.rc_section_fibers(conc_tag=3, steel_tag=2), GJ=1.1e14)
        m.elements("dispBeamColumn", [1], [(1, 2)], 1, 1, 5)
        m.load(2, 1e3, 0.0, 0.0, 0.0, 0.0, 0.0)
        m.system("GpuBlockPCG", 1e-14, 10000)
        m.test("NormDispIncr", 1e-10, 40)
        m.integrator("DisplacementControl", 2, 1, 4.0)
        m.setup()
        t = []
        for step in range(10):
            assert m.analyze(1) == 0
            t.append((m.trial_disp().copy(), m.load_factor))
        res.append(t)
    for (ug, lg), (uo, lo) in zip(*res):
        assert relerr(ug, uo) <= RTOL
        assert abs(lg - lo) <= RTOL * abs(lo)


def test_two_materials_in_one_element_group():
    """stdBricks alternating between two J2 materials (different yield
    stresses) in one group: per-element material parameters."""
    print("case: two J2 materials", flush=True)
    mesh = meshes.brick_cube(3)
    tags, conn = mesh["ele_tags"], mesh["conn"]
    odd = (np.arange(len(tags)) % 2) == 1
    npts = 8 * len(tags)
    out = []
    for m in (Model(3, 3), _oracle.oracle_model(3, 3)):
        m.nodes(mesh["node_tags"], mesh["crds"])
        fx = mesh["fixed_nodes"]
        m.fix_many(fx, np.ones((len(fx), 3), np.int32))
        j2(m, 1)
        j2(m, 2, sig0_scale=0.6)
        m.elements("stdBrick", tags[~odd], conn[~odd], 1)
        m.elements("stdBrick", tags[odd], conn[odd], 2)
        m.system("GpuPCG", 1e-13, 20000)
        m.setup()
        rng = np.random.default_rng(2)
        U = 4e-3 * rng.normal(size=(m.num_nodes, 3))
        U[np.isin(mesh["node_tags"], fx)] = 0
        m.set_trial_disp(U)
        m.update()
        out.append((m.point_data("stress", npts), m.point_data("plastic", npts), m.form_unbalance()))
    for a, b in zip(*out):
        assert relerr(a, b) <= RTOL
    assert (out[1][1][:, 6] > 0).any()


def test_unknown_material_is_an_error():
    print("case: unknown material", flush=True)
    mesh = meshes.brick_cube(1)
    for m in (Model(3, 3), _oracle.oracle_model(3, 3)):
        m.nodes(mesh["node_tags"], mesh["crds"])
        with pytest.raises(XaraError) as e:
            m.elements("stdBrick", mesh["ele_tags"], mesh["conn"], 7)
            m.setup()
        assert e.value.code == -1
