"""The oracle's direct differentiation of a frame transient (forward mode,
Newmark::computeSensitivities structure with the Steel02/Concrete02 partials
the reference lacks) against central finite differences of the oracle's own
transient: the checker of the GPU adjoint is pinned here, on the CPU."""
import numpy as np
import pytest

from tests import _oracle
from xara_amd import ground_motion, meshes

DT = 0.01


def run(fr, ag, theta, nsteps, direction=None):
    ndm, ndf = (3, 6) if "beams_x" in fr else (2, 3)
    o = _oracle.oracle_model(ndm, ndf)
    meshes.define_transient_frame(o, fr, theta, ag, DT, tol=1e-12, maxit=40)
    o.setup()
    d = np.zeros_like(theta) if direction is None else direction
    return _oracle.ddm_transient(o, nsteps, DT, d[:, :3], d[:, 3], meshes.roof_nodes(fr), 1)


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_ddm_matches_central_differences(case):
    if case == "2d":
        fr, nsteps, ag = meshes.rc_frame_2d(2, 1), 150, ground_motion.kanai_tajimi(1, 160, DT, 1.2)
    else:
        fr, nsteps, ag = meshes.rc_frame_3d(1, 1, 2), 60, ground_motion.kanai_tajimi(0, 70, DT, 1.2)
    ne = len(meshes.frame_elements(fr)[0])
    theta = meshes.default_theta(ne)
    rng = np.random.default_rng(2)
    dirs = [rng.standard_normal((ne, 4)) * theta * 0.1]
    for k in range(4):
        d = np.zeros((ne, 4))
        d[:, k] = theta[:, k]
        dirs.append(d)
    for d in dirs:
        L, dL = run(fr, ag, theta, nsteps, d)
        h = 1e-5
        fd = (run(fr, ag, theta + h * d, nsteps)[0] - run(fr, ag, theta - h * d, nsteps)[0]) / (2 * h)
        tol = 1e-6 * abs(fd) + 100 * np.finfo(float).eps * abs(L) / h
        assert abs(dL - fd) <= tol, (dL, fd)
    if case == "2d":
        # the motion yields the steel: fy and b have a first-order effect
        d = np.zeros((ne, 4))
        d[:, 0] = theta[:, 0]
        assert abs(run(fr, ag, theta, nsteps, d)[1]) > 1e-6 * L
