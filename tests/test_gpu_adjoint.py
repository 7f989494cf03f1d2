"""GPU adjoint of a frame transient (loss = sum over steps of squared roof
displacements; parameters = per-element Steel02 fy, E, b and Concrete02 fc):
checked against the oracle's direct differentiation (forward-mode, DDM
structure) along random and unit directions, against central finite
differences of the GPU forward run (h = 1e-6 relative), for invariance to
the checkpoint interval, and through the torch.autograd.Function."""
import numpy as np
import pytest
import torch

from tests import _oracle
from xara_amd import ground_motion, meshes
from xara_amd.adjoint import FrameTransient, transient_loss

pytestmark = pytest.mark.gpu
DT = 0.01


def problem_2d(nsteps=120, ckpt=25):
    fr = meshes.rc_frame_2d(2, 1)
    ag = ground_motion.kanai_tajimi(1, nsteps + 5, DT, 1.2)

    def build(m, theta):
        meshes.define_transient_frame(m, fr, theta, ag, DT, tol=1e-12, maxit=40)
        m.system("GpuBlockPCG", 1e-14, 100000)

    ne = len(meshes.frame_elements(fr)[0])
    return fr, ag, build, ne, FrameTransient(2, 3, build, meshes.roof_nodes(fr), nsteps, DT, ckpt=ckpt)


def problem_3d(nsteps=80, ckpt=20):
    fr = meshes.rc_frame_3d(1, 1, 2)
    ag = ground_motion.kanai_tajimi(0, nsteps + 5, DT, 1.2)

    def build(m, theta):
        meshes.define_transient_frame(m, fr, theta, ag, DT, tol=1e-12, maxit=40)
        m.system("GpuBlockPCG", 1e-14, 100000)

    ne = len(meshes.frame_elements(fr)[0])
    return fr, ag, build, ne, FrameTransient(3, 6, build, meshes.roof_nodes(fr), nsteps, DT, ckpt=ckpt)


def oracle_ddm(ndm, ndf, fr, ag, theta, direction, nsteps):
    o = _oracle.oracle_model(ndm, ndf)
    meshes.define_transient_frame(o, fr, theta, ag, DT, tol=1e-12, maxit=40)
    o.setup()
    return _oracle.ddm_transient(o, nsteps, DT, direction[:, :3], direction[:, 3], meshes.roof_nodes(fr), 1)


@pytest.mark.parametrize("dim", [2, 3])
def test_adjoint_matches_oracle_ddm(dim):
    print(f"case: adjoint vs oracle DDM, {dim}D frame", flush=True)
    fr, ag, build, ne, prob = problem_2d() if dim == 2 else problem_3d()
    ndm, ndf = (2, 3) if dim == 2 else (3, 6)
    theta = meshes.default_theta(ne)
    loss = prob.run(theta)
    grad = prob.gradient()
    rng = np.random.default_rng(3)
    dirs = [rng.standard_normal((ne, 4)) * theta]
    for k in range(4):  # every parameter kind of every element together
        d = np.zeros((ne, 4))
        d[:, k] = theta[:, k]
        dirs.append(d)
    for d in dirs:
        L, dL = oracle_ddm(ndm, ndf, fr, ag, theta, d, prob.nsteps)
        assert abs(loss - L) <= 1e-9 * abs(L)
        adj = float(np.sum(grad * d))
        assert abs(adj - dL) <= 1e-6 * abs(dL) + 1e-12 * abs(L), (adj, dL)
    # the motion drove the steel past yield: fy matters
    assert abs(np.sum(grad[:, 0] * theta[:, 0])) > 1e-6 * loss


def test_adjoint_matches_central_differences():
    print("case: adjoint vs central FD", flush=True)
    fr, ag, build, ne, prob = problem_2d()
    theta = meshes.default_theta(ne)
    loss = prob.run(theta)
    grad = prob.gradient()
    for e, k in [(0, 0), (0, 1), (2, 3), (1, 2)]:
        h = 1e-6 * theta[e, k]
        tp, tm = theta.copy(), theta.copy()
        tp[e, k] += h
        tm[e, k] -= h
        fd = (prob.run(tp) - prob.run(tm)) / (2 * h)
        scale = max(abs(grad[e, k]), 1e-3 * np.abs(grad[:, k]).max())
        # truncation O(h^2) plus the rounding of the loss itself (~eps |loss| / h)
        tol = 1e-5 * scale + 50 * np.finfo(float).eps * abs(loss) / abs(h)
        assert abs(fd - grad[e, k]) <= tol, (e, k, fd, grad[e, k])


def test_checkpoint_interval_does_not_change_gradient():
    print("case: adjoint checkpoint intervals", flush=True)
    grads = []
    for ckpt in (1, 7, 120):
        fr, ag, build, ne, prob = problem_2d(ckpt=ckpt)
        prob.run(meshes.default_theta(ne))
        grads.append(prob.gradient())
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


def test_autograd_function():
    print("case: adjoint through torch.autograd", flush=True)
    fr, ag, build, ne, prob = problem_2d(nsteps=60)
    theta = torch.tensor(meshes.default_theta(ne), dtype=torch.float64, requires_grad=True)
    loss = transient_loss(theta, prob)
    (3.0 * loss).backward()
    ref = FrameTransient(2, 3, build, meshes.roof_nodes(fr), 60, DT)
    ref.run(theta.detach().numpy())
    assert np.array_equal(theta.grad.numpy(), 3.0 * ref.gradient())
    st = prob.stats()
    assert st["forward_ms"] > 0 and st["backward_ms"] > 0
