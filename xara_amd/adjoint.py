"""Gradients of a frame transient's displacement loss with respect to
material parameters, by the GPU adjoint (xg_adjoint_* in the C-ABI).

``FrameTransient`` defines the problem: a builder that defines a frame model
(Newmark, nodal masses, UniformExcitation, fiber sections) from a parameter
array theta, the loss nodes and DOF (loss = sum over steps of the squared
displacements), the number of steps, dt and the checkpoint interval.
``transient_loss(theta, problem)`` is the torch.autograd.Function: forward
runs the transient on the GPU and returns the loss; backward runs the
adjoint (recomputing each checkpoint segment) and returns dloss/dtheta.

The default parameter layout is the one of
``meshes.build_rc_frame_per_element``: theta[e] = (fy, E, b, fc) of element
e + 1, whose Steel02 / Concrete02 materials are the e-th in ascending tag
order.  Other layouts pass ``grad_map(grad_steel, grad_conc) -> dtheta``.
"""
import ctypes as C

import numpy as np
import torch

from . import _lib
from .model import Model, XaraError


def per_element_grad(grad_steel, grad_conc):
    """(n, 3) Steel02 (Fy, E0, b) and (n,) Concrete02 fc gradients in
    material-tag order -> (n, 4) rows (fy, E, b, fc)."""
    return np.concatenate([grad_steel, grad_conc[:, None]], axis=1)


class FrameTransient:
    def __init__(self, ndm, ndf, build, loss_nodes, nsteps, dt, dof=1, ckpt=25, device=0, grad_map=None):
        self.ndm, self.ndf, self.build = ndm, ndf, build
        self.loss_nodes = np.ascontiguousarray(loss_nodes, np.int32)
        self.nsteps, self.dt, self.dof, self.ckpt = int(nsteps), float(dt), int(dof), int(ckpt)
        self.device = device
        self.grad_map = grad_map or per_element_grad
        self.model = None
        self.token = 0
        self._b = _lib.load()

    def _check(self, rc):
        if rc < 0:
            raise XaraError(rc, self._b.error())
        return rc

    def run(self, theta):
        """Define the model from theta, run the transient; returns the loss."""
        if self.model is not None:
            self.model.close()
        m = Model(self.ndm, self.ndf, device=self.device)
        self.build(m, np.asarray(theta, np.float64))
        m.setup()
        h = m._h
        self._check(self._b.adjoint_loss(h, len(self.loss_nodes), self.loss_nodes.ctypes.data_as(_lib.P_I),
                                         self.dof))
        loss = C.c_double(0)
        self._check(self._b.adjoint_forward(h, self.nsteps, self.dt, self.ckpt, C.byref(loss)))
        self.model = m
        self.token += 1
        return loss.value

    def gradient(self):
        """dloss/dtheta of the last run (adjoint sweep)."""
        if self.model is None:
            raise XaraError(-2, "run the forward problem first")
        ns = len(self.model.material_tags("Steel02"))
        nc = len(self.model.material_tags("Concrete02"))
        gs, gc = np.zeros(3 * max(ns, 1)), np.zeros(max(nc, 1))
        self._check(self._b.adjoint_backward(self.model._h, gs.ctypes.data_as(_lib.P_D),
                                             gc.ctypes.data_as(_lib.P_D)))
        return self.grad_map(gs[: 3 * ns].reshape(ns, 3), gc[:nc])

    def stats(self):
        f, b, cg = C.c_double(0), C.c_double(0), C.c_longlong(0)
        self._check(self._b.adjoint_stats(self.model._h, C.byref(f), C.byref(b), C.byref(cg)))
        return {"forward_ms": f.value, "backward_ms": b.value, "backward_cg_iters": cg.value}

    def close(self):
        if self.model is not None:
            self.model.close()
            self.model = None


class _TransientLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, problem):
        th = theta.detach().to("cpu", torch.float64).numpy().copy()  # own copy: theta may change in place
        loss = problem.run(th)
        ctx.problem, ctx.token, ctx.theta = problem, problem.token, th
        return torch.tensor(loss, dtype=theta.dtype, device=theta.device)

    @staticmethod
    def backward(ctx, gout):
        p = ctx.problem
        if p.token != ctx.token:  # the problem ran another forward since: rerun this one
            p.run(ctx.theta)
        g = torch.as_tensor(p.gradient(), dtype=gout.dtype, device=gout.device)
        return gout * g, None


def transient_loss(theta, problem):
    """Differentiable loss of `problem` at parameters `theta` (torch tensor)."""
    return _TransientLoss.apply(theta, problem)
