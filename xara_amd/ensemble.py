"""Batched ensemble of independent 2D frame models on one GPU.

Each member is an ordinary ``Model`` (2D dispBeamColumn frame with fiber
sections, nodal masses, UniformExcitation, Newmark, NormDispIncr test)
defined through the usual commands; ``Ensemble.add`` snapshots its
definition.  ``analyze`` advances every member by Newmark steps with one
persistent launch per Newton iteration, captured as one hipGraph per step.
"""
import ctypes as C

import numpy as np

from . import _lib
from .model import XaraError


class Ensemble:
    def __init__(self, device=0):
        self._b = _lib.load()
        h = C.c_void_p()
        self._check(self._b.ensemble_create(device, C.byref(h)))
        self._h = h
        self._nnodes = []

    def _check(self, rc):
        if rc < 0:
            raise XaraError(rc, self._b.error())
        return rc

    def close(self):
        if getattr(self, "_h", None):
            self._b.ensemble_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, model):
        self._check(self._b.ensemble_add(self._h, model._h))
        self._nnodes.append(model.num_nodes)

    def setup(self):
        self._check(self._b.ensemble_setup(self._h))

    @property
    def size(self):
        return self._b.ensemble_num_models(self._h)

    def tune(self):
        cap = 64
        table = np.zeros(cap * 4)
        n = C.c_int(0)
        self._check(self._b.ensemble_tune(self._h, table.ctypes.data_as(_lib.P_D), cap, C.byref(n)))
        return table[: 4 * n.value].reshape(n.value, 4)

    def set_mapping(self, block=256, grid=0, order=0, use_graph=True):
        self._check(self._b.ensemble_set_mapping(self._h, block, grid, order, 1 if use_graph else 0))

    def analyze(self, nsteps, dt):
        stopped = C.c_int(0)
        self._check(self._b.ensemble_analyze(self._h, nsteps, dt, C.byref(stopped)))
        return stopped.value

    def disp(self, k):
        U = np.zeros((self._nnodes[k], 3))
        nn = C.c_int(0)
        self._check(self._b.ensemble_model_disp(self._h, k, U.ctypes.data_as(_lib.P_D), C.byref(nn)))
        return U

    def status(self):
        n = self.size
        st = np.zeros(n, np.int32)
        it = np.zeros(n, np.int64)
        self._check(self._b.ensemble_status(self._h, st.ctypes.data_as(_lib.P_I),
                                            it.ctypes.data_as(C.POINTER(C.c_longlong))))
        return st, it
