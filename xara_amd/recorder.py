"""Recorder output in the reference's file format.

``NodeRecorder`` writes what ``recorder Node -file <path> [-time] -node
<tags> -dof <dofs> disp`` writes: one line per record with, when `time` is
set, the domain time (or load factor) first, then node by node the selected
DOFs, separated by spaces (commas for csv), each value in C++'s default
stream format with `precision` significant digits (6 unless
``-precision``): DataFileStream::write (DataFileStream.cpp:463-487) and the
stream precision set at open (:184).  Works with a ``Model``, an oracle
model or one member of an ``Ensemble`` (any object whose nodal
displacements come in ascending node-tag order).
"""
import numpy as np


def _fmt(v, precision):
    return f"{v:.{precision}g}"


class NodeRecorder:
    def __init__(self, path, nodes, dofs, time=True, precision=6, csv=False, node_tags=None):
        self.path, self.time, self.precision = path, time, int(precision)
        self.sep = "," if csv else " "
        self.nodes = np.asarray(nodes, np.int64)
        self.dofs = np.asarray(dofs, np.int64) - 1
        self.node_tags = None if node_tags is None else np.sort(np.asarray(node_tags))
        self._f = open(path, "w")

    def values(self, U, tags):
        idx = np.searchsorted(tags, self.nodes)
        if np.any(idx >= len(tags)) or np.any(tags[np.minimum(idx, len(tags) - 1)] != self.nodes):
            raise KeyError("recorder node not in the model")
        return U[idx][:, self.dofs].ravel()

    def record(self, model, t=0.0, U=None):
        """Append one line for the model's current trial displacements (or U)."""
        tags = self.node_tags if self.node_tags is not None else model._tags()
        vals = self.values(model.trial_disp() if U is None else np.asarray(U), tags)
        row = ([t] if self.time else []) + list(vals)
        self._f.write(self.sep.join(_fmt(float(v), self.precision) for v in row) + "\n")

    def close(self):
        if self._f:
            self._f.close()
            self._f = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
