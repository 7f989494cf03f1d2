"""xg_solve_noalloc (the solve whose PCG buffers come from a fixed device
array, for callers that cannot allocate) against xg_solve through the
C-ABI: the same solution, bit for bit, on two calls in a row -- a 1,944-
equation cube (kernel-per-step PCG) and then a smaller 144-equation one
(whole solve in one workgroup) -- so contents the first call left in the
shared array cannot leak into the second."""
import ctypes as C

import numpy as np
import pytest

from xara_amd import _lib, meshes
from xara_amd.model import Model

pytestmark = pytest.mark.gpu


def cube_system(n):
    mesh = meshes.brick_cube(n)
    m = Model(3, 3)
    meshes.build_brick_cube(m, mesh)
    m.system("GpuBlockPCG", 1e-12, 20000)
    m.setup()
    x, y, z = mesh["crds"].T
    U = np.zeros((len(x), 3))
    U[:, 2] = -3e-3 * z * (1.0 + 0.5 * np.sin(3 * x))
    U[:, 0] = 1e-3 * z * y
    m.set_trial_disp(U)
    m.update()
    m.form_tangent(values=False)
    m.set_load_factor(50.0)
    B = m.form_unbalance()
    return m, B


@pytest.mark.parametrize("n", [(8, 3)])
def test_solve_noalloc_matches_solve_twice(n):
    b = _lib.load()
    for size in n:
        print(f"case: solve_noalloc cube {size}^3", flush=True)
        m, B = cube_system(size)
        X1 = m.solve(B)
        X2 = np.zeros(m.neq)
        rc = b.solve_noalloc(m._h, B.ctypes.data_as(_lib.P_D), X2.ctypes.data_as(_lib.P_D))
        assert rc == 0, b.error()
        assert np.array_equal(X1, X2)
        assert np.abs(X1).max() > 0
        m.close()
