"""ctypes binding of the C-ABI declared in include/xara_gpu.h.

``load()`` opens the in-tree gfx950 library ``xara_amd/libxara_gpu.so`` and
raises if it is missing or lacks a symbol: there is no CPU fallback.
``bind(lib, prefix)`` attaches the argument/return types to any library that
exports the same ABI under another prefix.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libxara_gpu.so")

P_D = C.POINTER(C.c_double)
P_I = C.POINTER(C.c_int)
VP = C.c_void_p


class Counters(C.Structure):
    _fields_ = [
        ("updates", C.c_longlong),
        ("material_points", C.c_longlong),
        ("newton_iters", C.c_longlong),
        ("solves", C.c_longlong),
        ("cg_iters", C.c_longlong),
        ("last_cg_relres", C.c_double),
        ("last_norm", C.c_double),
        ("time_update_ms", C.c_double),
        ("time_tangent_ms", C.c_double),
        ("time_unbalance_ms", C.c_double),
        ("time_solve_ms", C.c_double),
        ("band_bw", C.c_longlong),
        ("band_fallbacks", C.c_longlong),
        ("band_reuses", C.c_longlong),
    ]


# name -> (restype, argtypes); every function declared in include/xara_gpu.h
SIGNATURES = {
    "create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(VP)]),
    "destroy": (C.c_int, [VP]),
    "last_error": (C.c_char_p, []),
    "device_name": (C.c_int, [C.c_char_p, C.c_int]),
    "nodes": (C.c_int, [VP, C.c_int, P_I, P_D]),
    "fix": (C.c_int, [VP, C.c_int, P_I, P_I]),
    "mass": (C.c_int, [VP, C.c_int, P_I, P_D]),
    "nd_j2plasticity": (C.c_int, [VP, C.c_int] + [C.c_double] * 7),
    "uniaxial_steel02": (C.c_int, [VP, C.c_int] + [C.c_double] * 11),
    "uniaxial_concrete02": (C.c_int, [VP, C.c_int] + [C.c_double] * 7),
    "uniaxial_elastic": (C.c_int, [VP, C.c_int, C.c_double]),
    "section_fiber": (C.c_int, [VP, C.c_int, C.c_int, P_I, P_D, P_D, P_D, C.c_double, C.c_int]),
    "geom_transf_linear": (C.c_int, [VP, C.c_int, P_D]),
    "elements_brick": (C.c_int, [VP, C.c_int, P_I, P_I, C.c_int, C.c_double, C.c_double, C.c_double]),
    "elements_quad": (C.c_int, [VP, C.c_int, P_I, P_I, C.c_double, C.c_int, C.c_double, C.c_double,
                                C.c_double, C.c_double]),
    "elements_dispbeam": (C.c_int, [VP, C.c_int, P_I, P_I, C.c_int, C.c_int, C.c_int, C.c_double]),
    "elements_forcebeam": (C.c_int, [VP, C.c_int, P_I, P_I, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double]),
    "nodal_loads": (C.c_int, [VP, C.c_int, P_I, P_D]),
    "uniform_excitation": (C.c_int, [VP, C.c_int, C.c_int, C.c_double, P_D, C.c_double]),
    "rayleigh": (C.c_int, [VP, C.c_double, C.c_double, C.c_double, C.c_double]),
    "system": (C.c_int, [VP, C.c_int, C.c_double, C.c_int]),
    "add_a": (C.c_int, [VP, C.c_int, C.c_int, P_D, P_I]),
    "solve_noalloc": (C.c_int, [VP, P_D, P_D]),
    "numberer": (C.c_int, [VP, C.c_int]),
    "spmv_config": (C.c_int, [VP, C.c_int, C.c_int]),
    "cg_config": (C.c_int, [VP, C.c_int, C.c_int]),
    "test": (C.c_int, [VP, C.c_int, C.c_double, C.c_int]),
    "integrator_load_control": (C.c_int, [VP, C.c_double]),
    "integrator_disp_control": (C.c_int, [VP, C.c_int, C.c_int, C.c_double]),
    "integrator_newmark": (C.c_int, [VP, C.c_double, C.c_double]),
    "analyze": (C.c_int, [VP, C.c_int, C.c_double]),
    "setup": (C.c_int, [VP]),
    "num_nodes": (C.c_int, [VP]),
    "num_elements": (C.c_int, [VP]),
    "num_eqn": (C.c_int, [VP]),
    "csr_nnz": (C.c_longlong, [VP]),
    "node_eqns": (C.c_int, [VP, P_I]),
    "csr_pattern": (C.c_int, [VP, P_I, P_I]),
    "set_trial_disp": (C.c_int, [VP, P_D]),
    "get_trial_disp": (C.c_int, [VP, P_D]),
    "incr_disp": (C.c_int, [VP, P_D]),
    "update": (C.c_int, [VP]),
    "form_tangent": (C.c_int, [VP, P_D]),
    "form_unbalance": (C.c_int, [VP, P_D]),
    "solve": (C.c_int, [VP, P_D, P_D]),
    "spmv": (C.c_int, [VP, P_D, P_D]),
    "commit": (C.c_int, [VP]),
    "revert_to_last_commit": (C.c_int, [VP]),
    "revert_to_start": (C.c_int, [VP]),
    "set_load_factor": (C.c_int, [VP, C.c_double]),
    "get_load_factor": (C.c_double, [VP]),
    "get_point_data": (C.c_int, [VP, C.c_int, P_D, P_I]),
    "element_tangent": (C.c_int, [VP, C.c_int, P_D, P_I]),
    "element_resisting_force": (C.c_int, [VP, C.c_int, P_D, P_I]),
    "get_counters": (C.c_int, [VP, C.POINTER(Counters)]),
    "time_op": (C.c_int, [VP, C.c_int, C.c_int, P_D]),
    "state_size": (C.c_longlong, [VP, C.c_int]),
    "get_state": (C.c_int, [VP, C.c_int, P_D]),
    "set_state": (C.c_int, [VP, C.c_int, P_D]),
    # batched ensemble (GPU library only)
    "ensemble_create": (C.c_int, [C.c_int, C.POINTER(VP)]),
    "ensemble_destroy": (C.c_int, [VP]),
    "ensemble_add": (C.c_int, [VP, VP]),
    "ensemble_setup": (C.c_int, [VP]),
    "ensemble_num_models": (C.c_int, [VP]),
    "ensemble_tune": (C.c_int, [VP, P_D, C.c_int, P_I]),
    "ensemble_set_mapping": (C.c_int, [VP, C.c_int, C.c_int, C.c_int, C.c_int]),
    "ensemble_analyze": (C.c_int, [VP, C.c_int, C.c_double, P_I]),
    "ensemble_model_disp": (C.c_int, [VP, C.c_int, P_D, P_I]),
    "ensemble_status": (C.c_int, [VP, P_I, C.POINTER(C.c_longlong)]),
    # adjoint (GPU library only)
    "adjoint_loss": (C.c_int, [VP, C.c_int, P_I, C.c_int]),
    "adjoint_forward": (C.c_int, [VP, C.c_int, C.c_double, C.c_int, P_D]),
    "adjoint_backward": (C.c_int, [VP, P_D, P_D]),
    "adjoint_stats": (C.c_int, [VP, P_D, P_D, C.POINTER(C.c_longlong)]),
    "material_vjp": (C.c_int, [C.c_int, P_D, P_D, C.c_int, C.c_double, P_D, P_D, P_D, P_D]),
}
GPU_ONLY = {k for k in SIGNATURES if k.startswith(("ensemble_", "adjoint_"))} | {
    "device_name", "material_vjp", "spmv_config", "cg_config", "solve_noalloc"}


class Binding:
    """Typed function table of one library exporting the ABI under `prefix`."""

    def __init__(self, lib, prefix, path=None):
        self.lib = lib
        self.prefix = prefix
        self.path = path
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, prefix + name, None)
            if fn is None:
                if prefix != "xg_" and name in GPU_ONLY:
                    continue
                raise RuntimeError(f"{path or lib}: missing symbol {prefix + name}")
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def error(self):
        msg = self.last_error()
        return msg.decode() if msg else ""


def bind(lib, prefix, path=None):
    return Binding(lib, prefix, path)


_GPU = None


def load():
    """Return the binding of the gfx950 library; raise if it is not built."""
    global _GPU
    if _GPU is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is missing: build it with `python -m xara_amd.build` "
                "(there is no CPU fallback)")
        _GPU = Binding(C.CDLL(LIBPATH), "xg_", LIBPATH)
    return _GPU
