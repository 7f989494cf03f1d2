"""Test-side loaders for the CPU oracle (oracle/liboracle.so) and the
reference-source harness (oracle/_ref/libxararef.so).  Only tests,
__graft_entry__.smoke() and bench.py's cpu_baseline use these."""
import ctypes as C
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libxararef.so")

_orc = None


def build_oracle():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboracle.so"], check=True)


def load():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE):
            build_oracle()
        from xara_amd import _lib
        lib = C.CDLL(ORACLE)
        _orc = _lib.bind(lib, "orc_", ORACLE)
        P_D, P_I = _lib.P_D, _lib.P_I
        for name, args in {
            "orc_j2_3d_path": [P_D, C.c_int, P_D, P_I, P_D, P_D],
            "orc_j2_ps_path": [P_D, C.c_int, P_D, P_I, P_D, P_D],
            "orc_uniaxial_path": [C.c_int, P_D, C.c_int, P_D, P_I, P_D, P_D],
            "orc_fiber_section_path": [C.c_int, C.c_int, P_I, P_D, P_D, P_D, P_D, C.c_double, C.c_int,
                                       C.c_int, P_D, P_I, P_D, P_D],
            "orc_rcm": [C.c_int, P_I, P_I, P_I],
        }.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = C.c_int
            setattr(_orc, name[4:], f)
    return _orc


def oracle_model(ndm, ndf=None):
    from xara_amd.model import Model
    return Model(ndm, ndf, _binding=load())


class Ref:
    def __init__(self, lib):
        from xara_amd import _lib
        P_D, P_I = _lib.P_D, _lib.P_I
        self.lib = lib
        for name, args in {
            "ref_j2_3d_path": [P_D, C.c_int, P_D, P_I, P_D, P_D],
            "ref_j2_ps_path": [P_D, C.c_int, P_D, P_I, P_D, P_D],
            "ref_uniaxial_path": [C.c_int, P_D, C.c_int, P_D, P_I, P_D, P_D],
            "ref_fiber_section_path": [C.c_int, C.c_int, P_I, P_D, P_D, P_D, P_D, C.c_double, C.c_int,
                                       C.c_int, P_D, P_I, P_D, P_D],
            "ref_rcm": [C.c_int, P_I, P_I, P_I],
        }.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = C.c_int
            setattr(self, name[4:], f)


def load_ref():
    if not os.path.exists(REF):
        return None
    return Ref(C.CDLL(REF, mode=os.RTLD_LAZY | os.RTLD_LOCAL))
