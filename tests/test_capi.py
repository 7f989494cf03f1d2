"""The C-ABI library loads and exports every entry point declared in
include/xara_gpu.h; the Python binding types all of them.  No compute calls
(these run without a GPU)."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xara_gpu.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xg_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared()
    for must in ("xg_create", "xg_setup", "xg_update", "xg_form_tangent", "xg_form_unbalance",
                 "xg_solve", "xg_commit", "xg_analyze", "xg_csr_pattern", "xg_node_eqns"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from xara_amd import _lib
    if not os.path.exists(_lib.LIBPATH):
        pytest.fail("libxara_gpu.so is not built: run `python -m xara_amd.build`")
    lib = C.CDLL(_lib.LIBPATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    b = _lib.load()
    assert set("xg_" + k for k in _lib.SIGNATURES) == set(declared())
    assert b.path == _lib.LIBPATH


def test_oracle_exports_the_same_abi():
    from tests import _oracle
    lib = C.CDLL(_oracle.ORACLE) if os.path.exists(_oracle.ORACLE) else None
    if lib is None:
        _oracle.build_oracle()
        lib = C.CDLL(_oracle.ORACLE)
    from xara_amd import _lib
    missing = [n for n in declared() if n[3:] not in _lib.GPU_ONLY and not hasattr(lib, "orc_" + n[3:])]
    assert not missing, missing


def test_no_gpu_means_a_loud_error():
    """Without a visible GPU the product path raises instead of computing on the CPU."""
    from xara_amd.model import Model, XaraError
    import subprocess
    probe = subprocess.run(["bash", "-c", "ls /dev/kfd >/dev/null 2>&1"], capture_output=True)
    if probe.returncode == 0:
        pytest.skip("a GPU device node is present")
    with pytest.raises(XaraError):
        Model(3, 3)
