"""The reference-side binding (xara_amd/adapter/XaraGpuAdapter.{h,cpp}:
GpuElement, GpuLinSOE + GpuPCGSolver, GpuIntegrator<LoadControl /
DisplacementControl / Newmark>) compiles against the reference's own
headers, as a maintainer would build it inside the reference tree (build
container only: the reference tree is not on the GPU host)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference/SRC"
ADAPTER = os.path.join(ROOT, "xara_amd", "adapter", "XaraGpuAdapter.cpp")


def ref_includes():
    dirs = [os.path.join(dp) for dp, _, _ in os.walk(REF) if "/opensees" not in dp]
    return [f"-I{d}" for d in sorted(dirs)] + ["-I/opt/conda/include"]


@pytest.mark.skipif(not os.path.isdir(REF), reason="needs the reference tree")
def test_adapter_compiles_against_reference_headers():
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Wno-unused-parameter", "-Werror",
           *ref_includes(), f"-I{os.path.join(ROOT, 'include')}", ADAPTER]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]


def test_adapter_overrides_only_what_the_reference_calls():
    """The integrator wrapper overrides exactly the two element loops the
    reference's algorithms reach (IncrementalIntegrator::formTangent and
    formElementResidual); the SOE implements the LinearSOE interface."""
    src = open(ADAPTER).read()
    hdr = open(ADAPTER.replace(".cpp", ".h")).read()
    assert "int formTangent(int statFlag) override;" in hdr
    assert "int formElementResidual() override;" in hdr
    for m in ("setSize(Graph& theGraph)", "addA(const Matrix& m, const ID& id, double fact",
              "addB(const Vector& v, const ID& id, double fact", "setB(const Vector& v, double fact",
              "zeroA()", "zeroB()", "getX()", "getB()", "normRHS()", "setX(int loc, double value)"):
        assert m in hdr, m
    for call in ("xg_form_tangent", "xg_form_unbalance", "xg_solve", "xg_update", "xg_commit", "xg_add_a"):
        assert call in src, call
