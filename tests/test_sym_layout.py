"""CPU checks of the block-symmetric SpMV layout (xg_sym_layout.h, the host
code that builds the index arrays the GPU kernels read): every index in
bounds, each neighbour block mapped to the stored block that holds it (or
its transpose), each off-diagonal stored block used by exactly its two
block rows, and the product through the layout equal to the CSR product on
the oracle's assembled tangents.  The layout code is compiled from the
library's header with g++ and run on patterns of the oracle's numbering."""
import os
import subprocess

import numpy as np
import pytest

from tests import _oracle
from xara_amd import meshes

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("symlayout") / "sym_layout_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "sym_layout_driver.cpp")],
                   check=True)
    return exe


def run_layout(driver, tmp_path, M, rowptr, nbc):
    nbr = (len(rowptr) - 1) // M
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    np.concatenate([[M, nbr, len(nbc)], rowptr, nbc]).astype(np.int32).tofile(fin)
    subprocess.run([driver, fin, fout], check=True)
    out = np.fromfile(fout, dtype=np.int32)
    ok, nup = int(out[0]), int(out[1])
    if not ok:
        return None
    o = 2
    ptr = out[o:o + nbr + 1]; o += nbr + 1
    col = out[o:o + len(nbc)]; o += len(nbc)
    blk = out[o:o + len(nbc)]; o += len(nbc)
    src = out[o:o + nup]; o += nup
    ln = out[o:o + nup]; o += nup
    assert o == len(out)
    return dict(nup=nup, ptr=ptr, col=col, blk=blk, src=src, len=ln)


def node_block_columns(M, rowptr, colind):
    """One block column per (block row, neighbour), as the library forms it."""
    nbr = (len(rowptr) - 1) // M
    out = []
    for I in range(nbr):
        b, e = rowptr[M * I], rowptr[M * I + 1]
        assert b == M * M * len(out)
        cols = colind[b:e]
        assert np.array_equal(cols.reshape(-1, M) - cols[::M, None], np.tile(np.arange(M), (len(cols) // M, 1)))
        out.extend(cols[::M] // M)
    return np.array(out, np.int32)


def models():
    mesh = meshes.brick_cube(4)
    m = _oracle.oracle_model(3, 3)
    meshes.build_brick_cube(m, mesh, pressure=1.0)
    yield "brick 4^3", 3, m
    mesh = meshes.quad_strip(10, 3)
    m = _oracle.oracle_model(2, 2)
    meshes.build_quad_strip(m, mesh)
    yield "quad 10x3", 2, m


@pytest.mark.parametrize("case", [0, 1])
def test_layout_indices_and_product(driver, tmp_path, case):
    name, M, m = list(models())[case]
    m.setup()
    rng = np.random.default_rng(5)
    m.set_trial_disp(1e-3 * rng.normal(size=(m.num_nodes, m.ndf)))
    m.update()
    vals = m.form_tangent()
    rowptr, colind = m.csr_pattern()
    nnz = len(colind)
    nbc = node_block_columns(M, rowptr, colind)
    L = run_layout(driver, tmp_path, M, rowptr, nbc)
    assert L is not None, name
    nbr = (len(rowptr) - 1) // M
    ptr, col, blk, src, ln, nup = L["ptr"], L["col"], L["blk"], L["src"], L["len"], L["nup"]
    assert ptr[0] == 0 and ptr[-1] == len(nbc) and np.all(np.diff(ptr) > 0)
    assert np.array_equal(col, nbc)
    assert np.all((blk >= 0) & (blk < nup))
    assert np.all(src >= 0) and np.all(src + (M - 1) * ln + (M - 1) < nnz)
    uses = np.zeros(nup, int)
    for I in range(nbr):
        for e in range(ptr[I], ptr[I + 1]):
            J, k = int(col[e]), int(blk[e])
            uses[k] += 1
            r, c = (I, J) if J >= I else (J, I)  # the stored block is (r, c), c >= r
            b0, e0 = rowptr[M * r], rowptr[M * r + 1]
            assert b0 <= src[k] and src[k] + M - 1 < e0 and ln[k] == e0 - b0
            for a in range(M):  # row M*r + a of the block, its M columns
                s = src[k] + a * ln[k]
                assert rowptr[M * r + a] <= s < rowptr[M * r + a + 1]
                assert np.array_equal(colind[s:s + M], M * c + np.arange(M))
    # stored blocks: diagonal ones used once, the others by both block rows
    diag = np.array([col[ptr[I]:ptr[I + 1]].tolist().index(I) + ptr[I] for I in range(nbr)])
    assert np.all(uses[blk[diag]] == 1)
    assert np.all(np.delete(uses, blk[diag]) == 2)
    # the product through the layout against the CSR product
    sv = np.stack([vals[src + a * ln + c] for a in range(M) for c in range(M)], axis=1).reshape(-1, M, M)
    x = rng.normal(size=len(rowptr) - 1)
    y = np.zeros_like(x)
    rows = np.repeat(np.arange(nbr), np.diff(ptr))
    B = sv[blk]
    lower = col < rows
    B[lower] = np.transpose(B[lower], (0, 2, 1))
    np.add.at(y.reshape(nbr, M), rows, np.einsum("eij,ej->ei", B, x.reshape(-1, M)[col]))
    y_csr = np.zeros_like(x)
    np.add.at(y_csr, np.repeat(np.arange(len(x)), np.diff(rowptr)), vals * x[colind])
    assert np.max(np.abs(y - y_csr)) <= 1e-12 * np.max(np.abs(y_csr))


def test_unsymmetric_block_pattern_is_refused(driver, tmp_path):
    """A block row that lists a neighbour which does not list it back."""
    M = 2
    # block rows 0, 1, 2; block 0 lists 0, 1, 2; block 1 lists 0, 1; block 2
    # lists 2 only (missing its partner 0)
    lists = [[0, 1, 2], [0, 1], [2]]
    rowptr = [0]
    nbc = []
    for lst in lists:
        for _ in range(M):
            rowptr.append(rowptr[-1] + M * len(lst))
        nbc.extend(lst)
    assert run_layout(driver, tmp_path, M, np.array(rowptr), np.array(nbc)) is None
    lists[2] = [0, 2]
    rowptr, nbc = [0], []
    for lst in lists:
        for _ in range(M):
            rowptr.append(rowptr[-1] + M * len(lst))
        nbc.extend(lst)
    assert run_layout(driver, tmp_path, M, np.array(rowptr), np.array(nbc)) is not None
