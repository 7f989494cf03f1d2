"""CPU checks of the block-symmetric SpMV layout (xg_sym_layout.h, the host
code that builds the index arrays the GPU kernels read): every index in
bounds, each neighbour block mapped to the stored block that holds it (or
its transpose), each off-diagonal stored block used by exactly its two
block rows, the stored blocks numbered in sweep order, and the product
through the layout equal to the CSR product on the oracle's assembled
tangents -- for the equation order and for swept orders of the block rows
(the Morton order of the nodes, a random one).  The layout code is compiled
from the library's header with g++ and run on patterns of the oracle's
numbering."""
import os
import subprocess

import numpy as np
import pytest

from tests import _oracle
from xara_amd import meshes

HERE = os.path.dirname(os.path.abspath(__file__))
LOWER = np.int32(-2 ** 31)


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("symlayout") / "sym_layout_driver")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "sym_layout_driver.cpp")],
                   check=True)
    return exe


def run_layout(driver, tmp_path, M, rowptr, nbc, ord=None):
    nbr = (len(rowptr) - 1) // M
    fin, fout, ford = str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), str(tmp_path / "ord.bin")
    np.concatenate([[M, nbr, len(nbc)], rowptr, nbc]).astype(np.int32).tofile(fin)
    args = [driver, fin, fout]
    if ord is not None:
        np.asarray(ord, np.int32).tofile(ford)
        args.append(ford)
    subprocess.run(args, check=True)
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


def run_morton(driver, tmp_path, xyz):
    xyz = np.asarray(xyz, np.float64)
    nbr, dim = xyz.shape
    fin, fout = str(tmp_path / "pts.bin"), str(tmp_path / "ord.bin")
    with open(fin, "wb") as f:
        np.array([nbr, dim], np.int32).tofile(f)
        xyz.tofile(f)
    subprocess.run([driver, "--morton", fin, fout], check=True)
    return np.fromfile(fout, dtype=np.int32)


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
    yield "brick 4^3", 3, m, mesh["crds"]
    mesh = meshes.quad_strip(10, 3)
    m = _oracle.oracle_model(2, 2)
    meshes.build_quad_strip(m, mesh)
    yield "quad 10x3", 2, m, mesh["crds"]


def block_row_coords(m, M, crds):
    """Coordinates of each block row's node (all of a node's DOFs free)."""
    ne = m.node_eqns()
    free = ne[:, 0] >= 0
    nbr = int(ne[free, 0].max()) // M + 1
    xyz = np.zeros((nbr, crds.shape[1]))
    xyz[ne[free, 0] // M] = crds[free]
    return xyz


@pytest.mark.parametrize("order", ["equation", "morton", "random"])
@pytest.mark.parametrize("case", [0, 1])
def test_layout_indices_and_product(driver, tmp_path, case, order):
    name, M, m, crds = list(models())[case]
    m.setup()
    rng = np.random.default_rng(5)
    m.set_trial_disp(1e-3 * rng.normal(size=(m.num_nodes, m.ndf)))
    m.update()
    vals = m.form_tangent()
    rowptr, colind = m.csr_pattern()
    nnz = len(colind)
    nbc = node_block_columns(M, rowptr, colind)
    nbr = (len(rowptr) - 1) // M
    if order == "equation":
        ord = np.arange(nbr)
        L = run_layout(driver, tmp_path, M, rowptr, nbc)
    else:
        ord = (run_morton(driver, tmp_path, block_row_coords(m, M, crds)) if order == "morton"
               else rng.permutation(nbr))
        L = run_layout(driver, tmp_path, M, rowptr, nbc, ord)
    assert L is not None, name
    pos = np.empty(nbr, int)
    pos[ord] = np.arange(nbr)
    first = np.concatenate([[0], np.cumsum(np.diff(rowptr)[::M] // M)])
    ptr, col, blk, src, ln, nup = L["ptr"], L["col"], L["blk"], L["src"], L["len"], L["nup"]
    assert ptr[0] == 0 and ptr[-1] == len(nbc) and np.all(np.diff(ptr) > 0)
    lower = blk < 0
    k_of = blk & 0x7FFFFFFF
    assert np.all(k_of < nup)
    uses = np.zeros(nup, int)
    own = []
    for p in range(nbr):
        I = int(ord[p])
        seg = slice(ptr[p], ptr[p + 1])
        # block row I's neighbours, ascending, as in the CSR pattern
        assert np.array_equal(col[seg], nbc[first[I]:first[I + 1]])
        for e in range(ptr[p], ptr[p + 1]):
            J, k = int(col[e]), int(k_of[e])
            assert lower[e] == (pos[J] < p)
            uses[k] += 1
            if not lower[e]:
                own.append(k)
            r, c = (J, I) if lower[e] else (I, J)  # the stored block is (r, c) in the rows of r
            b0, e0 = rowptr[M * r], rowptr[M * r + 1]
            assert b0 <= src[k] and src[k] + M - 1 < e0 and ln[k] == e0 - b0
            for a in range(M):  # row M*r + a of the block, its M columns
                s_ = src[k] + a * ln[k]
                assert rowptr[M * r + a] <= s_ < rowptr[M * r + a + 1]
                assert np.array_equal(colind[s_:s_ + M], M * c + np.arange(M))
    # a sweep meets its own stored blocks in storage order
    assert own == list(range(nup))
    # stored blocks: diagonal ones used once, the others by both block rows
    diag = np.array([col[ptr[p]:ptr[p + 1]].tolist().index(ord[p]) + ptr[p] for p in range(nbr)])
    assert np.all(uses[k_of[diag]] == 1)
    assert np.all(np.delete(uses, k_of[diag]) == 2)
    # the product through the layout against the CSR product
    sv = np.stack([vals[src + a * ln + c] for a in range(M) for c in range(M)], axis=1).reshape(-1, M, M)
    x = rng.normal(size=len(rowptr) - 1)
    y = np.zeros_like(x)
    rows = np.repeat(ord, np.diff(ptr))
    B = sv[k_of]
    B[lower] = np.transpose(B[lower], (0, 2, 1))
    np.add.at(y.reshape(nbr, M), rows, np.einsum("eij,ej->ei", B, x.reshape(-1, M)[col]))
    y_csr = np.zeros_like(x)
    np.add.at(y_csr, np.repeat(np.arange(len(x)), np.diff(rowptr)), vals * x[colind])
    assert np.max(np.abs(y - y_csr)) <= 1e-12 * np.max(np.abs(y_csr))


@pytest.mark.parametrize("dim", [2, 3])
def test_morton_order_groups_aligned_cells(driver, tmp_path, dim):
    """The Morton order of a 2^k grid: a permutation in which every aligned
    run of 2^(dim*j) positions covers one aligned cell of side 2^j, on a
    grid whose block rows are numbered at random and whose coordinates are
    scaled and shifted (the quantisation spans the bounding box)."""
    n = 8
    g = np.stack(np.meshgrid(*([np.arange(n)] * dim), indexing="ij"), -1).reshape(-1, dim)
    rng = np.random.default_rng(3)
    perm = rng.permutation(len(g))
    pts = g[perm]  # block row i sits at grid point pts[i]
    ord = run_morton(driver, tmp_path, 2.5 * pts - 7.0)
    assert np.array_equal(np.sort(ord), np.arange(len(g)))
    for j in (1, 2):
        run = 2 ** (dim * j)
        for s in range(0, len(g), run):
            cell = pts[ord[s:s + run]] // 2 ** j
            assert (cell == cell[0]).all(), (j, s)


def test_morton_order_degenerate_inputs(driver, tmp_path):
    """A flat axis (every z equal) orders like the 2D grid of the other two;
    coincident points keep block-row order."""
    n = 4
    g2 = np.stack(np.meshgrid(np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 2)
    o2 = run_morton(driver, tmp_path, g2.astype(float))
    flat = np.concatenate([g2, np.full((len(g2), 1), 3.0)], axis=1)
    o3 = run_morton(driver, tmp_path, flat)
    assert np.array_equal(o2, o3)
    same = np.zeros((6, 3))
    assert np.array_equal(run_morton(driver, tmp_path, same), np.arange(6))


def test_layout_refuses_an_order_that_is_not_a_permutation(driver, tmp_path):
    M, lists = 2, [[0, 1], [0, 1, 2], [1, 2]]
    rowptr, nbc = [0], []
    for lst in lists:
        for _ in range(M):
            rowptr.append(rowptr[-1] + M * len(lst))
        nbc.extend(lst)
    rowptr, nbc = np.array(rowptr), np.array(nbc)
    assert run_layout(driver, tmp_path, M, rowptr, nbc, [2, 0, 1]) is not None
    assert run_layout(driver, tmp_path, M, rowptr, nbc, [2, 0, 2]) is None


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
