"""Pins of the sparsity closed form and the greedy colouring (CPU).

* ``oracle.sparsity.pattern`` against a brute-force set construction
  (dict of Python sets filled element by element, as the preallocator is
  filled in ``pyop2/sparsity.pyx:263-284``) on random maps with negative
  entries, several map pairs, rectangular shapes and with/without diagonal;
* the counts of the workload meshes: nnz = V + 2E for triangles (460,289 at
  256^2), and the P2 tetrahedral block pattern of the 8^3 cube against a
  scipy product of incidence matrices;
* the greedy colouring: validity (no two elements of a colour share a
  target), first-fit minimality of each colour, and the 32-colour pass
  structure on a clique larger than 32.
"""
import numpy as np
import pytest
import scipy.sparse as ss

from oracle import sparsity as osp
from pyop2_amd.meshes import unit_cube_tet_p2, unit_square_tri, workload_rng


def brute(nrows, ncols, pairs, diag):
    rows = [set() for _ in range(nrows)]
    for rmap, cmap, size in pairs:
        for e in range(size):
            for r in rmap[e]:
                for c in cmap[e]:
                    if r >= 0 and c >= 0:
                        rows[r].add(int(c))
    if diag:
        for i in range(min(nrows, ncols)):
            rows[i].add(i)
    rp = np.zeros(nrows + 1, dtype=np.int64)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.array([c for r in rows for c in sorted(r)], dtype=np.int64)
    return rp, ci


@pytest.mark.parametrize("seed", range(4))
def test_pattern_vs_brute_force(seed):
    rng = np.random.default_rng([2026, 99, seed])
    nrows, ncols = rng.integers(5, 60, size=2)
    pairs = []
    for _ in range(rng.integers(1, 4)):
        ne = rng.integers(1, 40)
        ra, ca = rng.integers(1, 5, size=2)
        rmap = rng.integers(-1, nrows, size=(ne, ra)).astype(np.int32)
        cmap = rng.integers(-1, ncols, size=(ne, ca)).astype(np.int32)
        pairs.append((rmap, cmap, int(rng.integers(0, ne + 1))))
    for diag in (True, False):
        rp, ci = osp.pattern(int(nrows), int(ncols), pairs, set_diag=diag)
        brp, bci = brute(int(nrows), int(ncols), pairs, diag)
        assert np.array_equal(rp, brp) and np.array_equal(ci, bci)


def test_triangle_counts():
    m = unit_square_tri(256, rng=workload_rng(1))
    rp, ci = osp.pattern(m.nverts, m.nverts, [(m.cell_vertex, m.cell_vertex, m.ncells)])
    nedges = 3 * 256 * 256 + 2 * 256
    assert len(ci) == m.nverts + 2 * nedges == 460289
    assert np.diff(rp).min() == 3 and np.diff(rp).max() == 7


def test_tet_block_pattern_vs_incidence_product():
    m = unit_cube_tet_p2(8, rng=workload_rng(3))
    rp, ci = osp.pattern(m.nnodes, m.nnodes, [(m.cell_node, m.cell_node, m.ncells)])
    C = ss.csr_matrix((np.ones(m.cell_node.size), m.cell_node.ravel(),
                       np.arange(0, m.cell_node.size + 1, 10)), shape=(m.ncells, m.nnodes))
    P = (C.T @ C).tocsr()
    P.sort_indices()
    assert np.array_equal(rp, P.indptr) and np.array_equal(ci, P.indices)
    assert np.diff(rp).min() == 10 and np.diff(rp).max() == 65


def test_dof_nnz_unwinding():
    rp = np.array([0, 2, 5], dtype=np.int32)
    assert list(osp.dof_nnz(rp, 3, 3, True)) == [2, 3]
    assert list(osp.dof_nnz(rp, 3, 3, False)) == [6, 6, 6, 9, 9, 9]
    assert list(osp.dof_nnz(rp, 2, 3, True)) == [6, 6, 9, 9]


def _check_colouring(maps, colours, nelem):
    for c in np.unique(colours):
        idx = np.nonzero(colours == c)[0]
        for m in maps:
            t = m[idx].ravel()
            t = t[t >= 0]
            assert len(np.unique(t)) == len(t), f"colour {c} has a conflict"


def test_greedy_colouring_valid_and_first_fit():
    m = unit_square_tri(32, rng=workload_rng(1), permute=True)
    colours, nc = osp.greedy_colour(m.ncells, [m.cell_vertex], [0], [m.nverts])
    _check_colouring([m.cell_vertex], colours, m.ncells)
    # first fit: every element of colour c > 0 conflicts with an earlier element of each lower colour
    touching = {}
    for e in range(m.ncells):
        for v in m.cell_vertex[e]:
            touching.setdefault(int(v), []).append(e)
    for e in range(m.ncells):
        c = colours[e]
        earlier = {colours[f] for v in m.cell_vertex[e] for f in touching[int(v)] if f < e}
        assert all(k in earlier for k in range(c)), (e, c)


def test_greedy_colouring_passes_of_32():
    # 40 elements all touching node 0: a 40-clique needs 40 colours (two passes)
    maps = [np.zeros((40, 1), dtype=np.int32)]
    colours, nc = osp.greedy_colour(40, maps, [0], [1])
    assert nc == 40 and sorted(colours) == list(range(40))


def _random_block_pattern(rng, diag):
    nrows, ncols = (int(v) for v in rng.integers(1, 40, size=2))
    ne = int(rng.integers(0, 30))
    ra, ca = (int(v) for v in rng.integers(1, 4, size=2))
    rmap = rng.integers(-1, nrows, size=(ne, ra)).astype(np.int32)
    cmap = rng.integers(-1, ncols, size=(ne, ca)).astype(np.int32)
    rp, ci = osp.pattern(nrows, ncols, [(rmap, cmap, ne)], set_diag=diag)
    return nrows, ncols, rp, ci


@pytest.mark.parametrize("seed", range(6))
@pytest.mark.parametrize("rdim,cdim", [(1, 2), (2, 1), (2, 3), (3, 3), (1, 1)])
def test_expand_block_pattern_is_the_kronecker_pattern(seed, rdim, cdim):
    """The dof pattern of rdim x cdim blocks (sparsity.pyx:168-173: every
    block entry of every block) is the pattern of kron(B, ones(rdim, cdim))
    computed by scipy, on random block patterns with empty rows; its row
    lengths are dof_nnz's unwinding."""
    rng = np.random.default_rng([2026, 98, seed])
    nrows, ncols, rp, ci = _random_block_pattern(rng, diag=bool(seed % 2))
    B = ss.csr_matrix((np.ones(len(ci)), ci, rp), shape=(nrows, ncols))
    D = ss.kron(B, np.ones((rdim, cdim))).tocsr()
    D.sort_indices()
    erp, eci = osp.expand_block_pattern(rp, ci, rdim, cdim)
    assert np.array_equal(erp, D.indptr) and np.array_equal(eci, D.indices)
    assert np.array_equal(np.diff(erp), osp.dof_nnz(rp, rdim, cdim, block_sparse=False))


def test_expand_block_pattern_catches_transposed_block_shape():
    """Planted mistake: the block shape given as (cdim, rdim) yields another
    pattern than the Kronecker one."""
    rng = np.random.default_rng([2026, 97])
    nrows, ncols, rp, ci = _random_block_pattern(rng, diag=True)
    B = ss.csr_matrix((np.ones(len(ci)), ci, rp), shape=(nrows, ncols))
    D = ss.kron(B, np.ones((1, 2))).tocsr()
    D.sort_indices()
    erp, eci = osp.expand_block_pattern(rp, ci, 2, 1)
    assert not (np.array_equal(erp, D.indptr) and np.array_equal(eci, D.indices))


def test_expand_block_pattern_reference_counts():
    """test_matrices.py:555-566: nnz [1, 2, 1, 2] at block level for the
    diagonal-augmented pattern, and its dof unwinding for 2x2 blocks."""
    m = np.array([[1, 3]], dtype=np.int32)
    rp, ci = osp.pattern(4, 4, [(m, m, 1)], set_diag=True)
    assert list(np.diff(rp)) == [1, 2, 1, 2]
    erp, _ = osp.expand_block_pattern(rp, ci, 2, 2)
    assert list(np.diff(erp)) == [2, 2, 4, 4, 2, 2, 4, 4]


@pytest.mark.parametrize("seed", range(4))
def test_monolithic_pattern_is_the_block_matrix_pattern(seed):
    """The monolithic dof pattern equals scipy's pattern of bmat of the
    Kronecker-expanded blocks (empty blocks included)."""
    rng = np.random.default_rng([2026, 96, seed])
    nr, nc = int(rng.integers(1, 4)), int(rng.integers(1, 4))
    rsz = [int(v) for v in rng.integers(1, 15, size=nr)]
    csz = [int(v) for v in rng.integers(1, 15, size=nc)]
    rd = [int(v) for v in rng.integers(1, 4, size=nr)]
    cd = [int(v) for v in rng.integers(1, 4, size=nc)]
    ne = 20
    blocks, dense = [], []
    for i in range(nr):
        brow, drow = [], []
        for j in range(nc):
            if rng.random() < 0.25:
                brow.append(None)
                drow.append(ss.csr_matrix((rsz[i] * rd[i], csz[j] * cd[j])))
                continue
            rmap = rng.integers(-1, rsz[i], size=(ne, 2)).astype(np.int32)
            cmap = rng.integers(-1, csz[j], size=(ne, 3)).astype(np.int32)
            rp, ci = osp.pattern(rsz[i], csz[j], [(rmap, cmap, ne)], set_diag=(i == j))
            brow.append((rp, ci))
            B = ss.csr_matrix((np.ones(len(ci)), ci, rp), shape=(rsz[i], csz[j]))
            drow.append(ss.kron(B, np.ones((rd[i], cd[j]))))
        blocks.append(brow)
        dense.append(drow)
    M = ss.bmat(dense).tocsr()
    M.eliminate_zeros()
    M.sort_indices()
    rp, ci = osp.monolithic_pattern(blocks, rd, cd, rsz, csz)
    assert np.array_equal(rp, M.indptr) and np.array_equal(ci, M.indices)
