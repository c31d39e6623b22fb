"""Device against oracle on the paths the other GPU tests leave open.

* dof-level Mats: a vector Mat built without block sparsity (AIJ over the
  dofs, ``sparsity.pyx:168-173``) and a rectangular rdim=1 x cdim=2 Mat,
  pattern bit-exact and values within 1e-12;
* the fused P2 residual + BSR(3) assembly on the 32^3 cube (196,608 cells),
  the whole Mat and the whole residual, float64 within 1e-12 and float32
  within 1e-5 of the float64 oracle, and the split loop with its residual
  overlapped on a side stream, two passes in a row;
* two successive temporary Subsets of one set with indirect INC and WRITE
  arguments (plans must follow the Subset's contents);
* the arity-3 row-list assembly on a mesh with a node of valence > 31 (the
  unpacked row-list words);
* the reference's Mat known answers for ``zero_rows``,
  ``set_local_diagonal_entries``, ``addto_values`` / ``set_values`` and the
  assembly states they leave (``test_matrices.py:583-601, 740-768,
  798-852``).
"""
import numpy as np
import pytest

import kat_cases
from helpers import TetProblem, oracle_p2, relerr
from oracle import seqwrap, sparsity as osp
from pyop2_amd import elements, op2
from pyop2_amd.configuration import configuration
from pyop2_amd.meshes import unit_cube_tet_p2, unit_square_tri, workload_rng

pytestmark = pytest.mark.gpu


def _tri_ops(m):
    verts = op2.Set(m.nverts, "verts")
    cells = op2.Set(m.ncells, "cells")
    c2v = op2.Map(cells, verts, 3, m.cell_vertex, "c2v")
    X = op2.Dat(verts ** 2, m.coords, name="X")
    return verts, cells, c2v, X


# ---------------------------------------------------------------------------
# dof-level Mats

@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
def test_vector_mat_without_block_sparsity(device, strategy):
    """block_sparsity off: the 2x2 vector mass matrix lives in a scalar CSR
    over the dofs; pattern = the block pattern unwound, values = oracle."""
    configuration["strategy"] = strategy
    configuration["block_sparsity"] = False
    m = unit_square_tri(24, rng=workload_rng(1), permute=True)
    verts, cells, c2v, X = _tri_ops(m)
    sp = op2.Sparsity((verts ** 2, verts ** 2), [(c2v, c2v, None)])
    A = op2.Mat(sp, np.float64)
    assert A.block_shape == (1, 1)
    op2.par_loop(op2.Kernel(kat_cases.VEC_MASS_KERNEL, "p1_vmass_matrix"), cells,
                 A(op2.INC, (c2v, c2v)), X(op2.READ, c2v))
    A.assemble()
    brp, bci = osp.pattern(m.nverts, m.nverts, [(m.cell_vertex, m.cell_vertex, m.ncells)])
    rp, ci = osp.expand_block_pattern(brp, bci, 2, 2)
    grp, gci, gv = A.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert np.array_equal(sp.nnz, osp.dof_nnz(brp, 2, 2, block_sparse=False))
    ref = seqwrap.CSRMat(rp, ci, np.zeros(len(ci)))
    seqwrap.run(kat_cases.VEC_MASS_KERNEL, "p1_vmass_matrix",
                [seqwrap.MatArg(ref, "INC", m.cell_vertex, m.cell_vertex, 2, 2),
                 seqwrap.DatArg(m.coords, "READ", m.cell_vertex)], [(0, m.ncells)])
    assert relerr(gv, ref.values) < 1e-12


def test_vector_mat_without_block_sparsity_known_answer(device):
    """test_matrices.py:468-486 with block_sparsity off: kron(M, I)."""
    configuration["block_sparsity"] = False
    kat_cases.case_vector_mass_matrix(op2)


RECT_KERNEL = r"""
static void rect(double A[3][6], const double X[3][2])
{
  const double det = (X[1][0] - X[0][0]) * (X[2][1] - X[0][1])
                   - (X[2][0] - X[0][0]) * (X[1][1] - X[0][1]);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int c = 0; c < 2; ++c)
        A[i][2 * j + c] += (1.0 + i + 3.0 * c) * X[j][c] * det;
}
"""


@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
def test_rectangular_dof_mat(device, strategy):
    """rdim = 1 rows by cdim = 2 columns on the same vertex set: no diagonal
    (the DataSets differ), dof-level pattern, values = oracle."""
    configuration["strategy"] = strategy
    m = unit_square_tri(20, rng=workload_rng(1), permute=True)
    verts, cells, c2v, X = _tri_ops(m)
    sp = op2.Sparsity((verts ** 1, verts ** 2), [(c2v, c2v, None)])
    A = op2.Mat(sp, np.float64)
    op2.par_loop(op2.Kernel(RECT_KERNEL, "rect"), cells, A(op2.INC, (c2v, c2v)), X(op2.READ, c2v))
    A.assemble()
    brp, bci = osp.pattern(m.nverts, m.nverts, [(m.cell_vertex, m.cell_vertex, m.ncells)], set_diag=False)
    rp, ci = osp.expand_block_pattern(brp, bci, 1, 2)
    grp, gci, gv = A.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    ref = seqwrap.CSRMat(rp, ci, np.zeros(len(ci)))
    seqwrap.run(RECT_KERNEL, "rect", [seqwrap.MatArg(ref, "INC", m.cell_vertex, m.cell_vertex, 1, 2),
                                      seqwrap.DatArg(m.coords, "READ", m.cell_vertex)], [(0, m.ncells)])
    assert relerr(gv, ref.values) < 1e-12


# ---------------------------------------------------------------------------
# the fused P2 loop at 32^3, whole Mat and residual

@pytest.fixture(scope="module")
def tet32():
    return unit_cube_tet_p2(32, rng=workload_rng(4))


@pytest.fixture(scope="module")
def tet32_oracle(tet32):
    print(f"[tet32] oracle residual + BSR(3) Mat, {tet32.ncells} cells", flush=True)
    return (oracle_p2("p2_elasticity_residual", tet32), oracle_p2("p2_elasticity_matrix", tet32))


@pytest.mark.parametrize("strategy", ["auto", "split"])
def test_p2_fused_32cube_f64(device, tet32, tet32_oracle, strategy):
    configuration["strategy"] = strategy
    Rref, Aref = tet32_oracle
    p = TetProblem(tet32)
    A = op2.Mat(p.sparsity(), np.float64)
    R = op2.Dat(p.nodes ** 3, name="R")
    op2.par_loop(elements.p2_elasticity_fused(), p.cells, A(op2.INC, (p.c2n, p.c2n)), R(op2.INC, p.c2n),
                 p.X(op2.READ, p.c2v), p.u(op2.READ, p.c2n))
    A.assemble()
    rp, ci, v = A._host_csr()
    assert np.array_equal(rp, Aref.rowptr) and np.array_equal(ci, Aref.colidx)
    assert len(ci) == 7678721
    assert relerr(v, Aref.values) < 1e-12
    assert relerr(R.data_ro, Rref) < 1e-12


SPLIT_PARTS = {
    "defaults": None,
    "tiles384_atomics": [{"strategy": "rowgather_slots", "batch": [384, 480], "block": 192},
                         {"strategy": "staged", "tile": 384, "lds": 40960, "block": 192, "partials": False}],
    "tiles384_coloured": [{"strategy": "rowgather_slots", "batch": [384, 480], "block": 192},
                          {"strategy": "staged_coloured", "tile": 384, "lds": 40960, "block": 192}],
}


@pytest.mark.parametrize("overlap,parts", [("before", "defaults"), ("after", "defaults"),
                                           ("after", "tiles384_atomics"),
                                           ("after", "tiles384_coloured")])
def test_p2_fused_32cube_split_overlap(device, tet32, tet32_oracle, monkeypatch, overlap, parts):
    """The split loop with its residual overlapped on a side stream: two
    passes in a row (the second adds onto the first) equal twice the
    oracle, and the calling stream sees both parts' results."""
    from pyop2_amd import autotune
    configuration["strategy"] = "auto"
    cfg = {"strategy": "split", "overlap": overlap, "parts": SPLIT_PARTS[parts]}
    monkeypatch.setattr(autotune, "choice", lambda ex, pl: cfg)
    Rref, Aref = tet32_oracle
    p = TetProblem(tet32)
    A = op2.Mat(p.sparsity(), np.float64)
    R = op2.Dat(p.nodes ** 3, name="R")
    for _ in range(2):
        op2.par_loop(elements.p2_elasticity_fused(), p.cells, A(op2.INC, (p.c2n, p.c2n)), R(op2.INC, p.c2n),
                     p.X(op2.READ, p.c2v), p.u(op2.READ, p.c2n))
    A.assemble()
    _, _, v = A._host_csr()
    assert relerr(v, 2 * Aref.values) < 1e-12
    assert relerr(R.data_ro, 2 * Rref) < 1e-12


@pytest.mark.parametrize("overlap,residual", [(None, "staged"), ("after", "staged"),
                                              ("after", "staged_coloured")])
def test_p2_fused_split_captured_in_a_hip_graph(device, monkeypatch, overlap, residual):
    """The split P2 loop (and its side-stream overlap) captured in a HIP graph
    after one eager pass: every replay adds one more pass, checked against the
    oracle after each replay, with eager work and synchronisation between
    replays (device arrays read directly, past the Dats' host copies)."""
    import torch
    from pyop2_amd import autotune
    m = unit_cube_tet_p2(8, rng=workload_rng(3))
    Rref = oracle_p2("p2_elasticity_residual", m)
    Aref = oracle_p2("p2_elasticity_matrix", m)
    configuration["strategy"] = "auto"
    cfg = {"strategy": "split", "overlap": overlap,
           "parts": [{"strategy": "rowgather_slots"},
                     {"strategy": residual, "tile": 256, "lds": 65536, "block": 256}]}
    monkeypatch.setattr(autotune, "choice", lambda ex, pl: cfg)
    p = TetProblem(m)
    A = op2.Mat(p.sparsity(), np.float64)
    R = op2.Dat(p.nodes ** 3, name="R")
    pl = op2.ParLoop(elements.p2_elasticity_fused(), p.cells, A(op2.INC, (p.c2n, p.c2n)), R(op2.INC, p.c2n),
                     p.X(op2.READ, p.c2v), p.u(op2.READ, p.c2n))
    pl()
    torch.cuda.synchronize()
    capture = torch.cuda.Stream()
    capture.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(capture):
        with torch.cuda.graph(g, stream=capture):
            pl()
    torch.cuda.current_stream().wait_stream(capture)
    torch.cuda.synchronize()
    rows, cols, _ = A._host_csr()
    assert np.array_equal(rows, Aref.rowptr) and np.array_equal(cols, Aref.colidx)
    for k in range(3):
        g.replay()
        torch.cuda.synchronize()
        scratch = torch.ones(1024, dtype=torch.float64, device=R._device_array().device).sum()
        torch.cuda.synchronize()
        assert float(scratch) == 1024.0
        got_r = R._device_array().cpu().numpy()
        _, _, got_a = A._host_csr()
        assert relerr(got_r, (k + 2) * Rref) < 1e-12, k
        assert relerr(got_a, (k + 2) * Aref.values) < 1e-12, k


def test_p2_fused_32cube_f32(device, tet32):
    Rref = oracle_p2("p2_elasticity_residual", tet32, coords=tet32.coords.astype(np.float32).astype(np.float64),
                     u=tet32.u.astype(np.float32).astype(np.float64))
    Aref = oracle_p2("p2_elasticity_matrix", tet32, coords=tet32.coords.astype(np.float32).astype(np.float64))
    p = TetProblem(tet32, dtype=np.float32)
    A = op2.Mat(p.sparsity(), np.float32)
    R = op2.Dat(p.nodes ** 3, dtype=np.float32, name="R")
    op2.par_loop(elements.p2_elasticity_fused(), p.cells, A(op2.INC, (p.c2n, p.c2n)), R(op2.INC, p.c2n),
                 p.X(op2.READ, p.c2v), p.u(op2.READ, p.c2n))
    A.assemble()
    _, _, v = A._host_csr()
    assert relerr(v, Aref.values) < 1e-5
    assert relerr(R.data_ro, Rref) < 1e-5


# ---------------------------------------------------------------------------
# successive temporary Subsets

SCATTER = "static void sc(double *r, const double *w) { for (int k = 0; k < 3; ++k) r[k] += w[0]; }"
STAMP = "static void st(double *r, const double *w) { for (int k = 0; k < 3; ++k) r[k] = 7.0 + 0.0 * w[0]; }"


@pytest.mark.parametrize("mode", ["INC", "WRITE"])
def test_successive_temporary_subsets(device, mode):
    """Loops over two temporary Subsets of equal size, one after the other:
    each result equals the oracle on that Subset's elements (a plan kept for
    the first Subset must not be reused for the second)."""
    configuration["strategy"] = "coloured"
    m = unit_square_tri(16, rng=workload_rng(1), permute=True)
    verts, cells, c2v, _ = _tri_ops(m)
    w = op2.Dat(cells, np.arange(1, m.ncells + 1, dtype=np.float64), np.float64)
    code, name = (SCATTER, "sc") if mode == "INC" else (STAMP, "st")
    for first in (0, 1):
        idx = np.arange(first, m.ncells, 2)[:200]
        r = op2.Dat(verts, np.zeros(m.nverts), np.float64)
        op2.par_loop(op2.Kernel(code, name), op2.Subset(cells, idx), r(getattr(op2, mode), c2v),
                     w(op2.READ))
        ref = np.zeros(m.nverts)
        seqwrap.run(code, name, [seqwrap.DatArg(ref, mode, m.cell_vertex),
                                 seqwrap.DatArg(np.arange(1, m.ncells + 1, dtype=np.float64), "READ")],
                    [(0, len(idx))], subset=idx.astype(np.int32))
        # INC sums; WRITE marks exactly the vertices the Subset's cells touch
        assert np.array_equal(r.data_ro, ref) if mode == "WRITE" else relerr(r.data_ro, ref) < 1e-14


# ---------------------------------------------------------------------------
# arity-3 row-list assembly with a high-valence node

def _fan_mesh(nfan=40, n=12, seed=5):
    """A triangulated square plus a fan of ``nfan`` triangles around one
    extra vertex joined to a ring of new vertices (valence nfan > 31)."""
    base = unit_square_tri(n, rng=workload_rng(1), permute=True)
    nv = base.nverts
    centre = nv
    ring = nv + 1 + np.arange(nfan)
    ang = np.linspace(0, 2 * np.pi, nfan, endpoint=False)
    coords = np.vstack([base.coords, [[2.0, 2.0]], np.c_[2.0 + np.cos(ang), 2.0 + np.sin(ang)]])
    fan = np.c_[np.full(nfan, centre), ring, np.roll(ring, -1)].astype(np.int32)
    cells = np.vstack([base.cell_vertex, fan]).astype(np.int32)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(cells))
    return cells[perm], coords, len(coords)


def test_rowlist_assembly_with_high_valence_node(device):
    configuration["strategy"] = "rowgather_rows"
    cv, coords, nverts = _fan_mesh()
    verts, cells = op2.Set(nverts), op2.Set(len(cv))
    c2v = op2.Map(cells, verts, 3, cv)
    X = op2.Dat(verts ** 2, coords)
    A = op2.Mat(op2.Sparsity((verts ** 1, verts ** 1), [(c2v, c2v, None)]), np.float64)
    pl = op2.ParLoop(elements.p1_laplace_matrix(), cells, A(op2.INC, (c2v, c2v)), X(op2.READ, c2v))
    pl()
    A.assemble()
    assert pl.global_kernel.compile().strategy_used == "rowgather_rows"
    rp, ci = osp.pattern(nverts, nverts, [(cv, cv, len(cv))])
    assert np.diff(rp).max() > 31
    ref = seqwrap.CSRMat(rp, ci, np.zeros(len(ci)))
    seqwrap.run(elements.kernel_source("p1_laplace_matrix"), "p1_laplace_matrix",
                [seqwrap.MatArg(ref, "INC", cv, cv), seqwrap.DatArg(coords, "READ", cv)], [(0, len(cv))])
    grp, gci, gv = A.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)
    assert relerr(gv, ref.values) < 1e-12


# ---------------------------------------------------------------------------
# Mat known answers

def _mass_mat():
    nodes, elements_, en, coords = kat_cases._two_triangles(op2)
    A = op2.Mat(op2.Sparsity((nodes ** 1, nodes ** 1), [(en, en, None)]), np.float64)
    op2.par_loop(op2.Kernel(kat_cases.MASS_KERNEL, "p1_mass_matrix"), elements_, A(op2.INC, (en, en)),
                 coords(op2.READ, en))
    A.assemble()
    return nodes, A


def test_zero_rows(device):
    nodes, A = _mass_mat()
    want = kat_cases.EXPECTED_MASS.copy()
    want[0] = [12.0, 0.0, 0.0, 0.0]
    A.zero_rows([0], 12.0)
    assert np.allclose(A.values, want, rtol=1e-5)


def test_zero_rows_subset_then_last_row(device):
    nodes, A = _mass_mat()
    want = kat_cases.EXPECTED_MASS.copy()
    want[0] = [12.0, 0.0, 0.0, 0.0]
    A.zero_rows(op2.Subset(nodes, [0]), 12.0)
    assert np.allclose(A.values, want, rtol=1e-5)
    want[3] = [0.0, 0.0, 0.0, 4.0]
    A.zero_rows([3], 4.0)
    assert np.allclose(A.values, want, rtol=1e-5)


@pytest.mark.parametrize("n", [1, 2])
def test_set_local_diagonal_entries(device, n):
    nodes, _, en, _ = kat_cases._two_triangles(op2)
    A = op2.Mat(op2.Sparsity((nodes ** n, nodes ** n), [(en, en, None)]), np.float64)
    A.set_local_diagonal_entries(list(range(A.nblock_rows)))
    assert A.assembly_state is op2.Mat.INSERT_VALUES
    A.assemble()
    assert (A.values == np.identity(4 * n)).all()
    A.set_local_diagonal_entries(list(range(A.nblock_rows)))
    A.assemble()
    assert (A.values == np.identity(4 * n)).all()


def _mixed_nested():
    elem, node = op2.Set(3), op2.Set(4)
    mset = op2.MixedSet((elem, node))
    mmap = op2.MixedMap((op2.Map(elem, elem, 1, [0, 1, 2]), op2.Map(elem, node, 2, [0, 1, 1, 2, 2, 3])))
    sp = op2.Sparsity((mset ** 1, mset ** 1),
                      {(i, j): [(rm, cm, None)] for i, rm in enumerate(mmap) for j, cm in enumerate(mmap)})
    return op2.Mat(sp)


def test_mixed_mat_insert_and_add(device):
    """test_matrices.py:824-838 (nested)."""
    mat = _mixed_nested()
    assert mat.assembly_state is op2.Mat.ASSEMBLED
    mat[0, 0].addto_values(0, 0, [1])
    assert mat[0, 0].assembly_state is op2.Mat.ADD_VALUES
    assert mat[1, 1].assembly_state is op2.Mat.ASSEMBLED
    mat[1, 1].addto_values(1, 1, [3])
    mat[1, 1].set_values(0, 0, [2])
    mat[0, 0].set_values(1, 1, [4])
    mat[1, 1].addto_values(0, 0, [3])
    mat.assemble()
    assert np.allclose(mat[0, 0].values, np.diag([1, 4, 0]))
    assert np.allclose(mat[1, 1].values, np.diag([5, 3, 0, 0]))
    assert np.allclose(mat[0, 1].values, 0)
    assert np.allclose(mat[1, 0].values, 0)


def test_flush_between_insert_and_add(device):
    """test_matrices.py:840-862 (nested): switching ADD -> INSERT flushes once
    (the reference inserts at (1, 0) with new-nonzero errors switched off;
    the device pattern is fixed, so the insertion goes to (1, 1))."""
    import types
    mat = _mixed_nested()
    count = [0]
    for blk in mat:
        old = blk._flush_assembly

        def flush(self, old=old):
            old()
            count[0] += 1
        blk._flush_assembly = types.MethodType(flush, blk)
    mat[0, 0].addto_values(0, 0, [1])
    assert count[0] == 0
    mat[0, 0].set_values(1, 1, [2])
    assert count[0] == 1
    mat.assemble()
    assert count[0] == 1
