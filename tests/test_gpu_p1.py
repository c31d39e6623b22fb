"""P1 triangle loops on the device against the oracle (FP64, 1e-12 normwise).

Workloads: the 256^2 jittered square (131,072 cells) for the lumped mass and
the Laplace residual, the permuted 1024^2 square for the CSR assembly at
reduced size (parity) and at full size (pattern count and row-sum property).
"""
import numpy as np
import pytest

from helpers import TriProblem, oracle_p1, relerr, use_strategy
from oracle import sparsity as osp
from pyop2_amd import elements, op2
from pyop2_amd.configuration import configuration
from pyop2_amd.meshes import unit_square_tri, workload_rng

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def mesh256():
    return unit_square_tri(256, rng=workload_rng(1))


@pytest.mark.parametrize("strategy", ["atomic", "coloured", "staged", "staged_atomic", "staged_coloured",
                                      "auto"])
@pytest.mark.parametrize("split", ["all", "nocore", "half"])
def test_p1_residual_and_mass(device, mesh256, strategy, split):
    n = mesh256.ncells
    core = {"all": None, "nocore": 0, "half": n // 2}[split]
    use_strategy(strategy)
    p = TriProblem(mesh256, core=core)
    R = op2.Dat(p.verts, name="R")
    M = op2.Dat(p.verts, name="M")
    op2.par_loop(elements.p1_laplace_residual(), p.cells, R(op2.INC, p.c2v), p.X(op2.READ, p.c2v),
                 p.u(op2.READ, p.c2v))
    op2.par_loop(elements.p1_lumped_mass(), p.cells, M(op2.INC, p.c2v), p.X(op2.READ, p.c2v))
    assert relerr(R.data_ro, oracle_p1("p1_laplace_residual", mesh256)) < TOL
    assert relerr(M.data_ro, oracle_p1("p1_lumped_mass", mesh256)) < TOL
    assert abs(M.data_ro.sum() - 1.0) < 1e-12


@pytest.mark.parametrize("strategy", ["atomic", "staged", "staged_coloured"])
def test_p1_inc_accumulates(device, mesh256, strategy):
    """INC adds into existing values (two loops = twice the residual)."""
    configuration["strategy"] = strategy
    p = TriProblem(mesh256)
    R = op2.Dat(p.verts, name="R")
    for _ in range(2):
        op2.par_loop(elements.p1_laplace_residual(), p.cells, R(op2.INC, p.c2v), p.X(op2.READ, p.c2v),
                     p.u(op2.READ, p.c2v))
    assert relerr(R.data_ro, 2 * oracle_p1("p1_laplace_residual", mesh256)) < TOL


@pytest.mark.parametrize("n,permute", [(16, False), (64, True), (200, True)])
def test_p1_sparsity_bit_exact(device, n, permute):
    m = unit_square_tri(n, rng=workload_rng(2), permute=permute)
    p = TriProblem(m)
    sp = p.sparsity()
    rp, ci = osp.pattern(m.nverts, m.nverts, [(m.cell_vertex, m.cell_vertex, m.ncells)])
    from pyop2_amd import device as dev
    grp, gci = sp.block_pattern
    assert np.array_equal(dev.to_host(grp), rp)
    assert np.array_equal(dev.to_host(gci), ci)
    assert np.array_equal(sp.nnz, np.diff(rp))


@pytest.mark.parametrize("strategy", ["rowgather_rows", "rowgather_cells", "rowgather_staged", "rowgather_slots", "rowgather_thread", "rowgather", "atomic", "coloured"])
def test_p1_csr_assembly(device, strategy):
    m = unit_square_tri(128, rng=workload_rng(2), permute=True)
    configuration["strategy"] = strategy
    p = TriProblem(m)
    A = op2.Mat(p.sparsity(), np.float64)
    op2.par_loop(elements.p1_laplace_matrix(), p.cells, A(op2.INC, (p.c2v, p.c2v)), p.X(op2.READ, p.c2v))
    A.assemble()
    ref = oracle_p1("p1_laplace_matrix", m)
    rp, ci, v = A.csr()
    assert np.array_equal(rp, ref.rowptr) and np.array_equal(ci, ref.colidx)
    assert relerr(v, ref.values) < TOL


def test_p1_csr_full_size_properties(device):
    """1024^2 permuted: nnz = V + 2E = 7,346,177; K 1 = 0 (rows sum to 0)."""
    m = unit_square_tri(1024, rng=workload_rng(2), permute=True)
    p = TriProblem(m)
    sp = p.sparsity()
    assert sp.block_nnz == 7346177
    A = op2.Mat(sp, np.float64)
    op2.par_loop(elements.p1_laplace_matrix(), p.cells, A(op2.INC, (p.c2v, p.c2v)), p.X(op2.READ, p.c2v))
    A.assemble()
    torch = __import__("torch")
    vals = A.device_values
    rowptr = A.rowptr.long()
    rows = torch.repeat_interleave(torch.arange(m.nverts, device=vals.device), rowptr[1:] - rowptr[:-1])
    rowsum = torch.zeros(m.nverts, dtype=vals.dtype, device=vals.device).index_add_(0, rows, vals)
    diag_scale = vals.abs().max().item()
    assert rowsum.abs().max().item() < 1e-12 * diag_scale


def test_user_written_copy_runs_the_hand_written_loop(device, mesh256):
    """A reformatted and renamed copy of the P1 residual kernel, built with
    op2.Kernel directly, runs the hand-written entry point and gives the
    oracle's residual."""
    import re
    src = elements.kernel_source("p1_laplace_residual").replace("p1_laplace_residual", "laplace_res")
    src = "/* user copy */\n" + re.sub(r"[ \t]+", "  ", src.replace(";", " ;\n"))
    use_strategy("auto")
    p = TriProblem(mesh256)
    R = op2.Dat(p.verts, name="R")
    pl = op2.ParLoop(op2.Kernel(src, "laplace_res"), p.cells, R(op2.INC, p.c2v), p.X(op2.READ, p.c2v),
                     p.u(op2.READ, p.c2v))
    pl()
    assert pl.global_kernel.is_hand_written
    assert relerr(R.data_ro, oracle_p1("p1_laplace_residual", mesh256)) < TOL
