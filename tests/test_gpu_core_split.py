"""Matrix-writing loops over a set with a core/owned split, on the device
against the oracle.

The reference runs a loop as two calls of the generated wrapper, one over
the core elements and one over the owned ones (``pyop2/parloop.py:243-260``);
every strategy must then add exactly the contributions of the elements of
each call's range.  The split points fall inside row batches and tiles
(a third or a half of the cells), and a loop with no core elements calls the
wrapper once.  FP64 within 1e-12 normwise, FP32 within 1e-5 of the FP64
oracle run on the float32-rounded inputs.
"""
import numpy as np
import pytest

from helpers import TetProblem, TriProblem, oracle_p1, oracle_p2, relerr
from pyop2_amd import elements, op2
from pyop2_amd.configuration import configuration
from pyop2_amd.meshes import unit_cube_tet_p2, unit_square_tri, workload_rng

pytestmark = pytest.mark.gpu

SPLITS = {"third": 3, "half": 2, "nocore": None}


def _core(n, split):
    return 0 if SPLITS[split] is None else n // SPLITS[split]


@pytest.fixture(scope="module")
def tri128():
    return unit_square_tri(128, rng=workload_rng(2), permute=True)


@pytest.fixture(scope="module")
def tet8():
    return unit_cube_tet_p2(8, rng=workload_rng(3))


@pytest.mark.parametrize("strategy", ["rowgather_rows", "rowgather_slots", "rowgather_cells", "atomic",
                                      "coloured"])
@pytest.mark.parametrize("split", ["third", "half", "nocore"])
def test_p1_csr_assembly_over_core_and_owned(device, tri128, strategy, split):
    configuration["strategy"] = strategy
    p = TriProblem(tri128, core=_core(tri128.ncells, split))
    A = op2.Mat(p.sparsity(), np.float64)
    op2.par_loop(elements.p1_laplace_matrix(), p.cells, A(op2.INC, (p.c2v, p.c2v)), p.X(op2.READ, p.c2v))
    A.assemble()
    ref = oracle_p1("p1_laplace_matrix", tri128)
    rp, ci, v = A.csr()
    assert np.array_equal(rp, ref.rowptr) and np.array_equal(ci, ref.colidx)
    assert relerr(v, ref.values) < 1e-12


@pytest.mark.parametrize("strategy", ["split", "rowgather_slots", "rowgather_rows", "atomic"])
@pytest.mark.parametrize("split", ["third", "half"])
def test_p2_fused_over_core_and_owned(device, tet8, strategy, split):
    configuration["strategy"] = strategy
    p = TetProblem(tet8, core=_core(tet8.ncells, split))
    A = op2.Mat(p.sparsity(), np.float64)
    R = op2.Dat(p.nodes ** 3, name="R")
    op2.par_loop(elements.p2_elasticity_fused(), p.cells, A(op2.INC, (p.c2n, p.c2n)), R(op2.INC, p.c2n),
                 p.X(op2.READ, p.c2v), p.u(op2.READ, p.c2n))
    A.assemble()
    _, _, v = A._host_csr()
    assert relerr(v, oracle_p2("p2_elasticity_matrix", tet8).values) < 1e-12
    assert relerr(R.data_ro, oracle_p2("p2_elasticity_residual", tet8)) < 1e-12


@pytest.mark.parametrize("overlap,residual", [(None, "staged"), ("before", "staged"),
                                              ("after", "staged_coloured")])
def test_p2_fused_split_overlap_over_core_and_owned(device, tet8, monkeypatch, overlap, residual):
    """The split loop with its residual on a side stream, called once per
    part: the owned call's residual must follow the core call's, and the
    calling stream must see both."""
    from pyop2_amd import autotune
    configuration["strategy"] = "auto"
    cfg = {"strategy": "split", "parts": [{"strategy": "rowgather_slots"},
                                          {"strategy": residual, "tile": 256, "lds": 65536, "block": 256}]}
    if overlap:
        cfg["overlap"] = overlap
    monkeypatch.setattr(autotune, "choice", lambda ex, pl: cfg)
    p = TetProblem(tet8, core=_core(tet8.ncells, "third"))
    A = op2.Mat(p.sparsity(), np.float64)
    R = op2.Dat(p.nodes ** 3, name="R")
    for _ in range(2):
        op2.par_loop(elements.p2_elasticity_fused(), p.cells, A(op2.INC, (p.c2n, p.c2n)), R(op2.INC, p.c2n),
                     p.X(op2.READ, p.c2v), p.u(op2.READ, p.c2n))
    A.assemble()
    _, _, v = A._host_csr()
    assert relerr(v, 2 * oracle_p2("p2_elasticity_matrix", tet8).values) < 1e-12
    assert relerr(R.data_ro, 2 * oracle_p2("p2_elasticity_residual", tet8)) < 1e-12


@pytest.mark.parametrize("strategy", ["split", "rowgather_slots"])
def test_p2_fused_f32_over_core_and_owned(device, tet8, strategy):
    configuration["strategy"] = strategy
    p = TetProblem(tet8, dtype=np.float32, core=_core(tet8.ncells, "third"))
    A = op2.Mat(p.sparsity(), np.float32)
    R = op2.Dat(p.nodes ** 3, dtype=np.float32, name="R")
    op2.par_loop(elements.p2_elasticity_fused(), p.cells, A(op2.INC, (p.c2n, p.c2n)), R(op2.INC, p.c2n),
                 p.X(op2.READ, p.c2v), p.u(op2.READ, p.c2n))
    A.assemble()
    _, _, v = A._host_csr()
    X32 = tet8.coords.astype(np.float32).astype(np.float64)
    u32 = tet8.u.astype(np.float32).astype(np.float64)
    assert relerr(v, oracle_p2("p2_elasticity_matrix", tet8, coords=X32).values) < 1e-5
    assert relerr(R.data_ro, oracle_p2("p2_elasticity_residual", tet8, coords=X32, u=u32)) < 1e-5
