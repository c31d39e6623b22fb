"""Batched P1 Laplace assembly + residual of many meshes in one persistent
launch (HIP graph replays) against the oracle and against per-mesh loops."""
import numpy as np
import pytest

from helpers import TriProblem, oracle_p1, relerr
from pyop2_amd import elements, op2
from pyop2_amd.batch import BatchedLaplace
from pyop2_amd.configuration import configuration
from pyop2_amd.meshes import batched_tri_meshes

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _pairs(meshes):
    out = []
    for m in meshes:
        p = TriProblem(m)
        A = op2.Mat(p.sparsity(), np.float64)
        R = op2.Dat(p.verts, name="R")
        pm = op2.ParLoop(elements.p1_laplace_matrix(), p.cells, A(op2.INC, (p.c2v, p.c2v)), p.X(op2.READ, p.c2v))
        pr = op2.ParLoop(elements.p1_laplace_residual(), p.cells, R(op2.INC, p.c2v), p.X(op2.READ, p.c2v),
                         p.u(op2.READ, p.c2v))
        out.append((p, A, R, pm, pr))
    return out


@pytest.fixture(scope="module")
def meshes():
    return batched_tri_meshes(count=12, nmin=8, nmax=40, rng=np.random.default_rng([2026, 5, 1]))


@pytest.mark.parametrize("graph", [True, False])
def test_batched_matches_oracle_on_every_replay(device, meshes, graph):
    items = _pairs(meshes)
    refs = [(oracle_p1("p1_laplace_matrix", m), oracle_p1("p1_laplace_residual", m)) for m in meshes]
    batch = BatchedLaplace([(pm, pr) for _, _, _, pm, pr in items], graph=graph)
    for step in range(1, 4):
        batch()
        for (p, A, R, _, _), (rA, rR) in zip(items, refs):
            rp, ci, v = A.csr()
            assert np.array_equal(rp, rA.rowptr) and np.array_equal(ci, rA.colidx)
            assert relerr(v, step * rA.values) < TOL
            assert relerr(R.data_ro, step * rR) < TOL
        # eager work on the same stream between replays
        items[0][2].data[:] += 0.0


def test_batched_equals_plain_loops(device, meshes):
    configuration["strategy"] = "rowgather_slots"
    plain = _pairs(meshes)
    for _, _, _, pm, pr in plain:
        pm()
        pr()
    configuration["strategy"] = "auto"
    batched = _pairs(meshes)
    BatchedLaplace([(pm, pr) for _, _, _, pm, pr in batched])()
    for (_, A0, R0, _, _), (_, A1, R1, _, _) in zip(plain, batched):
        assert relerr(A1.csr()[2], A0.csr()[2]) < TOL
        assert relerr(R1.data_ro, R0.data_ro) < TOL


def test_batched_rejects_other_loops(device, meshes):
    p = TriProblem(meshes[0])
    M = op2.Dat(p.verts, name="M")
    pm = op2.ParLoop(elements.p1_lumped_mass(), p.cells, M(op2.INC, p.c2v), p.X(op2.READ, p.c2v))
    with pytest.raises(TypeError):
        BatchedLaplace([(pm, pm)])
