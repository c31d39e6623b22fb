"""The device par_loop (generated and hand-written paths) against the
reference's known answers: every ``kat_cases`` case through
``pyop2_amd.op2`` on the GPU, under each conflict-resolution strategy."""
import pytest

import kat_cases
from pyop2_amd import op2
from pyop2_amd.configuration import configuration

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
@pytest.mark.parametrize("split", kat_cases.SPLITS, ids=["all", "nocore", "half"])
@pytest.mark.parametrize("case", kat_cases.LOOP_CASES, ids=lambda c: c.__name__)
def test_loop_cases(device, case, split, strategy):
    configuration["strategy"] = strategy
    case(op2, split)


@pytest.mark.parametrize("name,case", kat_cases.OTHER_CASES + kat_cases.MAP_CASES,
                         ids=[n for n, _ in kat_cases.OTHER_CASES + kat_cases.MAP_CASES])
def test_other_cases(device, name, case):
    case(op2)


@pytest.mark.parametrize("split", kat_cases.DIRECT_SPLITS, ids=["all", "nocore", "half", "halo"])
@pytest.mark.parametrize("case", kat_cases.DIRECT_CASES, ids=lambda c: c.__name__)
def test_direct_cases(device, case, split):
    case(op2, split)


@pytest.mark.parametrize("split", kat_cases.SPLITS, ids=["all", "nocore", "half"])
@pytest.mark.parametrize("case", kat_cases.REDUCTION_CASES, ids=lambda c: c.__name__)
def test_reduction_cases(device, case, split):
    case(op2, split)


@pytest.mark.parametrize("name,dtype,op,start,fill,want", kat_cases.TYPED_REDUCTIONS,
                         ids=[c[0] for c in kat_cases.TYPED_REDUCTIONS])
def test_typed_reductions(device, name, dtype, op, start, fill, want):
    kat_cases.case_reduce_typed(op2, kat_cases.SPLITS[2], dtype, op, start, fill, want)


@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
def test_permuted_composed_under_strategies(device, strategy):
    """INC through permuted and composed maps under both conflict strategies:
    the device result equals the table-level restatement."""
    import numpy as np
    rng = np.random.default_rng(7)
    configuration["strategy"] = strategy
    cells, mid, nodes = op2.Set(3000), op2.Set(2000), op2.Set(900)
    base = op2.Map(mid, nodes, 3, rng.integers(0, 900, (2000, 3)).astype(np.int32))
    c2m = op2.Map(cells, mid, 1, rng.integers(0, 2000, (3000, 1)).astype(np.int32))
    cm = op2.ComposedMap(op2.PermutedMap(base, [2, 0, 1]), c2m)
    w = op2.Dat(cells ** 3, rng.random((3000, 3)), np.float64)
    r = op2.Dat(nodes, np.zeros(900), np.float64)
    k = op2.Kernel("static void sc(double *r, const double *w) { for (int k = 0; k < 3; ++k) r[k] += w[k]; }", "sc")
    op2.par_loop(k, cells, r(op2.INC, cm), w(op2.READ))
    tab = base.values[:, [2, 0, 1]][c2m.values[:, 0]]
    want = np.zeros(900)
    np.add.at(want, tab.ravel(), np.asarray(w.data).ravel())
    assert np.allclose(np.asarray(r.data), want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("split", kat_cases.SUBSET_SPLITS, ids=["all", "nocore", "half"])
@pytest.mark.parametrize("case", kat_cases.SUBSET_CASES, ids=lambda c: c.__name__)
def test_subset_cases(device, case, split):
    case(op2, split)


def test_subset_matrix(device):
    kat_cases.case_subset_matrix(op2)


@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
@pytest.mark.parametrize("name,case", kat_cases.MIXED_CASES, ids=[n for n, _ in kat_cases.MIXED_CASES])
def test_mixed_cases(device, name, case, strategy):
    configuration["strategy"] = strategy
    case(op2)


@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
def test_mixed_random_against_oracle(device, strategy):
    """A mixed (cell, vertex) space on a 16x16 triangle mesh: MixedMat and
    MixedDat INC of a non-symmetric element kernel, device against the
    oracle's sequential wrapper on the same inputs (1e-12)."""
    import numpy as np
    import oracle_op2
    from pyop2_amd.meshes import unit_square_tri, workload_rng
    configuration["strategy"] = strategy
    m = unit_square_tri(16, rng=workload_rng(1))
    rng = np.random.default_rng(11)
    cval, vval = rng.random(m.ncells), rng.random(m.nverts)
    code = ("static void mk(double A[16], double b[4], const double d[4]) {"
            " for (int i = 0; i < 4; ++i) { b[i] += d[i] * (i + 1);"
            " for (int j = 0; j < 4; ++j) A[i * 4 + j] += d[i] * d[j] + 0.5 * i - 0.25 * j; } }")
    out = []
    for ns in (op2, oracle_op2):
        cells, verts = ns.Set(m.ncells), ns.Set(m.nverts)
        ms = ns.MixedSet((cells, verts))
        mmap = ns.MixedMap((ns.Map(cells, cells, 1, np.arange(m.ncells, dtype=np.int32)),
                            ns.Map(cells, verts, 3, m.cell_vertex.astype(np.int32))))
        d = ns.MixedDat([ns.Dat(cells ** 1, cval.copy(), np.float64), ns.Dat(verts ** 1, vval.copy(), np.float64)])
        b = ns.MixedDat(ms)
        sp = ns.Sparsity((ms ** 1, ms ** 1), {(i, j): [(rm, cm, None)] for i, rm in enumerate(mmap)
                                              for j, cm in enumerate(mmap)})
        A = ns.Mat(sp)
        ns.par_loop(ns.Kernel(code, "mk"), cells, A(ns.INC, (mmap, mmap)), b(ns.INC, mmap), d(ns.READ, mmap))
        A.assemble()
        out.append(([np.asarray(A[i, j].values) for i in range(2) for j in range(2)],
                    [np.asarray(x.data_ro) for x in b]))
    (Ad, bd), (Ao, bo) = out
    for x, y in zip(Ad + bd, Ao + bo):
        assert np.linalg.norm(x - y) <= 1e-12 * np.linalg.norm(y)


@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
@pytest.mark.parametrize("name,case", kat_cases.EXTRUDED_CASES, ids=[n for n, _ in kat_cases.EXTRUDED_CASES])
def test_extruded_cases(device, name, case, strategy):
    configuration["strategy"] = strategy
    case(op2)
