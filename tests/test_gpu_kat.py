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
@pytest.mark.parametrize("name,case", kat_cases.MONOLITHIC_CASES, ids=[n for n, _ in kat_cases.MONOLITHIC_CASES])
def test_monolithic_cases(device, name, case, strategy):
    configuration["strategy"] = strategy
    case(op2)


@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
def test_monolithic_random_against_oracle(device, strategy):
    """A monolithic MixedMat (nest=False) on a (cell x2, vertex) mixed space
    of a 12x12 triangle mesh: the dof pattern equals the oracle's
    monolithic_pattern of the block patterns, and the values the oracle's
    nested blocks placed at their offsets (1e-12)."""
    import numpy as np
    import oracle_op2
    from oracle import sparsity as osp
    from pyop2_amd.meshes import unit_square_tri, workload_rng
    configuration["strategy"] = strategy
    m = unit_square_tri(12, rng=workload_rng(1), permute=True)
    rng = np.random.default_rng(12)
    cval, vval = rng.random((m.ncells, 2)), rng.random(m.nverts)
    code = ("static void mk(double A[5][5], const double d[5]) {"
            " for (int i = 0; i < 5; ++i) for (int j = 0; j < 5; ++j)"
            " A[i][j] += d[i] * d[j] + 0.5 * i - 0.25 * j; }")
    out = []
    for ns, nest in ((op2, False), (oracle_op2, True)):
        cells, verts = ns.Set(m.ncells), ns.Set(m.nverts)
        ms = ns.MixedSet((cells, verts))
        c2c = ns.Map(cells, cells, 1, np.arange(m.ncells, dtype=np.int32))
        c2v = ns.Map(cells, verts, 3, m.cell_vertex.astype(np.int32))
        mmap = ns.MixedMap((c2c, c2v))
        d = ns.MixedDat([ns.Dat(cells ** 2, cval.copy(), np.float64), ns.Dat(verts ** 1, vval.copy(), np.float64)])
        mds = ns.MixedDataSet((cells ** 2, verts ** 1))
        maps = {(i, j): [(rm, cm, None)] for i, rm in enumerate(mmap) for j, cm in enumerate(mmap)}
        sp = ns.Sparsity((mds, mds), maps, nest=nest) if ns is op2 else ns.Sparsity((mds, mds), maps)
        A = ns.Mat(sp)
        ns.par_loop(ns.Kernel(code, "mk"), cells, A(ns.INC, (mmap, mmap)), d(ns.READ, mmap))
        A.assemble()
        out.append(A)
    Ad, Ao = out
    dense = np.block([[np.asarray(Ao[i, j].values) for j in range(2)] for i in range(2)])
    assert np.linalg.norm(Ad.values - dense) <= 1e-12 * np.linalg.norm(dense)
    tri = m.cell_vertex.astype(np.int32)
    ident = np.arange(m.ncells, dtype=np.int32)[:, None]
    blocks = [[osp.pattern(m.ncells, m.ncells, [(ident, ident, m.ncells)]),
               osp.pattern(m.ncells, m.nverts, [(ident, tri, m.ncells)], set_diag=False)],
              [osp.pattern(m.nverts, m.ncells, [(tri, ident, m.ncells)], set_diag=False),
               osp.pattern(m.nverts, m.nverts, [(tri, tri, m.ncells)])]]
    rp, ci = osp.monolithic_pattern(blocks, [2, 1], [2, 1], [m.ncells, m.nverts], [m.ncells, m.nverts])
    grp, gci, _ = Ad.csr()
    assert np.array_equal(grp, rp) and np.array_equal(gci, ci)


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


@pytest.mark.parametrize("periodic", [False, True], ids=["plain", "periodic"])
@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
def test_extruded_regions_random_against_oracle(device, strategy, periodic):
    """Prism columns over a 6x6 triangle mesh with random variable layers
    (1 to 6 cells a column), plain or periodic (the top vertices of a
    column's last cell wrapping to its bottom ones): a vertex Mat and Dat
    over all cells, a cell Mat and Dat over interior facets (jumps of a
    random cell field), a vertex Dat over the bottom and top cells with the
    layer argument; device against the oracle's sequential wrapper on the
    same inputs (1e-12), and the same patterns."""
    import numpy as np
    import oracle_op2
    from pyop2_amd.meshes import unit_square_tri, workload_rng
    configuration["strategy"] = strategy
    m = unit_square_tri(6, rng=workload_rng(2))
    rng = np.random.default_rng(12)
    nl = 8                                   # vertex layers reserved a column
    bottom = rng.integers(0, 2, m.ncells).astype(np.int32)
    top = (bottom + rng.integers(2, 8, m.ncells)).astype(np.int32)
    lay = np.stack([bottom, np.minimum(top, nl)], axis=1).astype(np.int32)
    # a vertex's layers start at 0 in every column, a cell's at its column's bottom
    vbase = np.array([[v * nl + b + d for v in tri for d in (0, 1)] for tri, b in zip(m.cell_vertex, bottom)],
                     dtype=np.int32)
    cbase = (np.arange(m.ncells) * nl + bottom).astype(np.int32)
    u, q = rng.random(m.nverts * nl), rng.random(m.ncells * nl)
    k_all = ("static void pr(double A[6][6], double b[6], const double x[6]) {"
             " for (int i = 0; i < 6; ++i) { b[i] += x[i] * (i + 1);"
             " for (int j = 0; j < 6; ++j) A[i][j] += x[i] * x[j] + 0.5 * i - 0.25 * j; } }")
    k_fac = ("static void fc(double A[2][2], double c[2], const double q[2]) {"
             " c[0] += q[1] - 0.5 * q[0]; c[1] -= 0.25 * q[0] + q[1];"
             " A[0][0] += q[0]; A[0][1] -= q[1]; A[1][0] -= q[0] * q[1]; A[1][1] += 2.0 * q[1]; }")
    k_end = ("static void ed(double b[6], const double x[6], int layer) {"
             " for (int i = 0; i < 6; ++i) b[i] += x[i] * (layer + 1) + i; }")
    out = []
    for ns in (op2, oracle_op2):
        base = ns.Set(m.ncells)
        ext = ns.ExtrudedSet(base, lay, extruded_periodic=periodic)
        verts, cells = ns.Set(m.nverts * nl), ns.Set(m.ncells * nl)
        vmap = ns.Map(ext, verts, 6, vbase, offset=np.ones(6, dtype=np.int32),
                      offset_quotient=np.tile([0, 1], 3).astype(np.int32) if periodic else None)
        cmap = ns.Map(ext, cells, 1, cbase, offset=np.ones(1, dtype=np.int32))
        ud = ns.Dat(verts ** 1, u.copy(), np.float64)
        qd = ns.Dat(cells ** 1, q.copy(), np.float64)
        b, be = ns.Dat(verts ** 1, np.zeros(m.nverts * nl), np.float64), ns.Dat(verts ** 1, np.zeros(m.nverts * nl), np.float64)
        c = ns.Dat(cells ** 1, np.zeros(m.ncells * nl), np.float64)
        Av = ns.Mat(ns.Sparsity((verts ** 1, verts ** 1), [(vmap, vmap, None)]))
        Ac = ns.Mat(ns.Sparsity((cells ** 1, cells ** 1), [(cmap, cmap, (ns.ALL, ns.ON_INTERIOR_FACETS))]))
        ns.par_loop(ns.Kernel(k_all, "pr"), ext, Av(ns.INC, (vmap, vmap)), b(ns.INC, vmap), ud(ns.READ, vmap))
        ns.par_loop(ns.Kernel(k_fac, "fc"), ext, Ac(ns.INC, (cmap, cmap)), c(ns.INC, cmap), qd(ns.READ, cmap),
                    iteration_region=ns.ON_INTERIOR_FACETS)
        for region in (ns.ON_BOTTOM, ns.ON_TOP):
            ns.par_loop(ns.Kernel(k_end, "ed"), ext, be(ns.INC, vmap), ud(ns.READ, vmap),
                        iteration_region=region, pass_layer_arg=True)
        Av.assemble()
        Ac.assemble()
        out.append([np.asarray(Av.values), np.asarray(Ac.values), np.asarray(b.data_ro),
                    np.asarray(c.data_ro), np.asarray(be.data_ro)])
    for x, y in zip(*out):
        assert np.count_nonzero(y) > 0
        assert np.linalg.norm(x - y) <= 1e-12 * np.linalg.norm(y)
    assert np.array_equal(out[0][1] != 0, out[1][1] != 0)


@pytest.mark.parametrize("strategy", ["atomic", "coloured"])
@pytest.mark.parametrize("name,case", kat_cases.EXTRUDED_CASES, ids=[n for n, _ in kat_cases.EXTRUDED_CASES])
def test_extruded_cases(device, name, case, strategy):
    configuration["strategy"] = strategy
    case(op2)


def test_arglist_matches_argtypes(device):
    """Parloop.arglist (iteration set pointers, data pointers, distinct map
    pointers; parloop.py:203-212) has the length GlobalKernel.argtypes
    declares after start and end (global_kernel.py:410-417): plain, Subset,
    extruded, mixed and Mat arguments."""
    import numpy as np
    from pyop2_amd.parloop import LegacyParloop
    s, t = op2.Set(4), op2.Set(10)
    m = op2.Map(s, t, 2, np.arange(8, dtype=np.int32).reshape(4, 2))
    d = op2.Dat(t ** 1, dtype=np.float64)
    g = op2.Global(1, 0.0, np.float64)
    A = op2.Mat(op2.Sparsity((t ** 1, t ** 1), [(m, m, None)]), np.float64)
    ext = op2.ExtrudedSet(s, 3)
    me = op2.Map(ext, t, 2, np.array([[0, 1], [2, 3], [4, 5], [6, 7]], dtype=np.int32), offset=[1, 1])
    ms = op2.MixedSet((t, t))
    mm = op2.MixedMap((m, m))
    md = op2.MixedDat(ms)
    loops = [
        LegacyParloop(op2.Kernel("static void a1(double *x, double *g) {}", "a1"), s, d(op2.INC, m), g(op2.INC)),
        LegacyParloop(op2.Kernel("static void a2(double *x) {}", "a2"), s([1, 3]), d(op2.INC, m)),
        LegacyParloop(op2.Kernel("static void a3(double *x, double *y) {}", "a3"), ext, d(op2.INC, me),
                      d(op2.READ, me), iteration_region=op2.ON_INTERIOR_FACETS),
        LegacyParloop(op2.Kernel("static void a4(double *x) {}", "a4"), s, md(op2.INC, mm)),
        LegacyParloop(op2.Kernel("static void a5(double *a, double *x) {}", "a5"), s, A(op2.INC, (m, m)),
                      d(op2.READ, m)),
    ]
    for pl in loops:
        assert len(pl.global_kernel.argtypes) == 2 + len(pl.arglist)


def test_parloop_dat_versions(device):
    """test_dats.py:195-204 and 234-258: a par_loop writing a Dat (or each
    component of a MixedDat) counts once per written Dat."""
    import numpy as np
    s = op2.Set(5)
    d3 = op2.Dat(s ** 1, data=None, dtype=np.uint32)
    op2.par_loop(op2.Kernel("static void write(unsigned int *v) { *v = 1; }", "write"), s, d3(op2.WRITE))
    assert d3.dat_version == 1
    d3, d4 = op2.Dat(s ** 1, data=None, dtype=np.uint32), op2.Dat(s ** 1, data=None, dtype=np.uint32)
    d3d4 = op2.MixedDat([d3, d4])
    m = op2.Map(s, s, 1, values=[0, 1, 2, 3, 4])
    op2.par_loop(op2.Kernel("static void write2(unsigned int *v) { v[0] = 1; v[1] = 2; }", "write2"), s,
                 d3d4(op2.WRITE, op2.MixedMap([m, m])))
    assert d3.dat_version == 1 and d4.dat_version == 1 and d3d4.dat_version == 2
