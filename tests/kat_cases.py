"""Known-answer cases for the gather/scatter semantics of indirect par_loops.

Each case takes an ``op2``-like namespace (``pyop2_amd.op2`` on the device,
``oracle_op2`` on the CPU) and an iteration-set size triple, builds the
problem, runs it and checks the closed-form answer that the reference's own
unit tests assert (the source of each answer is cited).  The kernels are
written here; only the expected values come from the reference tests.
"""
import numpy as np

N = 4096
SPLITS = [(N, N, N), (0, N, N), (N // 2, N, N)]   # test_indirect_loop.py:46-50


def _iter(op2, split):
    return op2.Set(split, "iterset")


def case_write(op2, split):
    """test_indirect_loop.py:131-136 -- WRITE 42 through a reversed map."""
    it = _iter(op2, split)
    ind = op2.Set(N, "ind")
    x = op2.Dat(ind, np.arange(N, dtype=np.uint32), np.uint32)
    m = op2.Map(it, ind, 1, np.arange(N, dtype=np.int32)[::-1].copy())
    op2.par_loop(op2.Kernel("static void put42(unsigned int *v) { v[0] = 42u; }", "put42"), it, x(op2.WRITE, m))
    assert (np.asarray(x.data) == 42).all()


def case_rw(op2, split):
    """test_indirect_loop.py:138-143 -- RW increment: sum = n(n+1)/2."""
    it = _iter(op2, split)
    ind = op2.Set(N, "ind")
    x = op2.Dat(ind, np.arange(N, dtype=np.uint32), np.uint32)
    m = op2.Map(it, ind, 1, np.arange(N, dtype=np.int32)[::-1].copy())
    op2.par_loop(op2.Kernel("static void bump(unsigned int *v) { v[0] = v[0] + 1u; }", "bump"), it, x(op2.RW, m))
    assert int(np.asarray(x.data, dtype=np.int64).sum()) == N * (N + 1) // 2


def case_inc_unitset(op2, split):
    """test_indirect_loop.py:145-151 -- every element INCs the one target: nelems."""
    it = _iter(op2, split)
    unit = op2.Set(1, "unit")
    u = op2.Dat(unit, np.zeros(1, dtype=np.uint32), np.uint32)
    m = op2.Map(it, unit, 1, np.zeros(N, dtype=np.int32))
    op2.par_loop(op2.Kernel("static void one(unsigned int *v) { v[0] += 1u; }", "one"), it, u(op2.INC, m))
    assert int(u.data[0]) == N


def case_max(op2, split):
    """test_indirect_loop.py:153-161 -- MAX of -10 against -5 is -5."""
    it = _iter(op2, split)
    ind = op2.Set(N, "ind")
    m = op2.Map(it, ind, 1, np.arange(N, dtype=np.int32)[::-1].copy())
    a = op2.Dat(ind, np.full(N, -10, dtype=np.int32), np.int32)
    b = op2.Dat(ind, np.full(N, -5, dtype=np.int32), np.int32)
    op2.par_loop(op2.Kernel("static void upto(int *a, const int *b) { if (*b > *a) *a = *b; }", "upto"),
                 it, a(op2.MAX, m), b(op2.READ, m))
    assert (np.asarray(a.data) == -5).all()


def case_min(op2, split):
    """test_indirect_loop.py:163-171 -- MIN of 10 against 5 is 5."""
    it = _iter(op2, split)
    ind = op2.Set(N, "ind")
    m = op2.Map(it, ind, 1, np.arange(N, dtype=np.int32)[::-1].copy())
    a = op2.Dat(ind, np.full(N, 10, dtype=np.int32), np.int32)
    b = op2.Dat(ind, np.full(N, 5, dtype=np.int32), np.int32)
    op2.par_loop(op2.Kernel("static void downto(int *a, const int *b) { if (*b < *a) *a = *b; }", "downto"),
                 it, a(op2.MIN, m), b(op2.READ, m))
    assert (np.asarray(a.data) == 5).all()


def case_global_read(op2, split):
    """test_indirect_loop.py:173-183 -- divide by a Global 2."""
    it = _iter(op2, split)
    ind = op2.Set(N, "ind")
    x = op2.Dat(ind, np.arange(N, dtype=np.uint32), np.uint32)
    m = op2.Map(it, ind, 1, np.arange(N, dtype=np.int32)[::-1].copy())
    g = op2.Global(1, 2, np.uint32, "g")
    op2.par_loop(op2.Kernel("static void halve(unsigned int *v, const unsigned int *d) { v[0] /= d[0]; }",
                            "halve"), it, x(op2.RW, m), g(op2.READ))
    assert int(np.asarray(x.data, dtype=np.int64).sum()) == sum(v // 2 for v in range(N))


def case_global_inc(op2, split):
    """test_indirect_loop.py:185-200 -- RW + Global INC: both n(n+1)/2."""
    it = _iter(op2, split)
    ind = op2.Set(N, "ind")
    x = op2.Dat(ind, np.arange(N, dtype=np.uint32), np.uint32)
    m = op2.Map(it, ind, 1, np.arange(N, dtype=np.int32)[::-1].copy())
    g = op2.Global(1, 0, np.uint32, "g")
    op2.par_loop(op2.Kernel("static void bump_sum(unsigned int *v, unsigned int *s) { v[0] += 1u; s[0] += v[0]; }",
                            "bump_sum"), it, x(op2.RW, m), g(op2.INC))
    assert int(np.asarray(x.data, dtype=np.int64).sum()) == N * (N + 1) // 2
    assert int(g.data[0]) == N * (N + 1) // 2


def case_2d_dat(op2, split):
    """test_indirect_loop.py:202-207 -- WRITE both components of a dim-2 Dat."""
    it = _iter(op2, split)
    ind = op2.Set(N, "ind")
    x2 = op2.Dat(ind ** 2, np.stack([np.arange(N), np.arange(N)], axis=1).astype(np.uint32), np.uint32)
    m = op2.Map(it, ind, 1, np.arange(N, dtype=np.int32)[::-1].copy())
    op2.par_loop(op2.Kernel("static void pair(unsigned int *v) { v[0] = 42u; v[1] = 43u; }", "pair"),
                 it, x2(op2.WRITE, m))
    d = np.asarray(x2.data)
    assert (d[:, 0] == 42).all() and (d[:, 1] == 43).all()


def case_edge_sum(op2, split=None, mode="WRITE"):
    """test_indirect_loop.py:209-232, test_vector_map.py:109-126 (WRITE),
    test_iteration_space_dats.py:89-104 (INC): edge = sum of its two nodes,
    expected 1, 3, 5, ..."""
    nedges = N - 1
    nodes = op2.Set(N, "nodes")
    edges = op2.Set(nedges, "edges")
    nv = op2.Dat(nodes, np.arange(N, dtype=np.uint32), np.uint32)
    ev = op2.Dat(edges, np.zeros(nedges, dtype=np.uint32), np.uint32)
    e2n = op2.Map(edges, nodes, 2, np.stack([np.arange(nedges), np.arange(1, N)], axis=1).astype(np.int32))
    if mode == "WRITE":
        k = op2.Kernel("static void esum(unsigned int *e, const unsigned int *n) { e[0] = n[0] + n[1]; }", "esum")
        op2.par_loop(k, edges, ev(op2.WRITE), nv(op2.READ, e2n))
    else:
        k = op2.Kernel("static void esum_inc(unsigned int *e, const unsigned int *n)"
                       " { for (int i = 0; i < 2; ++i) e[0] += n[i]; }", "esum_inc")
        op2.par_loop(k, edges, ev(op2.INC), nv(op2.READ, e2n))
    assert (np.asarray(ev.data) == np.arange(1, 2 * nedges + 1, 2)).all()


NNODES = 4096
NELE = NNODES // 2


def _node2ele(op2):
    node = op2.Set(NNODES, "node")
    ele = op2.Set(NELE, "ele")
    m = op2.Map(node, ele, 1, (np.arange(NNODES) // 2).astype(np.int32))
    return node, ele, m


def case_vmap_read(op2, split=None, dim=1):
    """test_vector_map.py:128-137 / :163-175 -- every node reads its element."""
    node, ele, n2e = _node2ele(op2)
    if dim == 1:
        d = op2.Dat(node, np.zeros(NNODES, dtype=np.int32), np.int32)
        vd = op2.Dat(ele, np.arange(NELE, dtype=np.int32), np.int32)
        k = op2.Kernel("static void rd(int *d, const int *v) { d[0] = v[0]; }", "rd")
    else:
        d = op2.Dat(node ** 2, np.zeros((NNODES, 2), dtype=np.int32), np.int32)
        vd = op2.Dat(ele ** 2, np.arange(NELE * 2, dtype=np.int32).reshape(NELE, 2), np.int32)
        k = op2.Kernel("static void rd2(int d[2], const int v[1][2]) { d[0] = v[0][0]; d[1] = v[0][1]; }", "rd2")
    op2.par_loop(k, node, d(op2.WRITE), vd(op2.READ, n2e))
    dd, vv = np.asarray(d.data), np.asarray(vd.data)
    assert (dd[::2] == vv).all() and (dd[1::2] == vv).all()


def case_vmap_write(op2, split=None, dim=1):
    """test_vector_map.py:139-147 / :177-186 -- two nodes WRITE the same value
    into one element (conflicting writes of equal values)."""
    node, ele, n2e = _node2ele(op2)
    if dim == 1:
        vd = op2.Dat(ele, np.zeros(NELE, dtype=np.int32), np.int32)
        k = op2.Kernel("static void wr(int *v) { v[0] = 2; }", "wr")
        op2.par_loop(k, node, vd(op2.WRITE, n2e))
        assert (np.asarray(vd.data) == 2).all()
    else:
        vd = op2.Dat(ele ** 2, np.zeros((NELE, 2), dtype=np.int32), np.int32)
        k = op2.Kernel("static void wr2(int v[1][2]) { v[0][0] = 2; v[0][1] = 3; }", "wr2")
        op2.par_loop(k, node, vd(op2.WRITE, n2e))
        v = np.asarray(vd.data)
        assert (v[:, 0] == 2).all() and (v[:, 1] == 3).all()


def case_vmap_inc(op2, split=None, dim=1):
    """test_vector_map.py:149-161 / :188-208 -- two nodes INC into each element."""
    node, ele, n2e = _node2ele(op2)
    if dim == 1:
        vd = op2.Dat(ele, np.full(NELE, 3, dtype=np.int32), np.int32)
        d = op2.Dat(node, np.arange(NNODES, dtype=np.int32), np.int32)
        k = op2.Kernel("static void ad(int *v, const int *d) { v[0] += d[0]; }", "ad")
        op2.par_loop(k, node, vd(op2.INC, n2e), d(op2.READ))
        expected = 3 + np.arange(0, NNODES, 2) + np.arange(1, NNODES, 2)
        assert (np.asarray(vd.data) == expected).all()
    else:
        init = np.empty((NELE, 2), dtype=np.int32)
        init[:, 0], init[:, 1] = 3, 4
        vd = op2.Dat(ele ** 2, init, np.int32)
        d = op2.Dat(node ** 2, np.arange(2 * NNODES, dtype=np.int32).reshape(NNODES, 2), np.int32)
        k = op2.Kernel("static void ad2(int v[1][2], const int d[2]) { v[0][0] += d[0]; v[0][1] += d[1]; }", "ad2")
        op2.par_loop(k, node, vd(op2.INC, n2e), d(op2.READ))
        v = np.asarray(vd.data)
        e0 = 3 + np.arange(0, 2 * NNODES, 4) + np.arange(2, 2 * NNODES, 4)
        e1 = 4 + np.arange(1, 2 * NNODES, 4) + np.arange(3, 2 * NNODES, 4)
        assert (v[:, 0] == e0).all() and (v[:, 1] == e1).all()


# ---- matrices: the two-triangle mesh of test_matrices.py:47-58, 166-222 --------
ELEM_NODE = np.array([0, 1, 3, 2, 3, 1], dtype=np.int32)
COORDS = np.array([(0.0, 0.0), (2.0, 0.0), (1.0, 1.0), (0.0, 1.5)])
# test_matrices.py:461-467 (the reference's quadrature kernel matches to 1e-5)
EXPECTED_MASS = np.array([(0.25, 0.125, 0.0, 0.125),
                          (0.125, 0.291667, 0.0208333, 0.145833),
                          (0.0, 0.0208333, 0.0416667, 0.0208333),
                          (0.125, 0.145833, 0.0208333, 0.291667)])
# test_matrices.py:488-492
EXPECTED_RHS = np.array([0.9999999523522115, 1.3541666031724144, 0.2499999883507239, 1.6458332580869566])

# exact P1 consistent mass matrix: |T|/12 * (1 + delta_ij)
MASS_KERNEL = r"""
static void p1_mass_matrix(double A[3][3], const double X[3][2])
{
  const double det = (X[1][0] - X[0][0]) * (X[2][1] - X[0][1])
                   - (X[2][0] - X[0][0]) * (X[1][1] - X[0][1]);
  const double s = fabs(det) / 24.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      A[i][j] += (i == j ? 2.0 : 1.0) * s;
}
"""
RHS_KERNEL = r"""
static void p1_mass_rhs(double b[3], const double X[3][2], const double f[3])
{
  const double det = (X[1][0] - X[0][0]) * (X[2][1] - X[0][1])
                   - (X[2][0] - X[0][0]) * (X[1][1] - X[0][1]);
  const double s = fabs(det) / 24.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      b[i] += (i == j ? 2.0 : 1.0) * s * f[j];
}
"""
VEC_MASS_KERNEL = r"""
static void p1_vmass_matrix(double A[6][6], const double X[3][2])
{
  const double det = (X[1][0] - X[0][0]) * (X[2][1] - X[0][1])
                   - (X[2][0] - X[0][0]) * (X[1][1] - X[0][1]);
  const double s = fabs(det) / 24.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      for (int c = 0; c < 2; ++c)
        A[2 * i + c][2 * j + c] += (i == j ? 2.0 : 1.0) * s;
}
"""


def _two_triangles(op2):
    nodes = op2.Set(4, "nodes")
    elements = op2.Set(2, "elements")
    en = op2.Map(elements, nodes, 3, ELEM_NODE, "elem_node")
    coords = op2.Dat(nodes ** 2, COORDS, np.float64)
    return nodes, elements, en, coords


def case_mass_matrix(op2, split=None):
    """test_matrices.py:644-653 -- assembled P1 mass matrix, and nnz 14
    (test_matrices.py:770-772: nbytes = 14 * 8)."""
    nodes, elements, en, coords = _two_triangles(op2)
    sp = op2.Sparsity((nodes ** 1, nodes ** 1), [(en, en, None)])
    A = op2.Mat(sp, np.float64)
    op2.par_loop(op2.Kernel(MASS_KERNEL, "p1_mass_matrix"), elements, A(op2.INC, (en, en)),
                 coords(op2.READ, en))
    A.assemble()
    assert np.allclose(A.values, EXPECTED_MASS, rtol=1e-5, atol=0)
    assert int(np.sum(sp.nnz)) == 14
    return A.values


def case_vector_mass_matrix(op2, split=None):
    """test_matrices.py:468-486 -- the block (2x2) mass matrix = M (x) I."""
    nodes, elements, en, coords = _two_triangles(op2)
    sp = op2.Sparsity((nodes ** 2, nodes ** 2), [(en, en, None)])
    A = op2.Mat(sp, np.float64)
    op2.par_loop(op2.Kernel(VEC_MASS_KERNEL, "p1_vmass_matrix"), elements, A(op2.INC, (en, en)),
                 coords(op2.READ, en))
    A.assemble()
    assert np.allclose(A.values, np.kron(EXPECTED_MASS, np.eye(2)), rtol=1e-5, atol=1e-12)


def case_rhs(op2, split=None):
    """test_matrices.py:655-666 -- b = M f with f = (1, 2, 3, 4)."""
    nodes, elements, en, coords = _two_triangles(op2)
    f = op2.Dat(nodes, np.array([1.0, 2.0, 3.0, 4.0]), np.float64)
    b = op2.Dat(nodes, np.zeros(4), np.float64)
    op2.par_loop(op2.Kernel(RHS_KERNEL, "p1_mass_rhs"), elements, b(op2.INC, en), coords(op2.READ, en),
                 f(op2.READ, en))
    M = np.array([[0.25, 0.125, 0.0, 0.125], [0.125, 0.25 + 1 / 24, 1 / 48, 0.125 + 1 / 48],
                  [0.0, 1 / 48, 1 / 24, 1 / 48], [0.125, 0.125 + 1 / 48, 1 / 48, 0.25 + 1 / 24]])
    exact = M @ np.array([1.0, 2.0, 3.0, 4.0])
    assert np.allclose(np.asarray(b.data), exact, rtol=1e-14, atol=0)
    assert np.allclose(np.asarray(b.data), EXPECTED_RHS, rtol=1e-7, atol=0)


def case_write_vs_inc_matrix(op2, split=None):
    """test_matrices.py:680-696 -- INC ones: sum 9 * ncells; then WRITE ones
    over the same pattern: every one of the 14 entries is 1."""
    nodes, elements, en, coords = _two_triangles(op2)
    A = op2.Mat(op2.Sparsity((nodes ** 1, nodes ** 1), [(en, en, None)]), np.float64)
    g = op2.Global(1, 1.0, np.float64, "g")
    op2.par_loop(op2.Kernel("static void add_g(double e[3][3], const double *g)"
                            " { for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) e[i][j] += g[0]; }",
                            "add_g"), elements, A(op2.INC, (en, en)), g(op2.READ))
    A.assemble()
    assert A.values.sum() == 3 * 3 * 2
    op2.par_loop(op2.Kernel("static void put_g(double e[3][3], const double *g)"
                            " { for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) e[i][j] = g[0]; }",
                            "put_g"), elements, A(op2.WRITE, (en, en)), g(op2.READ))
    A.assemble()
    assert A.values.sum() == (3 * 3 - 2) * 2


def case_sparsity_nnz(op2, split=None):
    """test_matrices.py:555-566 -- diagonal space: nnz [1,2,1,2] and, for
    different row/column sets, [0,3,0,3]."""
    s = op2.Set(1)
    d = op2.Set(4)
    m = op2.Map(s, d, 2, [1, 3])
    d2 = op2.Set(4)
    m2 = op2.Map(s, d2, 3, [1, 2, 3])
    sp = op2.Sparsity((d ** 1, d ** 1), [(m, m, None)])
    sp2 = op2.Sparsity((d ** 1, d2 ** 1), [(m, m2, None)])
    assert list(sp.nnz) == [1, 2, 1, 2]
    assert list(sp2.nnz) == [0, 3, 0, 3]


LOOP_CASES = [case_write, case_rw, case_inc_unitset, case_max, case_min, case_global_read,
              case_global_inc, case_2d_dat]
OTHER_CASES = [
    ("edge_sum_write", lambda op2: case_edge_sum(op2, mode="WRITE")),
    ("edge_sum_inc", lambda op2: case_edge_sum(op2, mode="INC")),
    ("vmap_read_1d", lambda op2: case_vmap_read(op2, dim=1)),
    ("vmap_read_2d", lambda op2: case_vmap_read(op2, dim=2)),
    ("vmap_write_1d", lambda op2: case_vmap_write(op2, dim=1)),
    ("vmap_write_2d", lambda op2: case_vmap_write(op2, dim=2)),
    ("vmap_inc_1d", lambda op2: case_vmap_inc(op2, dim=1)),
    ("vmap_inc_2d", lambda op2: case_vmap_inc(op2, dim=2)),
    ("mass_matrix", case_mass_matrix),
    ("vector_mass_matrix", case_vector_mass_matrix),
    ("rhs", case_rhs),
    ("write_vs_inc_matrix", case_write_vs_inc_matrix),
    ("sparsity_nnz", case_sparsity_nnz),
]
