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
def case_empty_iterset(op2):
    """test_api.py:1593-1601: a loop over a zero-sized set whose map holds no
    values runs nothing.  Over the empty set an INC Dat keeps its values,
    INC and MIN Globals keep theirs (the INC adds a zero temporary,
    parloop.py:514-530), and a Mat built from a pair of empty maps holds its
    zero diagonal only."""
    s1, s2 = op2.Set(0), op2.Set(10)
    m = op2.Map(s1, s2, 3)
    d = op2.Dat(s2 ** 1, np.arange(10, dtype=np.int32), np.int32)
    op2.par_loop(op2.Kernel("static void k(int *x) {}", "k"), s1, d(op2.READ, m))
    op2.par_loop(op2.Kernel("static void ki(int *x) { x[0] += 1; }", "ki"), s1, d(op2.INC, m))
    assert list(np.asarray(d.data)) == list(range(10))
    g, h = op2.Global(1, 5.0, np.float64, "g"), op2.Global(1, 5.0, np.float64, "h")
    op2.par_loop(op2.Kernel("static void kg(double *g, double *h) { g[0] += 1.0; h[0] = h[0] < 0.0 ? h[0] : 0.0; }",
                            "kg"), s1, g(op2.INC), h(op2.MIN))
    assert float(np.asarray(g.data).reshape(-1)[0]) == 5.0
    assert float(np.asarray(h.data).reshape(-1)[0]) == 5.0
    e = op2.Map(s1, s2, 3, np.zeros((0, 3), dtype=np.int32))
    sp = op2.Sparsity((s2 ** 1, s2 ** 1), [(e, e, None)])
    A = op2.Mat(sp, np.float64)
    op2.par_loop(op2.Kernel("static void km(double a[3][3]) { a[0][0] += 1.0; }", "km"), s1, A(op2.INC, (e, e)))
    A.assemble()
    assert int(np.sum(sp.nnz)) == 10
    assert not np.asarray(A.values).any()


def case_dat_view(op2):
    """A DatView hands the kernel one component (dat.py:668-760,
    builder.py:303-410): INC of 1 into component 1 of a 2-vector through an
    arity-2 edge map counts each node's edges there and leaves component 0;
    READ of component 0 through the same map sums its two ends; a direct
    WRITE of component 1 sets it per node."""
    nodes, edges = op2.Set(5), op2.Set(4)
    e2n = op2.Map(edges, nodes, 2, np.array([[0, 1], [1, 2], [2, 3], [3, 4]], dtype=np.int32))
    v = op2.Dat(nodes ** 2, np.stack([np.arange(5.0), np.zeros(5)], axis=1), np.float64)
    inc = op2.Kernel("static void inc2(double x[2]) { x[0] += 1.0; x[1] += 1.0; }", "inc2")
    op2.par_loop(inc, edges, op2.DatView(v, 1)(op2.INC, e2n))
    assert list(np.asarray(v.data)[:, 1]) == [1.0, 2.0, 2.0, 2.0, 1.0]
    assert list(np.asarray(v.data)[:, 0]) == [0.0, 1.0, 2.0, 3.0, 4.0]
    s = op2.Dat(edges ** 1, np.zeros(4), np.float64)
    add = op2.Kernel("static void ends(double *s, const double x[2]) { s[0] = x[0] + x[1]; }", "ends")
    op2.par_loop(add, edges, s(op2.WRITE), op2.DatView(v, 0)(op2.READ, e2n))
    assert list(np.asarray(s.data)) == [1.0, 3.0, 5.0, 7.0]
    put = op2.Kernel("static void put(double *x) { x[0] = 9.0; }", "put")
    op2.par_loop(put, nodes, op2.DatView(v, 1)(op2.WRITE))
    assert list(np.asarray(v.data)[:, 1]) == [9.0] * 5 and list(np.asarray(v.data)[:, 0]) == [0.0, 1.0, 2.0, 3.0, 4.0]


def case_region_on_plain_set(op2):
    """A region other than interior facets on a set that is not extruded
    changes nothing (builder.py:771-795 only shape the layer loop): every
    element runs once."""
    it, tgt = op2.Set(6), op2.Set(3)
    m = op2.Map(it, tgt, 1, np.array([0, 1, 2, 0, 1, 2], dtype=np.int32))
    for region in (op2.ON_BOTTOM, op2.ON_TOP, op2.ALL):
        d = op2.Dat(tgt ** 1, np.zeros(3), np.float64)
        op2.par_loop(op2.Kernel("static void one(double *x) { *x += 1.0; }", "one"), it, d(op2.INC, m),
                     iteration_region=region)
        assert list(np.asarray(d.data)) == [2.0, 2.0, 2.0]


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
    ("empty_iterset", case_empty_iterset),
    ("region_on_plain_set", case_region_on_plain_set),
    ("dat_view", case_dat_view),
]


# ---- permuted and composed maps (test_indirect_loop.py:277-377) -------------
COPY4_INT = "static void copy4(int *dst, const int *src) { for (int k = 0; k < 4; ++k) dst[k] = src[k]; }"
COPY2_DBL = "static void copy2(double *dst, const double *src) { for (int k = 0; k < 2; ++k) dst[k] = src[k]; }"
NODAL_A = np.array([0.0, 0.1, 0.2, 0.3, 0.4, 0.5, 0.6, 0.7])


def case_permuted_map(op2, split=None):
    """test_indirect_loop.py:277-293 -- WRITE through PermutedMap(m, [3,2,0,1]),
    READ through m = [1,2,3,0]: dst[perm-table[k]] = src[m[k]] gives [1,3,0,2]."""
    one, four = op2.Set(1), op2.Set(4)
    src = op2.Dat(op2.DataSet(four, 1), dtype=np.int32)
    dst = op2.Dat(op2.DataSet(four, 1), dtype=np.int32)
    src.data[:] = np.arange(4, dtype=np.int32)
    m = op2.Map(one, four, 4, values=[1, 2, 3, 0])
    pm = op2.PermutedMap(m, [3, 2, 0, 1])
    op2.par_loop(op2.Kernel(COPY4_INT, "copy4"), one, dst(op2.WRITE, pm), src(op2.READ, m))
    assert (np.asarray(src.data) == np.arange(4)).all()
    assert list(np.asarray(dst.data)) == [1, 3, 0, 2]
    assert list(np.asarray(pm.values).ravel()) == [1, 2, 3, 0]      # values are the base map's


def case_permuted_map_both(op2, split=None):
    """test_indirect_loop.py:295-313 -- both arguments permuted: [2,1,3,0]."""
    one, four = op2.Set(1), op2.Set(4)
    src = op2.Dat(op2.DataSet(four, 1), dtype=np.int32)
    dst = op2.Dat(op2.DataSet(four, 1), dtype=np.int32)
    src.data[:] = np.arange(4, dtype=np.int32)
    m = op2.Map(one, four, 4, values=[0, 2, 1, 3])
    op2.par_loop(op2.Kernel(COPY4_INT, "copy4"), one, dst(op2.WRITE, op2.PermutedMap(m, [3, 2, 1, 0])),
                 src(op2.READ, op2.PermutedMap(m, [0, 2, 3, 1])))
    assert list(np.asarray(dst.data)) == [2, 1, 3, 0]


def case_composed_two(op2, split=None, permuted=False):
    """test_indirect_loop.py:316-344 -- ComposedMap(A0, A1): B-element b reads
    the nodes of A-element A1[b]; with A0 pre-permuted the pairs swap."""
    setB, nodesB = op2.Set(3), op2.Set(6)
    datB = op2.Dat(op2.DataSet(nodesB, 1), dtype=np.float64)
    mapB = op2.Map(setB, nodesB, 2, values=[[0, 1], [2, 3], [4, 5]])
    setA, nodesA = op2.Set(5), op2.Set(8)
    datA = op2.Dat(op2.DataSet(nodesA, 1), dtype=np.float64)
    datA.data[:] = NODAL_A
    mapA0 = op2.Map(setA, nodesA, 2, values=[[0, 1], [2, 3], [4, 5], [6, 7], [0, 1]])
    if permuted:
        mapA0 = op2.PermutedMap(mapA0, [1, 0])
    mapA = op2.ComposedMap(mapA0, op2.Map(setB, setA, 1, values=[3, 1, 2]))
    op2.par_loop(op2.Kernel(COPY2_DBL, "copy2"), setB, datB(op2.WRITE, mapB), datA(op2.READ, mapA))
    want = [0.7, 0.6, 0.3, 0.2, 0.5, 0.4] if permuted else [0.6, 0.7, 0.2, 0.3, 0.4, 0.5]
    assert list(np.asarray(datB.data)) == want


def case_composed_three(op2, split=None, nested="none", subset=False):
    """test_indirect_loop.py:347-386 -- three maps, flat or nested either way,
    over the whole set or the subset [1]."""
    setC, nodesC = op2.Set(2), op2.Set(4)
    datC = op2.Dat(op2.DataSet(nodesC, 1), dtype=np.float64)
    mapC = op2.Map(setC, nodesC, 2, values=[[0, 1], [2, 3]])
    setB, setA, nodesA = op2.Set(3), op2.Set(5), op2.Set(8)
    datA = op2.Dat(op2.DataSet(nodesA, 1), dtype=np.float64)
    datA.data[:] = NODAL_A
    a0 = op2.Map(setA, nodesA, 2, values=[[0, 1], [2, 3], [4, 5], [6, 7], [0, 1]])
    a1 = op2.Map(setB, setA, 1, values=[3, 1, 2])
    a2 = op2.Map(setC, setB, 1, values=[2, 0])
    mapA = {"none": lambda: op2.ComposedMap(a0, a1, a2),
            "first": lambda: op2.ComposedMap(op2.ComposedMap(a0, a1), a2),
            "last": lambda: op2.ComposedMap(a0, op2.ComposedMap(a1, a2))}[nested]()
    it = op2.Subset(setC, np.array([1], dtype=np.int32)) if subset else setC
    op2.par_loop(op2.Kernel(COPY2_DBL, "copy2"), it, datC(op2.WRITE, mapC), datA(op2.READ, mapA))
    want = [0.0, 0.0, 0.6, 0.7] if subset else [0.4, 0.5, 0.6, 0.7]
    assert list(np.asarray(datC.data)) == want


def case_composed_errors(op2, split=None):
    """map.py:216-260 -- composition checks, and values are not stored."""
    import pytest
    s1, s2, s3 = op2.Set(3), op2.Set(4), op2.Set(5)
    m = op2.Map(s2, s3, 2, values=np.zeros((4, 2), dtype=np.int32))
    with pytest.raises(op2.MapTypeError):
        op2.ComposedMap(m, op2.Map(s1, s3, 1, values=[0, 1, 2]))        # iterset/toset mismatch
    with pytest.raises(op2.MapTypeError):
        op2.ComposedMap(m, op2.Map(s1, s2, 2, values=np.zeros((3, 2), dtype=np.int32)))   # arity 2
    cm = op2.ComposedMap(m, op2.Map(s1, s2, 1, values=[0, 1, 2]))
    assert (cm.iterset, cm.toset, cm.arity) == (s1, s3, 2)
    with pytest.raises(RuntimeError):
        cm.values


MAP_CASES = [
    ("permuted_map", case_permuted_map),
    ("permuted_map_both", case_permuted_map_both),
    ("composed_two_none", lambda op2: case_composed_two(op2, permuted=False)),
    ("composed_two_pre", lambda op2: case_composed_two(op2, permuted=True)),
] + [(f"composed_three_{n}_{'subset' if s else 'all'}",
      (lambda n, s: lambda op2: case_composed_three(op2, nested=n, subset=s))(n, s))
     for n in ("none", "first", "last") for s in (False, True)] + [
    ("composed_errors", case_composed_errors),
]


# ---- direct loops (test_direct_loop.py:44-200) -------------------------------
DIRECT_SPLITS = SPLITS + [(0, N // 2, N)]          # test_direct_loop.py:46-50 (the last has a halo)


def _direct(op2, split, dim=1):
    elems = op2.Set(split, "elems")
    ds = op2.DataSet(elems, dim)
    init = np.arange(N, dtype=np.uint32)
    x = op2.Dat(ds, init if dim == 1 else np.stack([init, init], axis=1), np.uint32)
    return elems, x


def case_direct_wo(op2, split):
    """test_direct_loop.py:90-95 -- WRITE 42 everywhere owned."""
    elems, x = _direct(op2, split)
    op2.par_loop(op2.Kernel("static void wo(unsigned int *v) { *v = 42u; }", "wo"), elems, x(op2.WRITE))
    assert (np.asarray(x.data) == 42).all()


def case_direct_rw(op2, split):
    """test_direct_loop.py:107-115 -- RW +1: sum of the owned part n(n+1)/2."""
    elems, x = _direct(op2, split)
    op2.par_loop(op2.Kernel("static void plus1(unsigned int *v) { *v = *v + 1u; }", "plus1"), elems, x(op2.RW))
    n = elems.size
    assert int(np.asarray(x.data_ro, dtype=np.int64).sum()) == n * (n + 1) // 2


def case_direct_global_inc(op2, split):
    """test_direct_loop.py:117-125 -- RW and Global INC in one loop."""
    elems, x = _direct(op2, split)
    g = op2.Global(1, 0, np.uint32, "g")
    op2.par_loop(op2.Kernel("static void acc(unsigned int *v, unsigned int *s) { *v += 1u; *s += *v; }", "acc"),
                 elems, x(op2.RW), g(op2.INC))
    n = elems.size
    assert int(g.data[0]) == n * (n + 1) // 2


def case_direct_global_inc_init(op2, split):
    """test_direct_loop.py:127-132 -- INC into a Global that starts at 10."""
    elems, _ = _direct(op2, split)
    g = op2.Global(1, 0, np.uint32, "g")
    g.data[0] = 10
    op2.par_loop(op2.Kernel("static void count(unsigned int *s) { *s += 1u; }", "count"), elems, g(op2.INC))
    assert int(g.data[0]) == elems.size + 10


def case_direct_global_max(op2, split):
    """test_direct_loop.py:134-158 -- MAX: from 0 reduces to the Dat's max;
    from 2N (larger than every entry) stays 2N."""
    elems, x = _direct(op2, split)
    g = op2.Global(1, 0, np.uint32, "g")
    op2.par_loop(op2.Kernel("static void mx(unsigned int *m, const unsigned int *v) { if (*m < *v) *m = *v; }",
                            "mx"), elems, g(op2.MAX), x(op2.READ))
    assert int(g.data[0]) == (int(np.asarray(x.data).max()) if elems.size else 0)
    h = op2.Global(1, 2 * N, np.uint32, "h")
    op2.par_loop(op2.Kernel("static void mx2(const unsigned int *v, unsigned int *m) { if (*m < *v) *m = *v; }",
                            "mx2"), elems, x(op2.READ), h(op2.MAX))
    assert int(h.data[0]) == 2 * N


def case_direct_global_min(op2, split):
    """test_direct_loop.py:160-185 -- MIN: from 1000 reduces to the Dat's min;
    from 10 over values 11 stays 10."""
    elems, x = _direct(op2, split)
    g = op2.Global(1, 1000, np.uint32, "g")
    op2.par_loop(op2.Kernel("static void mn(unsigned int *m, const unsigned int *v) { if (*m > *v) *m = *v; }",
                            "mn"), elems, g(op2.MIN), x(op2.READ))
    assert int(g.data[0]) == (int(np.asarray(x.data).min()) if elems.size else 1000)
    x.data[:] = 11
    h = op2.Global(1, 10, np.uint32, "h")
    op2.par_loop(op2.Kernel("static void mn2(const unsigned int *v, unsigned int *m) { if (*m > *v) *m = *v; }",
                            "mn2"), elems, x(op2.READ), h(op2.MIN))
    assert int(h.data[0]) == 10


def case_direct_global_read(op2, split):
    """test_direct_loop.py:187-196 -- add a Global 1 to every entry."""
    elems, x = _direct(op2, split)
    h = op2.Global(1, 1, np.uint32, "h")
    op2.par_loop(op2.Kernel("static void addg(unsigned int *v, const unsigned int *c) { *v += *c; }", "addg"),
                 elems, x(op2.RW), h(op2.READ))
    n = elems.size
    assert int(np.asarray(x.data_ro, dtype=np.int64).sum()) == n * (n + 1) // 2


def case_direct_2d(op2, split):
    """test_direct_loop.py:198-204 -- WRITE both components of a dim-2 Dat."""
    elems, y = _direct(op2, split, dim=2)
    op2.par_loop(op2.Kernel("static void two(unsigned int *v) { v[0] = 42u; v[1] = 43u; }", "two"),
                 elems, y(op2.WRITE))
    d = np.asarray(y.data)
    assert (d[:, 0] == 42).all() and (d[:, 1] == 43).all()


def case_direct_host_write(op2, split):
    """test_direct_loop.py:206-223 -- host writes between loops are seen:
    INC of ones gives n, then of twos 2n (argument order swapped)."""
    elems, x = _direct(op2, split)
    g = op2.Global(1, 0, np.uint32, "g")
    x.data[:] = 1
    g.data[:] = 0
    op2.par_loop(op2.Kernel("static void s1(unsigned int *s, const unsigned int *v) { *s += *v; }", "s1"),
                 elems, g(op2.INC), x(op2.READ))
    assert int(g.data[0]) == elems.size
    x.data[:] = 2
    g.data[:] = 0
    op2.par_loop(op2.Kernel("static void s2(const unsigned int *v, unsigned int *s) { *s += *v; }", "s2"),
                 elems, x(op2.READ), g(op2.INC))
    assert int(g.data[0]) == 2 * elems.size


DIRECT_CASES = [case_direct_wo, case_direct_rw, case_direct_global_inc, case_direct_global_inc_init,
                case_direct_global_max, case_direct_global_min, case_direct_global_read, case_direct_2d,
                case_direct_host_write]


# ---- Global reductions (test_global_reduction.py:44-470) ---------------------
def _typed(dtype):
    return {np.uint32: "unsigned int", np.int32: "int", np.float32: "float", np.float64: "double"}[dtype]


def case_reduce_typed(op2, split, dtype, op, start, fill, want):
    """test_global_reduction.py:172-258 -- MIN/MAX of a constant Dat against a
    Global initial value, per scalar type."""
    s = op2.Set(split, "set")
    d = op2.Dat(op2.DataSet(s, 1), [fill] * N, dtype, "d")
    t = _typed(dtype)
    cmp = "<" if op == "MIN" else ">"
    k = op2.Kernel(f"static void red({t} *g, const {t} *x) {{ if (*x {cmp} *g) *g = *x; }}", "red")
    g = op2.Global(1, start, dtype, "g")
    op2.par_loop(k, s, g(getattr(op2, op)), d(op2.READ))
    assert np.asarray(g.data)[0] == dtype(want if s.size else start)


TYPED_REDUCTIONS = [
    ("min_uint32", np.uint32, "MIN", 8, 12, 8),
    ("min_int32", np.int32, "MIN", 8, -12, -12),
    ("max_int32", np.int32, "MAX", -42, -12, -12),
    ("min_float", np.float32, "MIN", -0.8, -12.0, -12.0),
    ("max_float", np.float32, "MAX", -42.8, -12.0, -12.0),
    ("min_double", np.float64, "MIN", -0.8, -12.0, -12.0),
    ("max_double", np.float64, "MAX", -42.8, -12.0, -12.0),
]


def _seq(op2, split, dim):
    s = op2.Set(split, "set")
    d = op2.Dat(op2.DataSet(s, dim), np.arange(dim * N, dtype=np.uint32) + 1, np.uint32)
    return s, d


K1 = {"INC": "static void r1(unsigned int *g, const unsigned int *x) { *g += *x; }",
      "MIN": "static void r1(unsigned int *g, const unsigned int *x) { if (*x < *g) *g = *x; }",
      "MAX": "static void r1(unsigned int *g, const unsigned int *x) { if (*x > *g) *g = *x; }"}
K2 = {"INC": "static void r2(unsigned int *g, const unsigned int *x) { g[0] += x[0]; g[1] += x[1]; }",
      "MIN": "static void r2(unsigned int *g, const unsigned int *x)"
             " { for (int c = 0; c < 2; ++c) if (x[c] < g[c]) g[c] = x[c]; }",
      "MAX": "static void r2(unsigned int *g, const unsigned int *x)"
             " { for (int c = 0; c < 2; ++c) if (x[c] > g[c]) g[c] = x[c]; }"}


def case_reduce_1d(op2, split):
    """test_global_reduction.py:260-336 -- READ (with and without initial
    data), INC, and MIN/MAX with the Dat or the Global holding the extreme."""
    s, d = _seq(op2, split, 1)
    n = s.size
    own = np.asarray(d.data).astype(np.int64)
    g = op2.Global(1, 1, dtype=np.uint32)
    op2.par_loop(op2.Kernel("static void w(unsigned int *x, const unsigned int *g) { *x = *g; }", "w"),
                 s, d(op2.WRITE), g(op2.READ))
    assert (np.asarray(d.data) == 1).all()
    g0 = op2.Global(1, dtype=np.uint32)
    d.data[:] = 100
    op2.par_loop(op2.Kernel("static void w(unsigned int *x, const unsigned int *g) { *x = *g; }", "w"),
                 s, d(op2.WRITE), g0(op2.READ))
    assert (np.asarray(g0.data) == 0).all() and (np.asarray(d.data) == 0).all()
    d.data[:] = own
    for start in (0, None):
        gi = op2.Global(1, start, dtype=np.uint32)
        op2.par_loop(op2.Kernel(K1["INC"], "r1"), s, gi(op2.INC), d(op2.READ))
        assert int(gi.data[0]) == own.sum()
    if n:
        for op, start, want in (("MIN", own.min() + 1, own.min()), ("MAX", own.max() - 1, own.max()),
                                ("MIN", own.min() - 1, own.min() - 1), ("MAX", own.max() + 1, own.max() + 1)):
            gm = op2.Global(1, start, dtype=np.uint32)
            op2.par_loop(op2.Kernel(K1[op], "r1"), s, gm(getattr(op2, op)), d(op2.READ))
            assert int(gm.data[0]) == want, (op, start)


def case_reduce_2d(op2, split):
    """test_global_reduction.py:270-277, 338-397 -- dim-2 READ, INC, MIN, MAX."""
    s, d = _seq(op2, split, 2)
    own = np.asarray(d.data).astype(np.int64)
    x = op2.Dat(op2.DataSet(s, 1), np.zeros(N, dtype=np.uint32), np.uint32)
    g = op2.Global(2, (1, 2), dtype=np.uint32)
    op2.par_loop(op2.Kernel("static void w2(unsigned int *x, const unsigned int *g) { *x = g[0] + g[1]; }", "w2"),
                 s, x(op2.WRITE), g(op2.READ))
    assert (np.asarray(x.data) == 3).all()
    gi = op2.Global(2, (0, 0), dtype=np.uint32)
    op2.par_loop(op2.Kernel(K2["INC"], "r2"), s, gi(op2.INC), d(op2.READ))
    assert list(np.asarray(gi.data).astype(np.int64)) == list(own.sum(axis=0))
    if s.size:
        for op, delta, red in (("MIN", 1, own.min(axis=0)), ("MAX", -1, own.max(axis=0))):
            gm = op2.Global(2, tuple(int(v) + delta for v in red), dtype=np.uint32)
            op2.par_loop(op2.Kernel(K2[op], "r2"), s, gm(getattr(op2, op)), d(op2.READ))
            assert list(np.asarray(gm.data).astype(np.int64)) == list(red), op
        for op, delta, red in (("MIN", -1, own.min(axis=0)), ("MAX", 1, own.max(axis=0))):
            start = tuple(int(v) + delta for v in red)
            gm = op2.Global(2, start, dtype=np.uint32)
            op2.par_loop(op2.Kernel(K2[op], "r2"), s, gm(getattr(op2, op)), d(op2.READ))
            assert tuple(int(v) for v in np.asarray(gm.data)) == start, op


def case_reduce_repeated(op2, split):
    """test_global_reduction.py:399-470 -- INC twice accumulates; setting the
    Global in between is seen; two Globals of different types; a ParLoop
    object computed repeatedly, with zero() in between."""
    s, d = _seq(op2, split, 1)
    n = s.size
    total = int(np.asarray(d.data).astype(np.int64).sum())
    g = op2.Global(1, 0, dtype=np.uint32)
    k = op2.Kernel(K1["INC"], "r1")
    op2.par_loop(k, s, g(op2.INC), d(op2.READ))
    op2.par_loop(k, s, g(op2.INC), d(op2.READ))
    assert int(g.data[0]) == 2 * total
    g.data = 10
    op2.par_loop(k, s, g(op2.INC), d(op2.READ))
    assert int(g.data[0]) == total + 10
    gu = op2.Global(1, [0], np.uint32, "gu")
    gd = op2.Global(1, [0.0], np.float64, "gd")
    op2.par_loop(op2.Kernel("static void both(unsigned int *i, double *x) { *i += 1u; *x += 1.0; }", "both"),
                 s, gu(op2.INC), gd(op2.INC))
    assert int(gu.data[0]) == n and float(gd.data[0]) == float(n)
    c = op2.Global(1, 0, dtype=np.uint32)
    loop = op2.ParLoop(op2.Kernel("static void one(unsigned int *g) { *g += 1u; }", "one"), s, c(op2.INC))
    loop.compute()
    loop.compute()
    assert int(c.data[0]) == 2 * n
    c.zero()
    loop.compute()
    assert int(c.data[0]) == n


REDUCTION_CASES = [case_reduce_1d, case_reduce_2d, case_reduce_repeated]


# ---- subsets (test_subset.py:42-258) -----------------------------------------
NS = 32
SUBSET_SPLITS = [(NS, NS, NS), (0, NS, NS), (NS // 2, NS, NS)]     # test_subset.py:43-48
INC1 = "static void inc1(unsigned int *v) { *v += 1u; }"


def case_subset_direct(op2, split):
    """test_subset.py:56-74 -- RW +1 over the even elements touches exactly
    them; over an empty subset touches nothing."""
    it = op2.Set(split, "iterset")
    even = np.arange(0, NS, 2, dtype=np.int32)
    d = op2.Dat(it ** 1, data=None, dtype=np.uint32)
    op2.par_loop(op2.Kernel(INC1, "inc1"), op2.Subset(it, even), d(op2.RW))
    assert list(np.nonzero(np.asarray(d.data))[0]) == list(even)
    e = op2.Dat(it ** 1, data=None, dtype=np.uint32)
    op2.par_loop(op2.Kernel(INC1, "inc1"), op2.Subset(it, []), e(op2.RW))
    assert not np.asarray(e.data).any()


def case_subset_complementary(op2, split):
    """test_subset.py:76-102, 184-207 -- even then odd subsets (built both
    ways) cover every element once, direct and through a map."""
    it = op2.Set(split, "iterset")
    even, odd = np.arange(0, NS, 2, dtype=np.int32), np.arange(1, NS, 2, dtype=np.int32)
    for sse, sso in ((op2.Subset(it, even), op2.Subset(it, odd)), (it(even), it(odd))):
        d = op2.Dat(it ** 1, data=None, dtype=np.uint32)
        op2.par_loop(op2.Kernel(INC1, "inc1"), sse, d(op2.RW))
        op2.par_loop(op2.Kernel(INC1, "inc1"), sso, d(op2.RW))
        assert (np.asarray(d.data) == 1).all()
    ind = op2.Set(NS, "indset")
    m = op2.Map(it, ind, 1, np.arange(NS, dtype=np.int32))
    d1 = op2.Dat(it ** 1, data=None, dtype=np.uint32)
    d2 = op2.Dat(ind ** 1, data=None, dtype=np.uint32)
    k = op2.Kernel("static void both(unsigned int *a, unsigned int *b) { *a += 1u; *b += 1u; }", "both")
    op2.par_loop(k, op2.Subset(it, even), d1(op2.RW), d2(op2.INC, m))
    op2.par_loop(k, op2.Subset(it, odd), d1(op2.RW), d2(op2.INC, m))
    assert int(np.asarray(d1.data).sum()) == NS and int(np.asarray(d2.data).sum()) == NS


def case_subset_nested(op2, split):
    """test_subset.py:104-136 -- a subset of a subset (every other of the
    evens) equals the subset of multiples of 4, built both ways."""
    it = op2.Set(split, "iterset")
    k = op2.Kernel(INC1, "inc1")
    ref = op2.Dat(it ** 1, data=None, dtype=np.uint32)
    op2.par_loop(k, op2.Subset(it, np.arange(0, NS, 4, dtype=np.int32)), ref(op2.RW))
    inner = np.arange(0, NS // 2, 2, dtype=np.int32)
    for sss in (op2.Subset(op2.Subset(it, np.arange(0, NS, 2, dtype=np.int32)), inner),
                it(np.arange(0, NS, 2, dtype=np.int32))(inner)):
        d = op2.Dat(it ** 1, data=None, dtype=np.uint32)
        op2.par_loop(k, sss, d(op2.RW))
        assert (np.asarray(d.data) == np.asarray(ref.data)).all()


def case_subset_indirect(op2, split):
    """test_subset.py:138-182 -- INC through a map over the evens: half the
    elements hit target 0; empty subset leaves zeros; a direct READ Dat is
    read at the subset's elements."""
    it = op2.Set(split, "iterset")
    even = np.arange(0, NS, 2, dtype=np.int32)
    ind = op2.Set(2, "indset")
    m = op2.Map(it, ind, 1, (np.arange(NS) % 2).astype(np.int32))
    d = op2.Dat(ind ** 1, data=None, dtype=np.uint32)
    op2.par_loop(op2.Kernel(INC1, "inc1"), op2.Subset(it, even), d(op2.INC, m))
    assert int(d.data[0]) == NS // 2
    e = op2.Dat(ind ** 1, data=None, dtype=np.uint32)
    op2.par_loop(op2.Kernel(INC1, "inc1"), op2.Subset(it, []), e(op2.INC, m))
    assert not np.asarray(e.data).any()
    vals = np.full(NS, 2976579765, dtype=np.uint32)      # odd entries must never be read
    vals[::2] = np.arange(NS)[::2] // 2
    src = op2.Dat(it ** 1, data=vals, dtype=np.uint32)
    dst = op2.Dat(ind ** 1, data=None, dtype=np.uint32)
    op2.par_loop(op2.Kernel("static void add(unsigned int *d, const unsigned int *s) { *d += *s; }", "add"),
                 op2.Subset(it, even), dst(op2.INC, m), src(op2.READ))
    assert int(dst.data[0]) == int(vals[::2].astype(np.int64).sum())


def case_subset_matrix(op2, split=None):
    """test_subset.py:209-252 -- Mat INC over the Set equals INC over the
    subsets [0, 1] and [1, 0] (bitwise: the same additions)."""
    it, idset, ind = op2.Set(2), op2.Set(2), op2.Set(4)
    dat = op2.Dat(idset ** 1, data=[0, 1], dtype=np.float64)
    m = op2.Map(it, ind, 4, [0, 1, 2, 3, 0, 1, 2, 3])
    idm = op2.Map(it, idset, 1, [0, 1])
    sp = op2.Sparsity((ind ** 1, ind ** 1), [(m, m, None)])
    k = op2.Kernel("static void uid(double a[4][4], const double *d)"
                   " { for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) a[i][j] += d[0] * 16 + i * 4 + j; }",
                   "uid")
    out = []
    for iters in (it, op2.Subset(it, [0, 1]), op2.Subset(it, [1, 0])):
        A = op2.Mat(sp, np.float64)
        A.zero()
        op2.par_loop(k, iters, A(op2.INC, (m, m)), dat(op2.READ, idm))
        A.assemble()
        out.append(np.asarray(A.values))
    want = np.array([[16.0 + 2 * (4 * i + j) for j in range(4)] for i in range(4)])
    assert (out[0] == want).all() and (out[1] == out[0]).all() and (out[2] == out[0]).all()


SUBSET_CASES = [case_subset_direct, case_subset_complementary, case_subset_nested, case_subset_indirect]


# ---- mixed Dats and Mats (test_matrices.py:502-531, 864-981) ----------------
# off-diagonal and lower-right blocks of the mixed mass-like matrix
# (test_matrices.py:869-876)
MIXED_OD = np.array([[1.0, 2.0, 0.0, 0.0], [0.0, 4.0, 6.0, 0.0], [0.0, 0.0, 9.0, 12.0]])
MIXED_LL = np.diag([1.0, 8.0, 18.0, 16.0]) + np.diag([2.0, 6.0, 12.0], -1) + np.diag([2.0, 6.0, 12.0], 1)


def _mixed(op2, dim=1):
    """test_matrices.py:502-524: 3 elements, 4 nodes on a line; element e
    maps to itself and to nodes (e, e+1); data 1..n on each component."""
    mset = op2.MixedSet((op2.Set(3), op2.Set(4)))
    elem, node = mset
    mmap = op2.MixedMap((op2.Map(elem, elem, 1, [0, 1, 2]), op2.Map(elem, node, 2, [0, 1, 1, 2, 2, 3])))
    if dim == 1:
        mdat = op2.MixedDat(op2.Dat(s, np.arange(1, s.size + 1, dtype=np.float64)) for s in mset)
    else:
        mdat = op2.MixedDat(op2.Dat(s ** 2, np.repeat(np.arange(1, s.size + 1, dtype=np.float64), 2).reshape(-1, 2))
                            for s in mset)
    return mset, mmap, mdat


def case_mixed_mat(op2, split=None):
    """test_matrices.py:878-916: INC of d d^T into a Mat on the 2x2 mixed
    sparsity; each block against the reference's closed form."""
    mset, mmap, mdat = _mixed(op2)
    sp = op2.Sparsity((mset ** 1, mset ** 1), {(i, j): [(rm, cm, None)] for i, rm in enumerate(mmap)
                                               for j, cm in enumerate(mmap)})
    mat = op2.Mat(sp)
    k = op2.Kernel("static void outer3(double v[9], const double d[3])"
                   " { for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) v[i * 3 + j] += d[i] * d[j]; }",
                   "outer3")
    op2.par_loop(k, mmap.iterset, mat(op2.INC, (mmap, mmap)), mdat(op2.READ, mmap))
    mat.assemble()
    assert np.allclose(mat[0, 0].values, np.diag([1.0, 4.0, 9.0]), rtol=1e-12, atol=0)
    assert np.allclose(mat[0, 1].values, MIXED_OD, rtol=1e-12, atol=0)
    assert np.allclose(mat[1, 0].values, MIXED_OD.T, rtol=1e-12, atol=0)
    assert np.allclose(mat[1, 1].values, MIXED_LL, rtol=1e-12, atol=0)
    assert np.allclose(mat.values, np.block([[np.diag([1.0, 4.0, 9.0]), MIXED_OD], [MIXED_OD.T, MIXED_LL]]))


def case_mixed_rhs(op2, split=None):
    """test_matrices.py:895-922: INC of the element's packs into a MixedDat."""
    mset, mmap, mdat = _mixed(op2)
    dat = op2.MixedDat(mset)
    k = op2.Kernel("static void add3(double v[3], const double d[3]) { for (int i = 0; i < 3; ++i) v[i] += d[i]; }",
                   "add3")
    op2.par_loop(k, mmap.iterset, dat(op2.INC, mmap), mdat(op2.READ, mmap))
    assert np.allclose(np.asarray(dat[0].data_ro), [1.0, 2.0, 3.0], rtol=1e-12, atol=0)
    assert np.allclose(np.asarray(dat[1].data_ro), [1.0, 4.0, 6.0, 4.0], rtol=1e-12, atol=0)


def case_mixed_rhs_vector(op2, split=None):
    """test_matrices.py:924-942: the same on dim-2 components."""
    mset, mmap, mvdat = _mixed(op2, dim=2)
    dat = op2.MixedDat(mset ** 2)
    k = op2.Kernel("static void add3v(double v[6], const double d[3][2])"
                   " { for (int i = 0; i < 3; ++i) { v[i * 2] += d[i][0]; v[i * 2 + 1] += d[i][1]; } }", "add3v")
    op2.par_loop(k, mmap.iterset, dat(op2.INC, mmap), mvdat(op2.READ, mmap))
    assert np.allclose(np.asarray(dat[0].data_ro), np.kron(np.array([[1.0], [2.0], [3.0]]), np.ones(2)), rtol=1e-12)
    assert np.allclose(np.asarray(dat[1].data_ro), np.kron(np.array([[1.0], [4.0], [6.0], [4.0]]), np.ones(2)),
                       rtol=1e-12)


def case_mixed_rw(op2, split=None):
    """RW through a MixedMap (coloured launch): every element adds 1 to its
    own entry and to each of its two nodes; nodes shared by two elements
    get 2.  Closed form, in the style of test_indirect_loop.py:138-143."""
    mset, mmap, _ = _mixed(op2)
    dat = op2.MixedDat(mset)
    k = op2.Kernel("static void bump3(double v[3]) { for (int i = 0; i < 3; ++i) v[i] = v[i] + 1.0; }", "bump3")
    op2.par_loop(k, mmap.iterset, dat(op2.RW, mmap))
    assert list(np.asarray(dat[0].data_ro)) == [1.0, 1.0, 1.0]
    assert list(np.asarray(dat[1].data_ro)) == [1.0, 2.0, 2.0, 1.0]


def case_mixed_blockwise_sparsity(op2, split=None):
    """test_matrices.py:945-981: a mixed sparsity given block by block, with
    ComposedMaps (one with a -1 entry, which adds nothing): per-block nnz."""
    setA, nodesetA, setB, nodesetB = op2.Set(3), op2.Set(4), op2.Set(2), op2.Set(3)
    nodesetAB = op2.MixedSet((nodesetA, nodesetB))
    dsAB = nodesetAB ** 1
    mapA = op2.Map(setA, nodesetA, 2, values=[[0, 1], [1, 2], [2, 3]])
    mapB = op2.Map(setB, nodesetB, 2, values=[[0, 1], [1, 2]])
    mapBA = op2.Map(setB, setA, 1, values=[1, 2])
    mapAB = op2.Map(setA, setB, 1, values=[-1, 0, 1])
    s = op2.Sparsity((dsAB, dsAB), {(1, 0): [(mapB, op2.ComposedMap(mapA, mapBA), None)],
                                    (0, 1): [(mapA, op2.ComposedMap(mapB, mapAB), None)]})
    assert list(s._blocks[0][0].nnz) == [1, 1, 1, 1]
    assert list(s._blocks[0][1].nnz) == [0, 2, 3, 2]
    assert list(s._blocks[1][0].nnz) == [2, 3, 2]
    assert list(s._blocks[1][1].nnz) == [1, 1, 1]


def case_mixed_mat_monolithic(op2, split=None):
    """test_matrices.py:878-916 on a monolithic Sparsity (nest=False): one
    matrix over the mixed dof space, whose blocks hold the same closed forms
    (the reference's _init_monolithic, mat.py:613-654)."""
    mset, mmap, mdat = _mixed(op2)
    sp = op2.Sparsity((mset ** 1, mset ** 1), {(i, j): [(rm, cm, None)] for i, rm in enumerate(mmap)
                                               for j, cm in enumerate(mmap)}, nest=False)
    assert not sp.nested and sp.shape == (2, 2)
    mat = op2.Mat(sp)
    assert (mat.nrows, mat.ncols) == (7, 7)
    k = op2.Kernel("static void outer3(double v[9], const double d[3])"
                   " { for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) v[i * 3 + j] += d[i] * d[j]; }",
                   "outer3")
    op2.par_loop(k, mmap.iterset, mat(op2.INC, (mmap, mmap)), mdat(op2.READ, mmap))
    mat.assemble()
    want = np.block([[np.diag([1.0, 4.0, 9.0]), MIXED_OD], [MIXED_OD.T, MIXED_LL]])
    assert np.allclose(mat.values, want, rtol=1e-12, atol=0)
    assert np.allclose(mat[0, 1].values, MIXED_OD, rtol=1e-12, atol=0)
    assert np.allclose(mat[1, 1].values, MIXED_LL, rtol=1e-12, atol=0)


def case_monolithic_states(op2, split=None):
    """test_matrices.py:798-838 for the non-nested fixture: a block's
    insertion sets the whole matrix's state, assembling a block alone is an
    error, and INSERT / ADD mix block by block."""
    mset, mmap, _ = _mixed(op2)
    sp = op2.Sparsity((mset ** 1, mset ** 1), {(i, j): [(rm, cm, None)] for i, rm in enumerate(mmap)
                                               for j, cm in enumerate(mmap)}, nest=False)
    mat = op2.Mat(sp)
    assert mat.assembly_state is op2.Mat.ASSEMBLED
    mat[0, 0].set_local_diagonal_entries([0])
    assert mat[0, 0].assembly_state is op2.Mat.INSERT_VALUES
    assert mat.assembly_state is op2.Mat.INSERT_VALUES
    try:
        mat[0, 0].assemble()
    except RuntimeError:
        pass
    else:
        raise AssertionError("assembling a MatBlock must raise")
    mat.assemble()
    mat.zero()
    mat[0, 0].addto_values(0, 0, [1])
    assert mat.assembly_state is op2.Mat.ADD_VALUES
    mat[1, 1].addto_values(1, 1, [3])
    mat[1, 1].set_values(0, 0, [2])
    mat[0, 0].set_values(1, 1, [4])
    mat[1, 1].addto_values(0, 0, [3])
    mat.assemble()
    assert np.allclose(mat[0, 0].values, np.diag([1, 4, 0]))
    assert np.allclose(mat[1, 1].values, np.diag([5, 3, 0, 0]))
    assert np.allclose(mat[0, 1].values, 0)
    assert np.allclose(mat[1, 0].values, 0)


MONOLITHIC_CASES = [
    ("mixed_mat_monolithic", case_mixed_mat_monolithic),
    ("monolithic_states", case_monolithic_states),
]


MIXED_CASES = [
    ("mixed_mat", case_mixed_mat),
    ("mixed_rhs", case_mixed_rhs),
    ("mixed_rhs_vector", case_mixed_rhs_vector),
    ("mixed_rw", case_mixed_rw),
    ("mixed_blockwise_sparsity", case_mixed_blockwise_sparsity),
]


# ---- extruded sets (test_extrusion.py:40-449) --------------------------------
XN = 32                     # base elements of the extruded strip
XLAYERS = 11
XWEDGES = XLAYERS - 1
XNODES = XN + 2


def _strip(op2, layers=XLAYERS):
    """test_extrusion.py:80-271: a strip of XN triangles (element i has
    nodes i, i+1, i+2, at (k // 2, k % 2)) extruded into XLAYERS layers.
    Node k of layer l is dof k * XLAYERS + l of a 2-vector Dat; the cell
    field of element i and cell l is entry i * XWEDGES + l; both maps step
    by 1 per layer."""
    base = op2.Set(XN)
    elements = op2.ExtrudedSet(base, layers=layers)
    node_set = op2.Set(XNODES * XLAYERS)
    elem_set = op2.Set(XN * XWEDGES)
    cvals = np.array([[m * XLAYERS + d for m in (i, i + 1, i + 2) for d in (0, 1)] for i in range(XN)],
                     dtype=np.int32)
    coords_map = op2.Map(elements, node_set, 6, cvals, "elem_dofs", np.ones(6, dtype=np.int32))
    field_map = op2.Map(elements, elem_set, 1, (np.arange(XN) * XWEDGES).astype(np.int32), "elem_elem",
                        np.ones(1, dtype=np.int32))
    xy = np.array([[k // 2, k % 2] for k in range(XNODES) for _ in range(XLAYERS)], dtype=np.float64)
    coords = op2.Dat(node_set ** 2, xy, np.float64, "coords")
    field = op2.Dat(elem_set ** 1, np.ones(XN * XWEDGES), np.float64, "field")
    return base, elements, node_set, elem_set, coords_map, field_map, coords, field


def case_extruded_volume(op2, split=None):
    """test_extrusion.py:340-356: Global INC of 0.05 per cell (triangle area
    1/2, height 0.1, field 1): 16 over 32 x 10 cells."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2)
    g = op2.Global(1, 0.0, np.float64, "g")
    k = op2.Kernel("""
static void vol(double A[1], const double x[6][2], const double y[1])
{
  double a = x[0][0] * (x[2][1] - x[4][1]) + x[2][0] * (x[4][1] - x[0][1]) + x[4][0] * (x[0][1] - x[2][1]);
  if (a < 0) a = -a;
  A[0] += 0.5 * a * 0.1 * y[0];
}""", "vol")
    op2.par_loop(k, elements, g(op2.INC), coords(op2.READ, cmap), field(op2.READ, fmap))
    assert abs(float(g.data[0]) - XWEDGES * 0.1 * (XN // 2)) < 1e-9


def case_extruded_direct_inc(op2, split=None):
    """test_extrusion.py:362-369: a direct Dat of the base set is incremented
    once per cell of its column: 9 cells of 10 layers."""
    base = op2.Set(XN)
    dat = op2.Dat(base ** 1, np.zeros(XN), np.float64)
    op2.par_loop(op2.Kernel("static void one(double *x) { *x += 1.0; }", "one"),
                 op2.ExtrudedSet(base, layers=10), dat(op2.INC))
    assert np.allclose(np.asarray(dat.data), 9.0)


def case_extruded_layer_arg(op2, split=None):
    """test_extrusion.py:371-387: the layer is passed to the kernel; cell l of
    column n writes l into entry n * XWEDGES + l."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2)
    f = op2.Dat(elem_set ** 1, -np.ones(XN * XWEDGES), np.float64)
    op2.par_loop(op2.Kernel("static void lay(double *x, int layer) { x[0] = layer; }", "lay"),
                 elements, f(op2.WRITE, fmap), pass_layer_arg=True)
    assert list(np.asarray(f.data)) == list(np.tile(np.arange(XWEDGES, dtype=np.float64), XN))


def case_extruded_write(op2, split=None):
    """test_extrusion.py:389-410: WRITE through the field and the coordinate
    maps reaches every cell entry and every node of every layer."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2)
    f = op2.Dat(elem_set ** 1, -np.ones(XN * XWEDGES), np.float64)
    op2.par_loop(op2.Kernel("static void wo(double *x) { x[0] = 42.0; }", "wo"), elements, f(op2.WRITE, fmap))
    assert (np.asarray(f.data) == 42).all()
    c = op2.Dat(node_set ** 2, np.zeros((XNODES * XLAYERS, 2)), np.float64)
    op2.par_loop(op2.Kernel("static void wo6(double x[6][2]) { for (int i = 0; i < 6; ++i) x[i][0] = x[i][1] = 42.0; }",
                            "wo6"), elements, c(op2.WRITE, cmap))
    assert (np.asarray(c.data) == 42).all()


def case_extruded_neighbours(op2, split=None):
    """test_extrusion.py:412-426: cell l of element i writes the sum of its
    six coordinate pairs: 2 * ((i + i + 1 + i + 2) // 2 terms) -- computed
    here from the node positions (k // 2) + (k % 2) of nodes i, i+1, i+2,
    each counted twice (layers l and l + 1)."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2)
    f = op2.Dat(elem_set ** 1, -np.ones(XN * XWEDGES), np.float64)
    k = op2.Kernel("static void wtf(double *y, const double x[6][2])"
                   " { double s = 0.0; for (int i = 0; i < 6; ++i) s += x[i][0] + x[i][1]; y[0] = s; }", "wtf")
    op2.par_loop(k, elements, f(op2.WRITE, fmap), coords(op2.READ, cmap))
    want = [2.0 * sum(kk // 2 + kk % 2 for kk in (i, i + 1, i + 2)) for i in range(XN) for _ in range(XWEDGES)]
    assert list(np.asarray(f.data)) == want


def case_extruded_rw(op2, split=None):
    """test_extrusion.py:428-446: RW through the coordinate map sets every
    still-zero pair to 1 (cells in order within a column): every pair 1."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2)
    c = op2.Dat(node_set ** 2, np.zeros((XNODES * XLAYERS, 2)), np.float64)
    k = op2.Kernel("""
static void inc(double y[6][2], const double x[6][2])
{
  for (int i = 0; i < 6; i++)
    if (y[i][0] == 0) { y[i][0] += 1; y[i][1] += 1; }
}""", "inc")
    op2.par_loop(k, elements, c(op2.RW, cmap), coords(op2.READ, cmap))
    assert float(np.asarray(c.data).sum()) == XNODES * XLAYERS * 2


def case_extruded_variable_layers(op2, split=None):
    """Variable layers ([bottom, top) per column): direct INC counts each
    column's cells, and an INC through a map stepping by 1 per layer covers
    exactly the cells above each column's bottom (set.py:305-346)."""
    base = op2.Set(4)
    lay = np.array([[0, 3], [1, 5], [2, 3], [0, 6]], dtype=np.int32)
    ext = op2.ExtrudedSet(base, lay)
    cnt = op2.Dat(base ** 1, np.zeros(4), np.float64)
    op2.par_loop(op2.Kernel("static void one(double *x) { *x += 1.0; }", "one"), ext, cnt(op2.INC))
    assert list(np.asarray(cnt.data)) == [2.0, 3.0, 0.0, 5.0]
    tgt = op2.Set(40)
    m = op2.Map(ext, tgt, 1, np.array([0, 10, 20, 30], dtype=np.int32), offset=np.ones(1, dtype=np.int32))
    d = op2.Dat(tgt ** 1, np.zeros(40), np.float64)
    op2.par_loop(op2.Kernel("static void tag(double *x, int layer) { x[0] += 100.0 + layer; }", "tag"),
                 ext, d(op2.INC, m), pass_layer_arg=True)
    want = np.zeros(40)
    for n, (b, t) in enumerate(lay):
        for lyr in range(t - b - 1):
            want[10 * n + lyr] += 100.0 + b + lyr
    assert list(np.asarray(d.data)) == list(want)


def case_extruded_matrix(op2, split=None):
    """An extruded Mat: the 6x6 coordinate-dof block per cell, INC of
    (1 + i) * (1 + j) into the extruded sparsity; each pattern entry holds
    the sum over the cells that couple its two dofs (counted here from the
    cell lists)."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2, layers=4)
    sp = op2.Sparsity((node_set ** 1, node_set ** 1), [(cmap, cmap, None)])
    A = op2.Mat(sp, np.float64)
    k = op2.Kernel("static void blk(double a[6][6])"
                   " { for (int i = 0; i < 6; ++i) for (int j = 0; j < 6; ++j) a[i][j] += (1 + i) * (1 + j); }",
                   "blk")
    op2.par_loop(k, elements, A(op2.INC, (cmap, cmap)))
    A.assemble()
    want = np.zeros((XNODES * XLAYERS, XNODES * XLAYERS))
    cv = np.array([[m * XLAYERS + d for m in (i, i + 1, i + 2) for d in (0, 1)] for i in range(XN)])
    for i in range(XN):
        for lyr in range(3):
            dofs = cv[i] + lyr
            for a in range(6):
                for b in range(6):
                    want[dofs[a], dofs[b]] += (1 + a) * (1 + b)
    assert np.array_equal(np.asarray(A.values), want)
    assert int(np.sum(sp.nnz)) == int(np.count_nonzero(want + np.eye(len(want))))


def case_extruded_bottom_top(op2, split=None):
    """Iteration regions (builder.py:771-795).  An INC of 1 through the
    6-entry dof map over the bottom cells reaches layers 0 and 1 of every
    node, over the top cells layers XLAYERS - 2 and XLAYERS - 1, each as
    often as base elements share the node (elements k - 2 .. k that exist);
    a direct INC counts one cell per column; a WRITE through the cell map on
    the top cells sets only the top cell of every column, and the layer
    argument there is XWEDGES - 1."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2)
    share = np.array([min(k, XN - 1) - max(k - 2, 0) + 1 for k in range(XNODES)], dtype=np.float64)
    one6 = op2.Kernel("static void one6(double x[6]) { for (int i = 0; i < 6; ++i) x[i] += 1.0; }", "one6")
    one = op2.Kernel("static void one(double *x) { *x += 1.0; }", "one")
    for region, lays in ((op2.ON_BOTTOM, (0, 1)), (op2.ON_TOP, (XLAYERS - 2, XLAYERS - 1))):
        cnt = op2.Dat(node_set ** 1, np.zeros(XNODES * XLAYERS), np.float64)
        op2.par_loop(one6, elements, cnt(op2.INC, cmap), iteration_region=region)
        want = np.zeros((XNODES, XLAYERS))
        for lay in lays:
            want[:, lay] = share
        assert np.array_equal(np.asarray(cnt.data), want.reshape(-1)), region
        col = op2.Dat(base ** 1, np.zeros(XN), np.float64)
        op2.par_loop(one, elements, col(op2.INC), iteration_region=region)
        assert np.array_equal(np.asarray(col.data), np.ones(XN)), region
    f = op2.Dat(elem_set ** 1, -np.ones(XN * XWEDGES), np.float64)
    k = op2.Kernel("static void top(double *y, int layer) { y[0] = 100.0 + layer; }", "top")
    op2.par_loop(k, elements, f(op2.WRITE, fmap), iteration_region=op2.ON_TOP, pass_layer_arg=True)
    want = [100.0 + lyr if lyr == XWEDGES - 1 else -1.0 for _ in range(XN) for lyr in range(XWEDGES)]
    assert list(np.asarray(f.data)) == want


def case_extruded_interior_facets(op2, split=None):
    """ON_INTERIOR_FACETS (builder.py:81-129, 771-795): every map names the
    cell below, then the cell above.  INC of 1 below and 10 above into the
    cell field gives cell l 1 if a cell lies above it plus 10 if one lies
    below; through the dof map the upper cell's dof ids exceed the lower's
    by one layer each (6 per facet); reading the cell field c = id^2 the
    jumps c[above] - c[below] telescope to c[top] - c[bottom] per column; a
    direct INC counts the XWEDGES - 1 facets of a column, and the layer
    argument runs 0 .. XWEDGES - 2."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2)
    fac = op2.ON_INTERIOR_FACETS
    f = op2.Dat(elem_set ** 1, np.zeros(XN * XWEDGES), np.float64)
    k = op2.Kernel("static void jmp(double c[2]) { c[0] += 1.0; c[1] += 10.0; }", "jmp")
    op2.par_loop(k, elements, f(op2.INC, fmap), iteration_region=fac)
    want = [(1.0 if lyr < XWEDGES - 1 else 0.0) + (10.0 if lyr > 0 else 0.0)
            for _ in range(XN) for lyr in range(XWEDGES)]
    assert list(np.asarray(f.data)) == want
    ids = op2.Dat(node_set ** 1, np.arange(XNODES * XLAYERS, dtype=np.float64), np.float64)
    g = op2.Global(1, 0.0, np.float64, "g")
    k2 = op2.Kernel("static void up(double *g, const double x[12])"
                    " { for (int i = 0; i < 6; ++i) g[0] += x[6 + i] - x[i]; }", "up")
    op2.par_loop(k2, elements, g(op2.INC), ids(op2.READ, cmap), iteration_region=fac)
    assert float(np.asarray(g.data).reshape(-1)[0]) == 6.0 * XN * (XWEDGES - 1)
    sq = op2.Dat(elem_set ** 1, np.arange(XN * XWEDGES, dtype=np.float64) ** 2, np.float64)
    h = op2.Global(1, 0.0, np.float64, "h")
    k3 = op2.Kernel("static void dif(double *h, const double c[2]) { h[0] += c[1] - c[0]; }", "dif")
    op2.par_loop(k3, elements, h(op2.INC), sq(op2.READ, fmap), iteration_region=fac)
    tele = sum(float((i * XWEDGES + XWEDGES - 1) ** 2 - (i * XWEDGES) ** 2) for i in range(XN))
    assert float(np.asarray(h.data).reshape(-1)[0]) == tele
    col = op2.Dat(base ** 1, np.zeros(XN), np.float64)
    lsum = op2.Global(1, 0.0, np.float64, "lsum")
    k4 = op2.Kernel("static void cnt(double *x, double *s, int layer) { *x += 1.0; s[0] += layer; }", "cnt")
    op2.par_loop(k4, elements, col(op2.INC), lsum(op2.INC), iteration_region=fac, pass_layer_arg=True)
    assert np.array_equal(np.asarray(col.data), np.full(XN, XWEDGES - 1.0))
    assert float(np.asarray(lsum.data).reshape(-1)[0]) == XN * sum(range(XWEDGES - 1))


def case_extruded_facet_matrix(op2, split=None):
    """A cell-to-cell Mat over interior facets (sparsity.pyx:285-393): the
    jump matrix [[1, -1], [-1, 1]] per facet gives, per column, the path
    Laplacian of its XWEDGES cells (diagonal 1 at the ends, 2 inside, -1
    between neighbours); a second pair over the bottom and top cells adds 5
    to those cells' diagonal.  The pattern is the diagonal plus the two
    neighbours."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2)
    sp = op2.Sparsity((elem_set ** 1, elem_set ** 1),
                      [(fmap, fmap, (op2.ON_INTERIOR_FACETS,)), (fmap, fmap, (op2.ON_BOTTOM, op2.ON_TOP))])
    A = op2.Mat(sp, np.float64)
    jump = op2.Kernel("static void jump(double a[2][2])"
                      " { a[0][0] += 1.0; a[0][1] -= 1.0; a[1][0] -= 1.0; a[1][1] += 1.0; }", "jump")
    op2.par_loop(jump, elements, A(op2.INC, (fmap, fmap)), iteration_region=op2.ON_INTERIOR_FACETS)
    five = op2.Kernel("static void five(double a[1][1]) { a[0][0] += 5.0; }", "five")
    for region in (op2.ON_BOTTOM, op2.ON_TOP):
        op2.par_loop(five, elements, A(op2.INC, (fmap, fmap)), iteration_region=region)
    A.assemble()
    n = XN * XWEDGES
    want = np.zeros((n, n))
    for i in range(XN):
        for lyr in range(XWEDGES):
            c = i * XWEDGES + lyr
            want[c, c] = (lyr > 0) + (lyr < XWEDGES - 1) + 5.0 * (lyr == 0) + 5.0 * (lyr == XWEDGES - 1)
            if lyr < XWEDGES - 1:
                want[c, c + 1] = want[c + 1, c] = -1.0
    assert np.array_equal(np.asarray(A.values), want)
    assert int(np.sum(sp.nnz)) == int(np.count_nonzero(want))


def case_extruded_variable_regions(op2, split=None):
    """Regions over variable layers [[0, 3], [1, 5], [2, 3], [0, 6]] (2, 3, 0
    and 5 cells; builder.py:771-795): per column ON_BOTTOM runs one layer
    (also on the column with no cell: its bottom layer, as the reference's
    bounds give), ON_TOP one (none on the empty column), ON_INTERIOR_FACETS
    cells - 1, ALL cells; through a map stepping by 1 from 10 * column,
    interior facets add 1 to the cell below and 2 to the cell above."""
    base = op2.Set(4)
    lay = np.array([[0, 3], [1, 5], [2, 3], [0, 6]], dtype=np.int32)
    ext = op2.ExtrudedSet(base, lay)
    one = op2.Kernel("static void one(double *x) { *x += 1.0; }", "one")
    for region, want in ((op2.ON_BOTTOM, [1, 1, 1, 1]), (op2.ON_TOP, [1, 1, 0, 1]),
                         (op2.ON_INTERIOR_FACETS, [1, 2, 0, 4]), (op2.ALL, [2, 3, 0, 5])):
        cnt = op2.Dat(base ** 1, np.zeros(4), np.float64)
        op2.par_loop(one, ext, cnt(op2.INC), iteration_region=region)
        assert list(np.asarray(cnt.data)) == want, region
    tgt = op2.Set(40)
    m = op2.Map(ext, tgt, 1, np.array([0, 10, 20, 30], dtype=np.int32), offset=np.ones(1, dtype=np.int32))
    d = op2.Dat(tgt ** 1, np.zeros(40), np.float64)
    k = op2.Kernel("static void jmp2(double c[2]) { c[0] += 1.0; c[1] += 2.0; }", "jmp2")
    op2.par_loop(k, ext, d(op2.INC, m), iteration_region=op2.ON_INTERIOR_FACETS)
    want = np.zeros(40)
    for c, ncell in enumerate([2, 3, 0, 5]):
        for s in range(ncell - 1):
            want[10 * c + s] += 1.0
            want[10 * c + s + 1] += 2.0
    assert np.array_equal(np.asarray(d.data), want)


PCELLS = 4                   # cells of a periodic column
PCOLS = 3
PNV = PCOLS * PCELLS         # vertices (and cells) of the periodic columns
PPAD = 2 * PCELLS            # unused targets past them: a wrong wrap stays in bounds


def _ppad(a):
    """``a`` (vector or square matrix over PNV targets) padded with the
    zero targets past them."""
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        return np.concatenate([a, np.zeros(PPAD)])
    out = np.zeros((PNV + PPAD, PNV + PPAD))
    out[:PNV, :PNV] = a
    return out


def _periodic(op2):
    """PCOLS periodic columns of PCELLS cells (builder.py:103-121): vertex
    v of column c is c * PCELLS + v, the top vertex of the last cell is the
    bottom one (offset quotient 1 on the top entry); cell l of column c is
    c * PCELLS + l."""
    base = op2.Set(PCOLS)
    ext = op2.ExtrudedSet(base, PCELLS + 1, extruded_periodic=True)
    vset, cset = op2.Set(PNV + PPAD), op2.Set(PNV + PPAD)
    c0 = np.arange(PCOLS, dtype=np.int32) * PCELLS
    vmap = op2.Map(ext, vset, 2, np.stack([c0, c0 + 1], axis=1), "vmap", np.ones(2, dtype=np.int32),
                   offset_quotient=np.array([0, 1], dtype=np.int32))
    cmap = op2.Map(ext, cset, 1, c0.reshape(-1, 1), "cmap", np.ones(1, dtype=np.int32))
    return base, ext, vset, cset, vmap, cmap


def _cycle(n, diag, off):
    """The n x n matrix with ``diag`` on the diagonal and ``off`` between
    cyclic neighbours."""
    a = np.eye(n) * diag
    for v in range(n):
        a[v, (v + 1) % n] += off
        a[(v + 1) % n, v] += off
    return a


def case_periodic_vertices(op2, split=None):
    """Periodic extrusion, all cells: an INC of 1 into both vertices of every
    cell gives every vertex 2 (each has a cell below and above it, the
    bottom one through the wrap); the 2x2 mass-like matrix [[2, 1], [1, 2]]
    assembles, per column, the cyclic matrix with 4 on the diagonal and 1
    between cyclic neighbours; the top cells reach vertices PCELLS - 1 and 0."""
    base, ext, vset, cset, vmap, cmap = _periodic(op2)
    cnt = op2.Dat(vset ** 1, np.zeros(PNV + PPAD), np.float64)
    two = op2.Kernel("static void two(double x[2]) { x[0] += 1.0; x[1] += 1.0; }", "two")
    op2.par_loop(two, ext, cnt(op2.INC, vmap))
    assert np.array_equal(np.asarray(cnt.data), _ppad(np.full(PNV, 2.0)))
    A = op2.Mat(op2.Sparsity((vset ** 1, vset ** 1), [(vmap, vmap, None)]), np.float64)
    mk = op2.Kernel("static void mm(double a[2][2]) { a[0][0] += 2.0; a[0][1] += 1.0; a[1][0] += 1.0; a[1][1] += 2.0; }",
                    "mm")
    op2.par_loop(mk, ext, A(op2.INC, (vmap, vmap)))
    A.assemble()
    want = np.kron(np.eye(PCOLS), _cycle(PCELLS, 4.0, 1.0))
    assert np.array_equal(np.asarray(A.values), _ppad(want))
    top = op2.Dat(vset ** 1, np.zeros(PNV + PPAD), np.float64)
    op2.par_loop(two, ext, top(op2.INC, vmap), iteration_region=op2.ON_TOP)
    want = np.zeros((PCOLS, PCELLS))
    want[:, [0, PCELLS - 1]] = 1.0
    assert np.array_equal(np.asarray(top.data), _ppad(want.reshape(-1)))


def case_periodic_facets(op2, split=None):
    """Periodic interior facets: as many as cells (the last joins the top
    cell to the bottom one, builder.py:781-786).  INC of 1 below and 10
    above gives every cell 11; the facet jump matrix assembles the cycle
    Laplacian (2 on the diagonal, -1 between cyclic neighbours); through the
    vertex map the lower cell's top vertex is the upper cell's bottom one
    on every facet, and the upper cell's top vertex ids sum, per column, to
    PCELLS * c * PCELLS + (0 + 1 + ... + PCELLS - 1)."""
    base, ext, vset, cset, vmap, cmap = _periodic(op2)
    fac = op2.ON_INTERIOR_FACETS
    c = op2.Dat(cset ** 1, np.zeros(PNV + PPAD), np.float64)
    jmp = op2.Kernel("static void pj(double c[2]) { c[0] += 1.0; c[1] += 10.0; }", "pj")
    op2.par_loop(jmp, ext, c(op2.INC, cmap), iteration_region=fac)
    assert np.array_equal(np.asarray(c.data), _ppad(np.full(PNV, 11.0)))
    A = op2.Mat(op2.Sparsity((cset ** 1, cset ** 1), [(cmap, cmap, (fac,))]), np.float64)
    jm = op2.Kernel("static void pjm(double a[2][2]) { a[0][0] += 1.0; a[0][1] -= 1.0; a[1][0] -= 1.0; a[1][1] += 1.0; }",
                    "pjm")
    op2.par_loop(jm, ext, A(op2.INC, (cmap, cmap)), iteration_region=fac)
    A.assemble()
    assert np.array_equal(np.asarray(A.values), _ppad(np.kron(np.eye(PCOLS), _cycle(PCELLS, 2.0, -1.0))))
    ids = op2.Dat(vset ** 1, np.arange(PNV + PPAD, dtype=np.float64), np.float64)
    g = op2.Global(2, [0.0, 0.0], np.float64, "g")
    k = op2.Kernel("static void pv(double *g, const double x[4])"
                   " { g[0] += (x[1] - x[2]) * (x[1] - x[2]); g[1] += x[3]; }", "pv")
    op2.par_loop(k, ext, g(op2.INC), ids(op2.READ, vmap), iteration_region=fac)
    got = np.asarray(g.data).reshape(-1)
    assert got[0] == 0.0
    assert got[1] == sum(PCELLS * cc * PCELLS + sum(range(PCELLS)) for cc in range(PCOLS))
    cnt = op2.Dat(base ** 1, np.zeros(PCOLS), np.float64)
    op2.par_loop(op2.Kernel("static void one(double *x) { *x += 1.0; }", "one"), ext, cnt(op2.INC),
                 iteration_region=fac)
    assert np.array_equal(np.asarray(cnt.data), np.full(PCOLS, float(PCELLS)))


def case_periodic_variable_layers(op2, split=None):
    """Periodic columns of 2 and 3 cells ([[0, 3], [1, 5]]): vertex k of
    column c is 10 * c + k; every vertex gets 2 from an INC of 1 into both
    vertices of each cell, and each column has as many interior facets as
    cells."""
    base = op2.Set(2)
    ext = op2.ExtrudedSet(base, np.array([[0, 3], [1, 5]], dtype=np.int32), extruded_periodic=True)
    tgt = op2.Set(20)
    m = op2.Map(ext, tgt, 2, np.array([[0, 1], [10, 11]], dtype=np.int32), offset=np.ones(2, dtype=np.int32),
                offset_quotient=np.array([0, 1], dtype=np.int32))
    d = op2.Dat(tgt ** 1, np.zeros(20), np.float64)
    op2.par_loop(op2.Kernel("static void two(double x[2]) { x[0] += 1.0; x[1] += 1.0; }", "two"), ext, d(op2.INC, m))
    want = np.zeros(20)
    want[[0, 1, 10, 11, 12]] = 2.0
    assert np.array_equal(np.asarray(d.data), want)
    cnt = op2.Dat(base ** 1, np.zeros(2), np.float64)
    op2.par_loop(op2.Kernel("static void one(double *x) { *x += 1.0; }", "one"), ext, cnt(op2.INC),
                 iteration_region=op2.ON_INTERIOR_FACETS)
    assert list(np.asarray(cnt.data)) == [2.0, 3.0]


def case_extruded_permuted_offsets(op2, split=None):
    """A PermutedMap of an extruded map with different offsets per entry
    keeps each entry's own offset (builder.py:148-172): entries (0 step 1,
    100 step 2) permuted to (100, 0); INC of 1 into the first and 10 into
    the second over 4 cells reaches 100, 102, 104, 106 and 0, 1, 2, 3."""
    base = op2.Set(1)
    ext = op2.ExtrudedSet(base, 5)
    tgt = op2.Set(110)
    m = op2.Map(ext, tgt, 2, np.array([[0, 100]], dtype=np.int32), offset=np.array([1, 2], dtype=np.int32))
    pm = op2.PermutedMap(m, [1, 0])
    d = op2.Dat(tgt ** 1, np.zeros(110), np.float64)
    op2.par_loop(op2.Kernel("static void ab(double x[2]) { x[0] += 1.0; x[1] += 10.0; }", "ab"), ext, d(op2.INC, pm))
    want = np.zeros(110)
    want[[100, 102, 104, 106]] = 1.0
    want[[0, 1, 2, 3]] = 10.0
    assert np.array_equal(np.asarray(d.data), want)


def case_extruded_subset(op2, split=None):
    """A Subset of an ExtrudedSet runs the chosen columns only (set.py:394-428):
    an INC of 1 through the 6-entry dof map over columns {1, 5, 6} reaches
    each node of those columns' triangles, every layer, as often as chosen
    elements share it; a WRITE of the layer through the cell map sets the
    chosen columns' cells only; a direct INC counts their cells."""
    base, elements, node_set, elem_set, cmap, fmap, coords, field = _strip(op2)
    chosen = [1, 5, 6]
    sub = elements(chosen)
    cnt = op2.Dat(node_set ** 1, np.zeros(XNODES * XLAYERS), np.float64)
    one6 = op2.Kernel("static void one6(double x[6]) { for (int i = 0; i < 6; ++i) x[i] += 1.0; }", "one6")
    op2.par_loop(one6, sub, cnt(op2.INC, cmap))
    want = np.zeros((XNODES, XLAYERS))
    for i in chosen:
        for k in (i, i + 1, i + 2):
            want[k, 0] += 1.0
            want[k, 1:XLAYERS - 1] += 2.0
            want[k, XLAYERS - 1] += 1.0
    assert np.array_equal(np.asarray(cnt.data), want.reshape(-1))
    f = op2.Dat(elem_set ** 1, -np.ones(XN * XWEDGES), np.float64)
    k = op2.Kernel("static void lay(double *y, int layer) { y[0] = layer; }", "lay")
    op2.par_loop(k, sub, f(op2.WRITE, fmap), pass_layer_arg=True)
    want = -np.ones((XN, XWEDGES))
    want[chosen, :] = np.arange(XWEDGES)
    assert np.array_equal(np.asarray(f.data), want.reshape(-1))
    col = op2.Dat(base ** 1, np.zeros(XN), np.float64)
    op2.par_loop(op2.Kernel("static void one(double *x) { *x += 1.0; }", "one"), sub, col(op2.INC))
    want = np.zeros(XN)
    want[chosen] = XWEDGES
    assert np.array_equal(np.asarray(col.data), want)


EXTRUDED_CASES = [
    ("extruded_volume", case_extruded_volume),
    ("extruded_direct_inc", case_extruded_direct_inc),
    ("extruded_layer_arg", case_extruded_layer_arg),
    ("extruded_write", case_extruded_write),
    ("extruded_neighbours", case_extruded_neighbours),
    ("extruded_rw", case_extruded_rw),
    ("extruded_variable_layers", case_extruded_variable_layers),
    ("extruded_matrix", case_extruded_matrix),
    ("extruded_bottom_top", case_extruded_bottom_top),
    ("extruded_interior_facets", case_extruded_interior_facets),
    ("extruded_facet_matrix", case_extruded_facet_matrix),
    ("extruded_variable_regions", case_extruded_variable_regions),
    ("extruded_permuted_offsets", case_extruded_permuted_offsets),
    ("periodic_vertices", case_periodic_vertices),
    ("periodic_facets", case_periodic_facets),
    ("periodic_variable_layers", case_periodic_variable_layers),
    ("extruded_subset", case_extruded_subset),
]
