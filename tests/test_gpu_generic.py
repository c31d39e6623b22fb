"""Generated (runtime-compiled) loops for arbitrary C kernels, checked with the
reference's own known answers (``test/unit/test_indirect_loop.py``), over the
same core/owned split parametrisation (``:46-50``)."""
import numpy as np
import pytest

from pyop2_amd import op2
from pyop2_amd.exceptions import MapValueError

pytestmark = pytest.mark.gpu

nelems = 4096


@pytest.fixture(params=[(nelems, nelems, nelems), (0, nelems, nelems), (nelems // 2, nelems, nelems)])
def iterset(request):
    return op2.Set(request.param, "iterset")


@pytest.fixture
def indset():
    return op2.Set(nelems, "indset")


@pytest.fixture
def unitset():
    return op2.Set(1, "unitset")


@pytest.fixture
def x(indset):
    return op2.Dat(indset, list(range(nelems)), np.uint32, "x")


@pytest.fixture
def x2(indset):
    return op2.Dat(indset ** 2, np.array([list(range(nelems)), list(range(nelems))], dtype=np.uint32),
                   np.uint32, "x2")


@pytest.fixture
def mapd():
    return list(range(nelems))[::-1]


@pytest.fixture
def iterset2indset(iterset, indset, mapd):
    return op2.Map(iterset, indset, 1, np.array(mapd, dtype=np.uint32), "iterset2indset")


@pytest.fixture
def iterset2unitset(iterset, unitset):
    return op2.Map(iterset, unitset, 1, np.zeros(nelems, dtype=np.uint32), "iterset2unitset")


def test_onecolor_wo(device, iterset, x, iterset2indset):
    op2.par_loop(op2.Kernel("static void kernel_wo(unsigned int* x) { *x = 42; }\n", "kernel_wo"),
                 iterset, x(op2.WRITE, iterset2indset))
    assert all(v == 42 for v in x.data)


def test_onecolor_rw(device, iterset, x, iterset2indset):
    op2.par_loop(op2.Kernel("static void rw(unsigned int* x) { (*x) = (*x) + 1; }\n", "rw"),
                 iterset, x(op2.RW, iterset2indset))
    assert sum(x.data) == nelems * (nelems + 1) // 2


def test_indirect_inc(device, iterset, unitset, iterset2unitset):
    u = op2.Dat(unitset, np.array([0], dtype=np.uint32), np.uint32, "u")
    op2.par_loop(op2.Kernel("static void inc(unsigned int* x) { (*x) = (*x) + 1; }\n", "inc"),
                 iterset, u(op2.INC, iterset2unitset))
    assert u.data[0] == nelems


def test_indirect_max(device, iterset, indset, iterset2indset):
    a = op2.Dat(indset, dtype=np.int32)
    b = op2.Dat(indset, dtype=np.int32)
    a.data[:] = -10
    b.data[:] = -5
    op2.par_loop(op2.Kernel("static void maxify(int *a, int *b) {*a = *a < *b ? *b : *a;}\n", "maxify"),
                 iterset, a(op2.MAX, iterset2indset), b(op2.READ, iterset2indset))
    assert np.allclose(a.data_ro, -5)


def test_indirect_min(device, iterset, indset, iterset2indset):
    a = op2.Dat(indset, dtype=np.int32)
    b = op2.Dat(indset, dtype=np.int32)
    a.data[:] = 10
    b.data[:] = 5
    op2.par_loop(op2.Kernel("static void minify(int *a, int *b) {*a = *a > *b ? *b : *a;}\n", "minify"),
                 iterset, a(op2.MIN, iterset2indset), b(op2.READ, iterset2indset))
    assert np.allclose(a.data_ro, 5)


def test_global_read(device, iterset, x, iterset2indset):
    g = op2.Global(1, 2, np.uint32, "g")
    op2.par_loop(op2.Kernel("static void global_read(unsigned int* x, unsigned int* g) { (*x) /= (*g); }\n",
                            "global_read"), iterset, x(op2.RW, iterset2indset), g(op2.READ))
    assert sum(x.data) == sum(v // 2 for v in range(nelems))


def test_global_inc(device, iterset, x, iterset2indset):
    g = op2.Global(1, 0, np.uint32, "g")
    op2.par_loop(op2.Kernel("""
        static void global_inc(unsigned int *x, unsigned int *inc) {
          (*x) = (*x) + 1; (*inc) += (*x);
        }""", "global_inc"), iterset, x(op2.RW, iterset2indset), g(op2.INC))
    assert sum(x.data) == nelems * (nelems + 1) // 2
    assert g.data[0] == nelems * (nelems + 1) // 2


def test_2d_dat(device, iterset, iterset2indset, x2):
    op2.par_loop(op2.Kernel("static void wo(unsigned int* x) { x[0] = 42; x[1] = 43; }\n", "wo"),
                 iterset, x2(op2.WRITE, iterset2indset))
    assert all(all(v == [42, 43]) for v in x2.data)


def test_2d_map(device):
    nedges = nelems - 1
    nodes = op2.Set(nelems, "nodes")
    edges = op2.Set(nedges, "edges")
    node_vals = op2.Dat(nodes, np.arange(nelems, dtype=np.uint32), np.uint32, "node_vals")
    edge_vals = op2.Dat(edges, np.zeros(nedges, dtype=np.uint32), np.uint32, "edge_vals")
    e_map = np.array([(i, i + 1) for i in range(nedges)], dtype=np.uint32)
    edge2node = op2.Map(edges, nodes, 2, e_map, "edge2node")
    op2.par_loop(op2.Kernel("""
        static void sum(unsigned int *edge, unsigned int *nodes) {
          *edge = nodes[0] + nodes[1];
        }""", "sum"), edges, edge_vals(op2.WRITE), node_vals(op2.READ, edge2node))
    assert all(np.arange(1, nedges * 2 + 1, 2) == edge_vals.data)


def test_uninitialized_map(iterset, indset, x):
    with pytest.raises(MapValueError):
        op2.par_loop(op2.Kernel("static void wo(unsigned int* x) { *x = 42; }\n", "wo"), iterset,
                     x(op2.WRITE, op2.Map(iterset, indset, 1)))
