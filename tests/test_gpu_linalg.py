"""Dat linear algebra on the device against the reference's known answers
(``test/unit/test_linalg.py``): pointwise operators with Dats and scalars,
their in-place forms, the result type (the left operand's), copies from
unary plus / minus, shape mismatches, a par_loop result used in an
expression, and the norm."""
import numpy as np
import pytest

from pyop2_amd import op2

pytestmark = pytest.mark.gpu

N = 8


@pytest.fixture
def dset():
    return op2.Set(N) ** 1


def _x(dset, values=None):
    return op2.Dat(dset, values, np.float64, "x")


def _y(dset):
    return op2.Dat(dset, np.arange(1, N + 1), np.float64, "y")


def _yi(dset):
    return op2.Dat(dset, np.arange(1, N + 1), np.int64, "yi")


def _vals(d):
    return np.asarray(d.data_ro)


@pytest.mark.parametrize("op,want", [
    (lambda x, y: x + y, lambda y: 3 * y),            # test_linalg.py:86-88
    (lambda x, y: x - y, lambda y: y),                # :90-92
    (lambda x, y: x * y, lambda y: 2 * y * y),        # :94-96
    (lambda x, y: x / y, lambda y: np.full_like(y, 2.0)),  # :98-100
])
def test_binary_ops(device, dset, op, want):
    y = _y(dset)
    x = _x(dset, 2 * _vals(y))
    assert np.array_equal(_vals(op(x, y)), want(_vals(y)))


@pytest.mark.parametrize("expr,want", [
    (lambda y: y + 1.0, lambda v: v + 1.0),           # :118-120
    (lambda y: 1.0 + y, lambda v: 1.0 + v),           # :122-124
    (lambda y: y - 1.0, lambda v: v - 1.0),           # :136-138
    (lambda y: 1.0 - y, lambda v: 1.0 - v),           # :140-142
    (lambda y: y * 2.0, lambda v: 2.0 * v),           # :144-146
    (lambda y: 2.0 * y, lambda v: 2.0 * v),           # :148-150
    (lambda y: y / 2.0, lambda v: v / 2.0),           # :152-154
])
def test_scalar_ops(device, dset, expr, want):
    y = _y(dset)
    assert np.array_equal(_vals(expr(y)), want(_vals(y)))


def test_unary_ops_copy(device, dset):
    """test_linalg.py:126-134: +y and -y are new Dats."""
    y = _y(dset)
    z = +y
    assert z is not y and np.array_equal(_vals(z), _vals(y))
    z = -y
    assert z is not y and np.array_equal(_vals(z), -_vals(y))


@pytest.mark.parametrize("op", ["add", "sub", "mul", "div"])
def test_result_type_is_the_left_operand(device, dset, op):
    """test_linalg.py:156-186 and 255-285: y op yi is float64, yi op y is
    int64, in place too."""
    f = {"add": lambda a, b: a + b, "sub": lambda a, b: a - b,
         "mul": lambda a, b: a * b, "div": lambda a, b: a / b}[op]
    assert f(_y(dset), _yi(dset)).data_ro.dtype == np.float64
    assert f(_yi(dset), _y(dset)).data_ro.dtype == np.int64
    y, yi = _y(dset), _yi(dset)
    inplace = {"add": "__iadd__", "sub": "__isub__", "mul": "__imul__", "div": "__itruediv__"}[op]
    y = getattr(y, inplace)(_yi(dset))
    yi = getattr(yi, inplace)(_y(dset))
    assert y.data_ro.dtype == np.float64 and yi.data_ro.dtype == np.int64


def test_integer_in_place_truncates(device, dset):
    """yi /= 2.5 * y keeps the integer type: each entry is the quotient
    truncated toward zero, as the generated C assignment does."""
    yi = _yi(dset)
    y = _y(dset)
    yi /= y * 2.5
    assert np.array_equal(_vals(yi), np.zeros(N, dtype=np.int64))
    yi = _yi(dset)
    yi -= y * 1.5
    assert np.array_equal(_vals(yi), np.trunc(-0.5 * np.arange(1, N + 1)).astype(np.int64))


@pytest.mark.parametrize("op,want", [
    ("__iadd__", lambda v: 3 * v),                    # test_linalg.py:203-206
    ("__isub__", lambda v: v),                        # :208-211
    ("__imul__", lambda v: 2 * v * v),                # :213-216
    ("__itruediv__", lambda v: np.full_like(v, 2.0)),  # :218-221
])
def test_in_place_ops(device, dset, op, want):
    y = _y(dset)
    x = _x(dset, 2 * _vals(y))
    x = getattr(x, op)(y)
    assert np.array_equal(_vals(x), want(_vals(y)))


@pytest.mark.parametrize("op", [lambda a, b: a + b, lambda a, b: a - b, lambda a, b: a * b,
                                lambda a, b: a / b])
def test_shape_mismatch(device, op):
    """test_linalg.py:102-116: a (1, 2) Dat against a (2, 1) Dat."""
    s = op2.Set(N)
    x2 = op2.Dat(s ** (1, 2), np.zeros(2 * N), np.float64, "x")
    y2 = op2.Dat(s ** (2, 1), np.zeros(2 * N), np.float64, "y")
    with pytest.raises(ValueError):
        op(x2, y2)


def test_linalg_and_parloop(device, dset):
    """test_linalg.py:188-194: a par_loop's output in an expression."""
    x = _x(dset, np.zeros(N))
    y = _y(dset)
    op2.par_loop(op2.Kernel("static void k(double *x) { *x = 1.0; }", "k"), dset.set, x(op2.WRITE))
    assert np.array_equal(_vals(x + y), _vals(y) + 1)


def test_norm(device):
    """test_linalg.py:294-297."""
    n = op2.Dat(op2.Set(2), [3, 4], np.float64, "n")
    assert abs(n.norm - 5) < 1e-12
