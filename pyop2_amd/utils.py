"""Small helpers with the semantics of ``pyop2/utils.py``."""
import functools
import inspect
from collections.abc import Iterable

import numpy as np

from pyop2_amd.exceptions import DataTypeError, DataValueError


class cached_property:
    """A read-only @property that is only evaluated once."""

    def __init__(self, fget, doc=None):
        self.fget = fget
        self.__doc__ = doc or fget.__doc__
        self.__name__ = fget.__name__

    def __get__(self, obj, cls):
        if obj is None:
            return self
        obj.__dict__[self.__name__] = result = self.fget(obj)
        return result


def as_tuple(item, type=None, length=None, allow_none=False):
    """``item`` as a tuple: None -> (), iterables -> tuple(item), anything
    else -> ``length`` (default 1) copies.  Checks the length and the item
    types when asked (``TypeError`` / ``ValueError``)."""
    if item is None:
        out = ()
    elif isinstance(item, str) or not hasattr(item, "__iter__"):
        out = (item,) * (length or 1) if not isinstance(item, str) else tuple(item)
    else:
        try:
            out = tuple(item)
        except (TypeError, NotImplementedError):
            out = (item,) * (length or 1)
    if length and len(out) != length:
        raise ValueError("Tuple needs to be of length %d" % length)
    if type is not None:
        ok = (lambda v: isinstance(v, type) or v is None) if allow_none else (lambda v: isinstance(v, type))
        if not all(ok(v) for v in out):
            raise TypeError("Items need to be of type %s" % type)
    return out


def tuplify(xs):
    """Turn a data structure into a tuple tree."""
    if isinstance(xs, Iterable) and not isinstance(xs, str):
        return tuple(tuplify(x) for x in xs)
    return xs


class validate_base:
    """Decorator checking argument types/values before the call."""

    def __init__(self, *checks):
        self._checks = checks

    def __call__(self, f):
        def wrapper(f, *args, **kwargs):
            if self._configuration_type_check():
                code_of = inspect.unwrap(f)   # stacked validators: the real signature
                self.nargs = code_of.__code__.co_argcount
                self.defaults = code_of.__defaults__ or ()
                self.varnames = code_of.__code__.co_varnames
                self.file = code_of.__code__.co_filename
                self.line = code_of.__code__.co_firstlineno + 1
                self.check_args(args, kwargs)
            return f(*args, **kwargs)
        return _decorate(f, wrapper)

    @staticmethod
    def _configuration_type_check():
        from pyop2_amd.configuration import configuration
        return configuration["type_check"]

    def check_args(self, args, kwargs):
        """Check each named argument that was actually passed; parameters the
        function does not have, arguments not given and arguments equal to
        the parameter's default are accepted."""
        for argname, argcond, exception in self._checks:
            if argname not in self.varnames[:self.nargs]:
                continue
            i = self.varnames.index(argname)
            if argname in kwargs and kwargs[argname] is not None:
                arg = kwargs[argname]
            elif i < len(args):
                arg = args[i]
            else:
                continue
            d = i - self.nargs + len(self.defaults)
            if d >= 0:
                try:
                    if arg is self.defaults[d] or bool(arg == self.defaults[d]):
                        continue
                except (ValueError, TypeError):
                    pass
            self.check_arg(arg, argcond, exception)


def _decorate(f, wrapper):
    @functools.wraps(f)
    def inner(*args, **kwargs):
        return wrapper(f, *args, **kwargs)
    inner.__signature__ = inspect.signature(f)
    return inner


class validate_type(validate_base):
    """Check that arguments are of the given types (None is allowed when
    it is the default)."""

    def check_arg(self, arg, argtype, exception):
        if not isinstance(arg, argtype):
            raise exception("%s:%d Parameter %s must be of type %r"
                            % (self.file, self.line, getattr(arg, "__name__", arg), argtype))


class validate_in(validate_base):
    """Check that arguments take one of the given values."""

    def check_arg(self, arg, values, exception):
        if arg not in values:
            raise exception("%s:%d %s must be one of %s" % (self.file, self.line, arg, values))


class validate_range(validate_base):
    """Check that arguments lie within a [lo, hi] range."""

    def check_arg(self, arg, range, exception):
        if not range[0] <= arg <= range[1]:
            raise exception("%s:%d %s must be within range %s" % (self.file, self.line, arg, range))


class validate_dtype(validate_base):
    """Check that a dtype argument is a numpy dtype (or None)."""

    def check_arg(self, arg, ignored, exception):
        if arg is None:
            return
        try:
            np.dtype(arg)
        except TypeError:
            raise exception("%s:%d %s must be a valid dtype" % (self.file, self.line, arg))


def verify_reshape(data, dtype, shape, allow_none=False):
    """``data`` as a numpy array of ``dtype`` with exactly ``shape``.

    Raises ``DataTypeError`` for an unusable dtype and ``DataValueError``
    for None (unless ``allow_none``: an empty array) or a shape mismatch.
    The array is not copied when it already has the dtype; its shape is set
    in place, which requires contiguous data.
    """
    try:
        target = None if dtype is None else np.dtype(dtype)
    except TypeError:
        raise DataTypeError("Invalid data type: %s" % dtype)
    if data is None:
        if allow_none:
            return np.asarray([], dtype=target)
        raise DataValueError("Invalid data: None is not allowed!")
    try:
        arr = np.asarray(data, dtype=target)
    except ValueError:
        raise DataValueError("Invalid data: cannot convert to %s!" % dtype)
    except TypeError:
        raise DataTypeError("Invalid data type: %s" % dtype)
    try:
        arr.shape = shape
    except ValueError:
        raise DataValueError("Expected data of shape %s, got %s" % (shape, arr.shape))
    return arr
