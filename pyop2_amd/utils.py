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
    # Empty list if we get passed None
    if item is None:
        t = ()
    else:
        # Convert iterable to tuple...
        try:
            t = tuple(item)
        # ... or create a list of a single item
        except (TypeError, NotImplementedError):
            t = (item,) * (length or 1)
    if length and not len(t) == length:
        raise ValueError("Tuple needs to be of length %d" % length)
    if type is not None:
        if allow_none:
            valid = all((isinstance(i, type) or i is None) for i in t)
        else:
            valid = all(isinstance(i, type) for i in t)
        if not valid:
            raise TypeError("Items need to be of type %s" % type)
    return t


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
        @functools.wraps(f)
        def wrapper(f, *args, **kwargs):
            if self._configuration_type_check():
                self.nargs = f.__code__.co_argcount
                self.defaults = f.__defaults__ or ()
                self.varnames = f.__code__.co_varnames
                self.file = f.__code__.co_filename
                self.line = f.__code__.co_firstlineno + 1
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
    """Verify data is of type dtype and try to reshape it to shape."""
    try:
        t = np.dtype(dtype) if dtype is not None else None
    except TypeError:
        raise DataTypeError("Invalid data type: %s" % dtype)
    if data is None and allow_none:
        return np.asarray([], dtype=t)
    elif data is None:
        raise DataValueError("Invalid data: None is not allowed!")
    else:
        try:
            a = np.asarray(data, dtype=t)
        except ValueError:
            raise DataValueError("Invalid data: cannot convert to %s!" % dtype)
        except TypeError:
            raise DataTypeError("Invalid data type: %s" % dtype)
        try:
            # Destructively modify shape.  Fails if data are not
            # contiguous, but that's what we want anyway.
            a.shape = shape
            return a
        except ValueError:
            raise DataValueError("Expected data of shape %s, got %s" % (shape, a.shape))
