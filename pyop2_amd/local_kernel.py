"""Local (per-element) kernels (``pyop2/local_kernel.py``).

``Kernel(code, name)`` with a C string.  Loopy kernels are not supported
(loopy is absent): they raise ``TypeError`` as any other unknown code type.
"""
import abc
import hashlib
from dataclasses import dataclass
from typing import Union

import numpy as np

from pyop2_amd import __version__
from pyop2_amd.exceptions import NameTypeError
from pyop2_amd.types.access import Access
from pyop2_amd.utils import cached_property, validate_type


@dataclass(frozen=True)
class LocalKernelArg:
    access: Access
    dtype: Union[np.dtype, str]


@validate_type(("name", str, NameTypeError))
def Kernel(code, name, **kwargs):
    """Construct a local kernel from a C string."""
    if isinstance(code, str):
        return CStringLocalKernel(code, name, **kwargs)
    raise TypeError("code argument is the wrong type")


class LocalKernel(abc.ABC):
    """The kernel executed per member of the iteration set.

    Keyword arguments as in the reference (``cpp``, ``flop_count``,
    ``headers``, ``include_dirs``, ``ldargs``, ``opts``,
    ``requires_zeroed_output_arguments``, ``user_code``, ``events``).
    """

    @validate_type(("name", str, NameTypeError))
    def __init__(self, code, name, accesses=None, *, cpp=False, flop_count=None, headers=(),
                 include_dirs=(), ldargs=(), opts=None, requires_zeroed_output_arguments=False,
                 user_code="", events=()):
        self.code = code
        self.name = name
        self.accesses = accesses
        self.cpp = cpp
        self.flop_count = flop_count
        self.headers = headers
        self.include_dirs = include_dirs
        self.ldargs = ldargs
        self.opts = opts or {}
        self.requires_zeroed_output_arguments = requires_zeroed_output_arguments
        self.user_code = user_code
        self.events = events

    @property
    @abc.abstractmethod
    def dtypes(self):
        """The dtypes of the kernel arguments."""

    @property
    def cache_key(self):
        return self._immutable_cache_key, self.accesses, self.dtypes

    @cached_property
    def _immutable_cache_key(self):
        key = (self.code, self.name, self.cpp, self.flop_count, self.headers, self.include_dirs,
               self.ldargs, sorted(self.opts.items()), self.requires_zeroed_output_arguments,
               self.user_code, __version__)
        return hashlib.md5(str(key).encode()).hexdigest()

    @property
    def arguments(self):
        assert len(self.accesses) == len(self.dtypes)
        return tuple(LocalKernelArg(acc, dtype) for acc, dtype in zip(self.accesses, self.dtypes))

    @cached_property
    def num_flops(self):
        return self.flop_count if self.flop_count is not None else 0

    def __eq__(self, other):
        if not isinstance(other, LocalKernel):
            return NotImplemented
        return self.cache_key == other.cache_key

    def __hash__(self):
        return hash(self.cache_key)

    def __str__(self):
        return f"OP2 Kernel: {self.name}"

    def __repr__(self):
        return 'Kernel("""%s""", %r)' % (self.code, self.name)


class CStringLocalKernel(LocalKernel):
    """A :class:`LocalKernel` whose ``code`` is a string of C code."""

    @validate_type(("code", str, TypeError))
    def __init__(self, code, name, accesses=None, dtypes=None, **kwargs):
        super().__init__(code, name, accesses, **kwargs)
        self._dtypes = dtypes

    @property
    def dtypes(self):
        return self._dtypes

    @dtypes.setter
    def dtypes(self, dtypes):
        self._dtypes = dtypes
