"""Map: iteration set -> target set indirection (``pyop2/types/map.py:17-168``).

Values are int32 ``(iterset.total_size, arity)`` (``map.py:41-43``); the
device copy is uploaded once, on first use by a par_loop, and cached.
"""
import numbers

import numpy as np

from pyop2_amd import exceptions as ex, utils
from pyop2_amd.datatypes import IntType
from pyop2_amd.types.set import Set


class Map:

    dtype = IntType

    @utils.validate_type(('iterset', Set, ex.SetTypeError), ('toset', Set, ex.SetTypeError),
                         ('arity', numbers.Integral, ex.ArityTypeError), ('name', str, ex.NameTypeError))
    def __init__(self, iterset, toset, arity, values=None, name=None, offset=None, offset_quotient=None):
        self._iterset = iterset
        self._toset = toset
        self.comm = toset.comm
        self._arity = int(arity)
        self._values = utils.verify_reshape(values, IntType, (iterset.total_size, arity), allow_none=True)
        self.shape = (iterset.total_size, arity)
        self._name = name or "map_#x%x" % id(self)
        if offset is not None and len(offset) > 0:
            raise NotImplementedError("extruded maps (offset) are not on the device path")
        self._offset = None
        self._offset_quotient = None
        self._cache = {}
        self._dev = None

    @utils.cached_property
    def _global_kernel_arg(self):
        from pyop2_amd.global_kernel import MapKernelArg
        return MapKernelArg(self.arity, None, None)

    @property
    def device_values(self):
        """Device copy (int32) of ``values_with_halo``, range-checked once:
        entries must be below ``toset.total_size`` (negative entries mean
        "no target" and are skipped by the loops)."""
        if self._dev is None:
            from pyop2_amd import device as dev
            v = np.ascontiguousarray(self._values.reshape(self.shape))
            self._has_negative = bool(v.size and v.min() < 0)
            if v.size and v.max() >= self._toset.total_size:
                raise ex.MapValueError(f"{self} has entries >= {self._toset.total_size} (target set size)")
            self._dev = dev.to_device(v)
        return self._dev

    @property
    def has_negative(self):
        """True if some entry is negative (masked target)."""
        self.device_values
        return self._has_negative

    @property
    def _kernel_args_(self):
        from pyop2_amd import device as dev
        return (dev.ptr(self.device_values),)

    def __iter__(self):
        yield self

    def __len__(self):
        return 1

    @utils.cached_property
    def split(self):
        return (self,)

    @utils.cached_property
    def iterset(self):
        return self._iterset

    @utils.cached_property
    def toset(self):
        return self._toset

    @utils.cached_property
    def arity(self):
        return self._arity

    @utils.cached_property
    def arities(self):
        return (self._arity,)

    @utils.cached_property
    def arange(self):
        return (0, self._arity)

    @utils.cached_property
    def values(self):
        """Map values of the owned iteration elements."""
        return self._values[:self.iterset.size]

    @utils.cached_property
    def values_with_halo(self):
        """All map values, halo elements included."""
        return self._values

    @utils.cached_property
    def name(self):
        return self._name

    @utils.cached_property
    def offset(self):
        return self._offset

    @utils.cached_property
    def offset_quotient(self):
        return self._offset_quotient

    @utils.cached_property
    def flattened_maps(self):
        return (self,)

    def __str__(self):
        return "OP2 Map: %s from (%s) to (%s) with arity %s" % (self._name, self._iterset, self._toset, self._arity)

    def __repr__(self):
        return "Map(%r, %r, %r, None, %r)" % (self._iterset, self._toset, self._arity, self._name)

    def __le__(self, o):
        return self == o
