"""Sets and iteration partitions (``pyop2/types/set.py``).

Only the single-rank non-extruded forms are on the device path: ``Set``
with its ``(core, owned, total)`` sizes and ``core_part``/``owned_part``
(``set.py:16-227``), ``Subset`` (``set.py:394-530``) and ``SetPartition``
(``set.py:533-537``).  ``MixedSet`` exists for API parity (iteration over
it is refused, as in the reference).
"""
import numbers

import numpy as np

from pyop2_amd import caching, exceptions as ex, utils
from pyop2_amd.datatypes import IntType


class Set:
    """OP2 set of ``size`` elements, or ``[core, owned, total]`` sizes.

    Elements are divided as ``[0, core)`` (owned, not touching the halo),
    ``[core, owned)`` and the halo ``[owned, total)``.  A par_loop computes
    ``[0, core)`` and then ``[core, owned)``; the halo is never iterated.
    """

    _CORE_SIZE = 0
    _OWNED_SIZE = 1
    _GHOST_SIZE = 2

    _extruded = False
    _extruded_periodic = False

    _kernel_args_ = ()

    @utils.validate_type(('size', (numbers.Integral, tuple, list, np.ndarray), ex.SizeTypeError),
                         ('name', str, ex.NameTypeError))
    def __init__(self, size, name=None, halo=None, comm=None, constrained_size=0):
        if isinstance(size, numbers.Integral):
            size = [size] * 3
        size = utils.as_tuple(size, numbers.Integral, 3)
        assert size[Set._CORE_SIZE] <= size[Set._OWNED_SIZE] <= size[Set._GHOST_SIZE], \
            "Set received invalid sizes: %s" % (size,)
        self._sizes = tuple(int(s) for s in size)
        self._name = name or "set_#x%x" % id(self)
        self._halo = halo
        self.comm = comm
        self._partition_size = 1024
        self._constrained_size = constrained_size
        self._cache = {}

    @property
    def indices(self):
        return range(self.total_size)

    @utils.cached_property
    def core_size(self):
        return self._sizes[Set._CORE_SIZE]

    @utils.cached_property
    def constrained_size(self):
        return self._constrained_size

    @utils.cached_property
    def size(self):
        return self._sizes[Set._OWNED_SIZE]

    @utils.cached_property
    def total_size(self):
        return self._sizes[Set._GHOST_SIZE]

    @utils.cached_property
    def sizes(self):
        return self._sizes

    @utils.cached_property
    def core_part(self):
        return SetPartition(self, 0, self.core_size)

    @utils.cached_property
    def owned_part(self):
        return SetPartition(self, self.core_size, self.size - self.core_size)

    @utils.cached_property
    def name(self):
        return self._name

    @utils.cached_property
    def halo(self):
        return self._halo

    @property
    def partition_size(self):
        return self._partition_size

    @partition_size.setter
    def partition_size(self, value):
        self._partition_size = value

    @utils.cached_property
    def layers(self):
        return None

    def __hash__(self):
        return hash((self._sizes, self._name))

    def __eq__(self, other):
        return isinstance(other, Set) and self._sizes == other._sizes and self._name == other._name

    def __iter__(self):
        yield self

    def __getitem__(self, idx):
        assert idx == 0
        return self

    def __len__(self):
        return 1

    def __str__(self):
        return "OP2 Set: %s with size %s" % (self._name, self.size)

    def __repr__(self):
        return "Set(%r, %r)" % (self._sizes, self._name)

    def __call__(self, *indices):
        """Build a :class:`Subset` from this :class:`Set`."""
        if len(indices) == 1:
            indices = indices[0]
            if np.isscalar(indices):
                indices = [indices]
        return Subset(self, indices)

    def __contains__(self, dset):
        from pyop2_amd.types.dataset import DataSet
        if isinstance(dset, DataSet):
            return dset.set is self
        return False

    def __pow__(self, e):
        from pyop2_amd.types.dataset import DataSet
        return DataSet(self, dim=e)

    def _check_operands(self, other):
        if type(other) is Set:
            if other is not self:
                raise ValueError("Unable to perform set operations between two unrelated sets: %s and %s." % (self, other))
        elif type(other) is Subset:
            if self is not other._superset:
                raise TypeError("Superset mismatch: self (%s) != other._superset (%s)" % (self, other._superset))
        else:
            raise TypeError("Unable to perform set operations between `Set` and %s." % (type(other),))

    def intersection(self, other):
        self._check_operands(other)
        return other

    def union(self, other):
        self._check_operands(other)
        return self

    def difference(self, other):
        self._check_operands(other)
        if other is self:
            return Subset(self, [])
        return type(other)(self, np.setdiff1d(np.asarray(range(self.total_size), dtype=IntType), other._indices))

    def symmetric_difference(self, other):
        self._check_operands(other)
        return self.difference(other)


class ExtrudedSet(Set):
    """A Set extruded into columns of cells (``pyop2/types/set.py:305-391``).

    ``layers`` is the number of layers of every column (``layers - 1`` cells
    each) or an ``(parent.total_size, 2)`` array of [bottom, top) layers per
    column.  A par_loop over it runs, for every column of the parent's
    partition, the cells of the column bottom to top; map entries step by
    the map's ``offset`` per layer.
    """

    _extruded = True

    @utils.validate_type(('parent', Set, TypeError))
    def __init__(self, parent, layers, extruded_periodic=False):
        if isinstance(parent, (ExtrudedSet, Subset)):
            raise NotImplementedError("only a plain Set can be extruded on the device path")
        self._parent = parent
        self.comm = parent.comm
        arr = np.asarray(layers)
        if arr.ndim == 2:
            arr = utils.verify_reshape(layers, IntType, (parent.total_size, 2))
            if arr.size and arr.min() < 0:
                raise ex.SizeTypeError("Bottom of layers must be >= 0")
            if np.any(arr[:, 1] - arr[:, 0] < 1):
                raise ex.SizeTypeError("Number of layers must be >= 0")
            self.constant_layers = False
        else:
            if arr.shape:
                raise ex.SizeTypeError("Specifying layers per entity, but provided %s, needed (%d, 2)"
                                       % (arr.shape, parent.total_size))
            if int(layers) < 2:
                raise ex.SizeTypeError("Need at least two layers, not %d" % int(layers))
            arr = np.asarray([[0, int(layers)]], dtype=IntType)
            self.constant_layers = True
        self._layers = np.ascontiguousarray(arr, dtype=IntType)
        # periodic: the top of every column's last cell is its bottom
        self._extruded_periodic = bool(extruded_periodic)
        self._cache = {}
        self._dev_layers = None
        self._partition_size = 1024
        self._constrained_size = 0

    @utils.cached_property
    def parent(self):
        return self._parent

    @utils.cached_property
    def layers(self):
        """The number of layers (constant layers only)."""
        if self.constant_layers:
            return int(self._layers[0, 1])
        raise ValueError("No single layer, use layers_array attribute")

    @utils.cached_property
    def layers_array(self):
        return self._layers

    @property
    def _kernel_args_(self):
        """The layers array's device pointer, the wrapper's first pointer
        argument (``set.py:349-351``)."""
        from pyop2_amd import device as dev
        return (dev.ptr(self.device_layers),)

    @property
    def device_layers(self):
        """Device copy of ``layers_array``."""
        if self._dev_layers is None:
            from pyop2_amd import device as dev
            self._dev_layers = dev.to_device(self._layers)
        return self._dev_layers

    def column_layers(self, columns):
        """(bottom, top) layer of each of the host column ids ``columns``."""
        cols = np.asarray(columns, dtype=np.int64)
        if self.constant_layers:
            return np.broadcast_to(self._layers[0], (len(cols), 2))
        return self._layers[cols]

    @property
    def _sizes(self):
        return self._parent.sizes

    @property
    def _name(self):
        return self._parent.name

    @property
    def _halo(self):
        return self._parent.halo

    def __contains__(self, set):
        return set is self._parent

    def __eq__(self, other):
        return self is other

    def __hash__(self):
        return id(self)

    def __str__(self):
        return "OP2 ExtrudedSet: %s with size %s (%s layers)" % (self._name, self.size, self._layers)

    def __repr__(self):
        return "ExtrudedSet(%r, %r)" % (self._parent, self._layers)


class Subset(Set):
    """Elements ``indices`` (sorted, de-duplicated) of ``superset``."""

    @utils.validate_type(('superset', Set, TypeError),
                         ('indices', (list, tuple, np.ndarray), TypeError))
    def __init__(self, superset, indices):
        self.comm = superset.comm
        indices = np.unique(indices)
        if isinstance(superset, Subset):
            indices = superset.indices[indices]
            superset = superset.superset
        assert type(superset) in (Set, ExtrudedSet), "Subset construction failed, should not happen"
        self._superset = superset
        self._indices = utils.verify_reshape(indices, IntType, (len(indices),))
        if len(self._indices) > 0 and (self._indices[0] < 0 or self._indices[-1] >= self._superset.total_size):
            raise ex.SubsetIndexOutOfBounds(
                'Out of bounds indices in Subset construction: [%d, %d) not [0, %d)' %
                (self._indices[0], self._indices[-1], self._superset.total_size))
        self._sizes = (int((self._indices < superset.core_size).sum()),
                       int((self._indices < superset.size).sum()),
                       len(self._indices))
        # a Subset of an ExtrudedSet runs the chosen columns (set.py:427-428)
        self._extruded = superset._extruded
        self._extruded_periodic = superset._extruded_periodic
        self._device_indices = None

    def __getattr__(self, name):
        if name.startswith("__") or name in ("_superset", "_device_indices"):
            raise AttributeError(name)
        return getattr(self._superset, name)

    def __pow__(self, e):
        raise NotImplementedError("Deriving a DataSet from a Subset is unsupported")

    def __str__(self):
        return "OP2 Subset: %s with sizes %s" % (self._name, self._sizes)

    def __repr__(self):
        return "Subset(%r, %r)" % (self._superset, self._indices)

    def __call__(self, *indices):
        if len(indices) == 1:
            indices = indices[0]
            if np.isscalar(indices):
                indices = [indices]
        return Subset(self, indices)

    @utils.cached_property
    def superset(self):
        return self._superset

    @utils.cached_property
    def indices(self):
        return self._indices

    @utils.cached_property
    def owned_indices(self):
        return self.indices[self.indices < self.superset.size]

    @property
    def _kernel_args_(self):
        """The superset's pointers, then the indices' device pointer
        (``set.py:432-434``)."""
        from pyop2_amd import device as dev
        return self._superset._kernel_args_ + (dev.ptr(self.device_indices),)

    @property
    def device_indices(self):
        """Device copy of the subset indices (int32)."""
        if self._device_indices is None:
            from pyop2_amd import device as dev
            self._device_indices = dev.to_device(self._indices)
        return self._device_indices

    def _check_operands(self, other):
        if type(other) is Set:
            if other is not self._superset:
                raise TypeError("Superset mismatch: self._superset (%s) != other (%s)" % (self._superset, other))
        elif type(other) is Subset:
            if self._superset is not other._superset:
                raise TypeError("Unable to perform set operation between subsets of mismatching supersets (%s != %s)" % (self._superset, other._superset))
        else:
            raise TypeError("Unable to perform set operations between `Subset` and %s." % (type(other),))

    def intersection(self, other):
        self._check_operands(other)
        if other is self._superset:
            return self
        return type(self)(self._superset, np.intersect1d(self._indices, other._indices))

    def union(self, other):
        self._check_operands(other)
        if other is self._superset:
            return other
        return type(self)(self._superset, np.union1d(self._indices, other._indices))

    def difference(self, other):
        self._check_operands(other)
        if other is self._superset:
            return Subset(other, [])
        return type(self)(self._superset, np.setdiff1d(self._indices, other._indices))

    def symmetric_difference(self, other):
        self._check_operands(other)
        if other is self._superset:
            return other.symmetric_difference(self)
        return type(self)(self._superset, np.setxor1d(self._indices, other._indices))


class SetPartition:
    """A contiguous part ``[offset, offset+size)`` of an iteration set."""

    def __init__(self, set, offset, size):
        self.set = set
        self.offset = offset
        self.size = size


class MixedSet(Set, caching.ObjectCached):
    """A tuple of Sets (API parity; par_loops cannot iterate it)."""

    def __init__(self, sets):
        if self._initialized:
            return
        self._sets = sets
        assert all(s is None or isinstance(s, Set) for s in sets), "All components must be Sets"
        self._initialized = True
        self._cache = {}

    @classmethod
    def _process_args(cls, sets, **kwargs):
        sets = [s for s in sets]
        try:
            sets = utils.as_tuple(sets, Set)
        except TypeError:
            raise ex.SetTypeError("All components of a MixedSet must be Sets")
        return (sets[0],) + (sets,), kwargs

    @classmethod
    def _cache_key(cls, sets, **kwargs):
        return sets

    def __getitem__(self, idx):
        return self._sets[idx]

    @utils.cached_property
    def split(self):
        return self._sets

    @utils.cached_property
    def core_size(self):
        return sum(s.core_size for s in self._sets)

    @utils.cached_property
    def size(self):
        return sum(0 if s is None else s.size for s in self._sets)

    @utils.cached_property
    def total_size(self):
        return sum(0 if s is None else s.total_size for s in self._sets)

    @utils.cached_property
    def sizes(self):
        return (self.core_size, self.size, self.total_size)

    @utils.cached_property
    def name(self):
        return tuple(s.name for s in self._sets)

    def __iter__(self):
        for s in self._sets:
            yield s

    def __len__(self):
        return len(self._sets)

    def __pow__(self, e):
        """A MixedDataSet with dimension ``e`` on every component."""
        from pyop2_amd.types.mixed import MixedDataSet
        return MixedDataSet(self, e)

    def __eq__(self, other):
        return type(self) is type(other) and self._sets == other._sets

    def __hash__(self):
        return hash(self._sets)

    def __str__(self):
        return "OP2 MixedSet composed of Sets: %s" % (self._sets,)

    def __repr__(self):
        return "MixedSet(%r)" % (self._sets,)
