"""Iteration sets of the device par_loop.

A set is an extent ``[0, total)`` of elements in three bands: ``[0, core)``
(owned, its loops read no halo data), ``[core, owned)`` (owned, reading
halo data) and ``[owned, total)`` (halo copies, never iterated).  A
par_loop computes the core band and then the owned band, as the reference
does (``pyop2/types/set.py:16-119`` for the sizes and partitions,
``:533-537`` for ``SetPartition``, ``:394-530`` for ``Subset``,
``:305-391`` for ``ExtrudedSet``).

Here a set is a small value object: its three band limits, a name, and a
per-object ``_cache`` on which plans built over it are kept.  Subsets are a
sorted index array into a plain (or extruded) superset; the set algebra of
the API (intersection, union, differences) is computed on those index
arrays by one table (``_ALGEBRA``) instead of per-class methods.
"""
import hashlib
import numbers

import numpy as np

from pyop2_amd import caching, exceptions as ex, utils
from pyop2_amd.datatypes import IntType

__all__ = ["Set", "ExtrudedSet", "Subset", "SetPartition", "MixedSet"]


def _band_limits(size):
    """(core, owned, total) from an integer (all three equal) or a sequence
    of three non-decreasing integers."""
    if isinstance(size, numbers.Integral):
        n = int(size)
        return (n, n, n)
    limits = utils.as_tuple(size, numbers.Integral, 3)
    core, owned, total = (int(v) for v in limits)
    assert 0 <= core <= owned <= total, \
        "Set band limits must satisfy 0 <= core <= owned <= total: %s" % (limits,)
    return (core, owned, total)


def _subset_from_call(base, indices):
    """``base(i)``, ``base(i, j, ...)`` or ``base(array)`` -> Subset."""
    if len(indices) == 1:
        (indices,) = indices
        if np.isscalar(indices):
            indices = [indices]
    return Subset(base, indices)


class SetPartition:
    """The contiguous run ``[offset, offset + size)`` of an iteration set
    that one ``Parloop._compute`` call covers."""

    __slots__ = ("set", "offset", "size")

    def __init__(self, set, offset, size):
        self.set = set
        self.offset = offset
        self.size = size

    def __repr__(self):
        return "SetPartition(%r, %d, %d)" % (self.set, self.offset, self.size)


class Set:
    """A set of ``size`` elements; ``size`` is one integer or the three band
    limits ``[core, owned, total]``."""

    _extruded = False
    _extruded_periodic = False
    _kernel_args_ = ()

    @utils.validate_type(('size', (numbers.Integral, tuple, list, np.ndarray), ex.SizeTypeError),
                         ('name', str, ex.NameTypeError))
    def __init__(self, size, name=None, halo=None, comm=None, constrained_size=0):
        self._sizes = _band_limits(size)
        self._name = name if name is not None else "set_#x%x" % id(self)
        self._halo = halo
        self.comm = comm
        self._partition_size = 1024
        self._constrained_size = constrained_size
        self._cache = {}

    # -- the bands ---------------------------------------------------------------
    @property
    def sizes(self):
        return self._sizes

    @property
    def core_size(self):
        return self._sizes[0]

    @property
    def size(self):
        return self._sizes[1]

    @property
    def total_size(self):
        return self._sizes[2]

    @property
    def constrained_size(self):
        return self._constrained_size

    @utils.cached_property
    def core_part(self):
        return SetPartition(self, 0, self.core_size)

    @utils.cached_property
    def owned_part(self):
        return SetPartition(self, self.core_size, self.size - self.core_size)

    @property
    def indices(self):
        return range(self.total_size)

    # -- descriptive attributes ----------------------------------------------------
    @property
    def name(self):
        return self._name

    @property
    def halo(self):
        return self._halo

    @property
    def layers(self):
        return None

    @property
    def partition_size(self):
        return self._partition_size

    @partition_size.setter
    def partition_size(self, value):
        self._partition_size = value

    def _plan_key(self):
        """Identifies the elements a loop over this set runs, for plans
        cached on the maps the loop reads (those maps keep their iteration
        set alive, so its identity is stable while the cache exists)."""
        return ("set", id(self))

    # -- container protocol: a plain set is a one-component mixed set ----------------
    def __iter__(self):
        yield self

    def __len__(self):
        return 1

    def __getitem__(self, idx):
        assert idx == 0, "a Set has one component"
        return self

    def __contains__(self, dset):
        from pyop2_amd.types.dataset import DataSet
        return isinstance(dset, DataSet) and dset.set is self

    def __pow__(self, e):
        from pyop2_amd.types.dataset import DataSet
        return DataSet(self, dim=e)

    def __call__(self, *indices):
        return _subset_from_call(self, indices)

    def __eq__(self, other):
        return isinstance(other, Set) and (self._sizes, self._name) == (other._sizes, other._name)

    def __hash__(self):
        return hash((self._sizes, self._name))

    def __str__(self):
        return "OP2 Set: %s with size %s" % (self._name, self.size)

    def __repr__(self):
        return "Set(%r, %r)" % (self._sizes, self._name)

    # -- set algebra -----------------------------------------------------------------
    def intersection(self, other):
        return _apply("intersection", self, other)

    def union(self, other):
        return _apply("union", self, other)

    def difference(self, other):
        return _apply("difference", self, other)

    def symmetric_difference(self, other):
        return _apply("symmetric_difference", self, other)


class ExtrudedSet(Set):
    """Columns of cells over the elements of ``parent``.

    ``layers`` is one integer (every column has ``layers`` layers, i.e.
    ``layers - 1`` cells) or a ``(parent.total_size, 2)`` array of
    ``[bottom, top)`` layers per column.  A par_loop runs, for every column
    of the parent's bands, the column's cells bottom to top; map entries
    step by the map's ``offset`` per layer (wrapping on a periodic set).
    """

    _extruded = True

    @utils.validate_type(('parent', Set, TypeError))
    def __init__(self, parent, layers, extruded_periodic=False):
        if isinstance(parent, (ExtrudedSet, Subset)):
            raise NotImplementedError("only a plain Set can be extruded on the device path")
        self._parent = parent
        self.comm = parent.comm
        self._layers, self.constant_layers = self._layer_table(parent, layers)
        self._extruded_periodic = bool(extruded_periodic)
        self._cache = {}
        self._dev_layers = None
        self._partition_size = 1024
        self._constrained_size = 0

    @staticmethod
    def _layer_table(parent, layers):
        """(int32 table, constant?): a per-column ``[bottom, top)`` table, or
        the single row ``[0, layers]`` for constant layers."""
        arr = np.asarray(layers)
        if arr.ndim == 2:
            table = utils.verify_reshape(layers, IntType, (parent.total_size, 2))
            if table.size and int(table.min()) < 0:
                raise ex.SizeTypeError("Bottom of layers must be >= 0")
            if np.any(table[:, 1] - table[:, 0] < 1):
                raise ex.SizeTypeError("Number of layers must be >= 0")
            return np.ascontiguousarray(table, dtype=IntType), False
        if arr.shape:
            raise ex.SizeTypeError("layers per column need shape (%d, 2), not %s"
                                   % (parent.total_size, arr.shape))
        if int(arr) < 2:
            raise ex.SizeTypeError("Need at least two layers, not %d" % int(arr))
        return np.array([[0, int(arr)]], dtype=IntType), True

    @property
    def _sizes(self):
        return self._parent.sizes

    @property
    def _name(self):
        return self._parent.name

    @property
    def _halo(self):
        return self._parent.halo

    @property
    def parent(self):
        return self._parent

    @property
    def layers(self):
        """The layer count of every column (constant layers only)."""
        if not self.constant_layers:
            raise ValueError("No single layer, use layers_array attribute")
        return int(self._layers[0, 1])

    @property
    def layers_array(self):
        return self._layers

    def column_layers(self, columns):
        """(bottom, top) layer of each host column id in ``columns``."""
        cols = np.asarray(columns, dtype=np.int64)
        if self.constant_layers:
            return np.broadcast_to(self._layers[0], (len(cols), 2))
        return self._layers[cols]

    @property
    def device_layers(self):
        """Device copy of ``layers_array``, indexed by column id."""
        if self._dev_layers is None:
            from pyop2_amd import device as dev
            self._dev_layers = dev.to_device(self._layers)
        return self._dev_layers

    @property
    def _kernel_args_(self):
        """The wrapper's first pointer: the layers table (``set.py:349-351``)."""
        from pyop2_amd import device as dev
        return (dev.ptr(self.device_layers),)

    def __contains__(self, other):
        return other is self._parent

    def __eq__(self, other):
        return self is other

    def __hash__(self):
        return id(self)

    def __str__(self):
        return "OP2 ExtrudedSet: %s with size %s (%s layers)" % (self._name, self.size, self._layers)

    def __repr__(self):
        return "ExtrudedSet(%r, %r)" % (self._parent, self._layers)


class Subset(Set):
    """The elements ``indices`` (sorted, duplicates dropped) of a Set or
    ExtrudedSet; a Subset of a Subset refers to the root superset."""

    @utils.validate_type(('superset', Set, TypeError),
                         ('indices', (list, tuple, np.ndarray), TypeError))
    def __init__(self, superset, indices):
        chosen = np.unique(np.asarray(indices))
        while isinstance(superset, Subset):
            chosen = superset.indices[chosen]
            superset = superset.superset
        assert type(superset) in (Set, ExtrudedSet), "a Subset needs a Set or ExtrudedSet superset"
        chosen = utils.verify_reshape(chosen, IntType, (len(chosen),))
        if len(chosen) and (chosen[0] < 0 or chosen[-1] >= superset.total_size):
            raise ex.SubsetIndexOutOfBounds(
                "Subset indices span [%d, %d], outside [0, %d)"
                % (chosen[0], chosen[-1], superset.total_size))
        self.comm = superset.comm
        self._superset = superset
        self._indices = chosen
        bounds = np.searchsorted(chosen, [superset.core_size, superset.size])
        self._sizes = (int(bounds[0]), int(bounds[1]), len(chosen))
        self._extruded = superset._extruded
        self._extruded_periodic = superset._extruded_periodic
        self._device_indices = None
        self._digest = None

    def __getattr__(self, name):
        # attributes a Subset does not define (layers, halo, ...) are its
        # superset's; private ones of its own that are not yet set are not
        if name.startswith("__") or name in ("_superset", "_device_indices", "_digest", "_indices"):
            raise AttributeError(name)
        return getattr(self._superset, name)

    @property
    def superset(self):
        return self._superset

    @property
    def indices(self):
        return self._indices

    @property
    def owned_indices(self):
        return self._indices[:self._sizes[1]]

    @property
    def name(self):
        return self._superset.name

    @property
    def _name(self):
        return self._superset.name

    @property
    def layers_array(self):
        """The superset's layers of the chosen columns (all of them when the
        layers are constant); the loops read the superset's table by column
        id (``device_layers``)."""
        sup = self._superset
        if not sup._extruded:
            raise AttributeError("layers_array")
        return sup.layers_array if sup.constant_layers else sup.layers_array[self._indices]

    @property
    def device_indices(self):
        """Device copy (int32) of the chosen indices."""
        if self._device_indices is None:
            from pyop2_amd import device as dev
            self._device_indices = dev.to_device(self._indices)
        return self._device_indices

    @property
    def _kernel_args_(self):
        """The superset's pointers, then the chosen indices (``set.py:432-434``)."""
        from pyop2_amd import device as dev
        return self._superset._kernel_args_ + (dev.ptr(self.device_indices),)

    def _plan_key(self):
        """Identifies the chosen elements by content: two Subsets with the same
        indices share plans, and a new Subset never picks up the plan of a
        freed one whose id it reuses."""
        if self._digest is None:
            self._digest = hashlib.blake2b(self._indices.tobytes(), digest_size=16).hexdigest()
        return ("subset", id(self._superset), len(self._indices), self._digest)

    def __call__(self, *indices):
        return _subset_from_call(self, indices)

    def __pow__(self, e):
        raise NotImplementedError("Deriving a DataSet from a Subset is unsupported")

    def __str__(self):
        return "OP2 Subset: %s with sizes %s" % (self.name, self._sizes)

    def __repr__(self):
        return "Subset(%r, %r)" % (self._superset, self._indices)


# ---------------------------------------------------------------------------
# Set algebra.  Operands are a plain set S (Set or ExtrudedSet) or a Subset of
# one; both must share the same plain set.  The result of each operation,
# by operand kinds (S: the plain set, A / B: Subsets of it):
#
#                     S op S     S op B        A op S        A op B
#   intersection      S          B             A             A & B
#   union             S          S             S             A | B
#   difference        empty      S - B         empty         A - B
#   symmetric diff.   empty      S - B         S - A         A ^ B
#
# ("S - B" and "empty" are new Subsets of S), which is the reference's
# behaviour (pyop2/types/set.py:197-227, 493-530).

def _kind(s):
    if type(s) is Subset:
        return "sub"
    if type(s) in (Set, ExtrudedSet):
        return "set"
    return None


def _check_algebra(a, b):
    ka, kb = _kind(a), _kind(b)
    if kb is None:
        raise TypeError("Unable to perform set operations between `%s` and %s."
                        % (type(a).__name__, type(b)))
    base_a = a if ka == "set" else a.superset
    base_b = b if kb == "set" else b.superset
    if base_a is not base_b:
        if ka == kb == "set":
            raise ValueError("Unable to perform set operations between two unrelated sets: %s and %s."
                             % (a, b))
        raise TypeError("Superset mismatch in a set operation: %s and %s" % (base_a, base_b))
    return ka, kb, base_a


def _complement(base, sub):
    return np.setdiff1d(np.arange(base.total_size, dtype=IntType), sub.indices)


_ALGEBRA = {
    ("intersection", "set", "set"): lambda a, b, s: a,
    ("intersection", "set", "sub"): lambda a, b, s: b,
    ("intersection", "sub", "set"): lambda a, b, s: a,
    ("intersection", "sub", "sub"): lambda a, b, s: Subset(s, np.intersect1d(a.indices, b.indices)),
    ("union", "set", "set"): lambda a, b, s: s,
    ("union", "set", "sub"): lambda a, b, s: s,
    ("union", "sub", "set"): lambda a, b, s: s,
    ("union", "sub", "sub"): lambda a, b, s: Subset(s, np.union1d(a.indices, b.indices)),
    ("difference", "set", "set"): lambda a, b, s: Subset(s, []),
    ("difference", "set", "sub"): lambda a, b, s: Subset(s, _complement(s, b)),
    ("difference", "sub", "set"): lambda a, b, s: Subset(s, []),
    ("difference", "sub", "sub"): lambda a, b, s: Subset(s, np.setdiff1d(a.indices, b.indices)),
    ("symmetric_difference", "set", "set"): lambda a, b, s: Subset(s, []),
    ("symmetric_difference", "set", "sub"): lambda a, b, s: Subset(s, _complement(s, b)),
    ("symmetric_difference", "sub", "set"): lambda a, b, s: Subset(s, _complement(s, a)),
    ("symmetric_difference", "sub", "sub"): lambda a, b, s: Subset(s, np.setxor1d(a.indices, b.indices)),
}


def _apply(op, a, b):
    ka, kb, base = _check_algebra(a, b)
    return _ALGEBRA[(op, ka, kb)](a, b, base)


class MixedSet(Set, caching.ObjectCached):
    """A tuple of Sets, the iteration space of mixed data (a par_loop cannot
    iterate it, as in the reference).  Cached on its first component."""

    def __init__(self, sets):
        if self._initialized:
            return
        self._sets = sets
        self._cache = {}
        self._initialized = True

    @classmethod
    def _process_args(cls, sets, **kwargs):
        try:
            sets = utils.as_tuple(list(sets), Set)
        except TypeError:
            raise ex.SetTypeError("All components of a MixedSet must be Sets")
        return (sets[0], sets), kwargs

    @classmethod
    def _cache_key(cls, sets, **kwargs):
        return sets

    @property
    def split(self):
        return self._sets

    def _summed(self, attr):
        return sum(getattr(s, attr) for s in self._sets if s is not None)

    @property
    def core_size(self):
        return self._summed("core_size")

    @property
    def size(self):
        return self._summed("size")

    @property
    def total_size(self):
        return self._summed("total_size")

    @property
    def sizes(self):
        return (self.core_size, self.size, self.total_size)

    @property
    def name(self):
        return tuple(s.name for s in self._sets)

    def __getitem__(self, idx):
        return self._sets[idx]

    def __iter__(self):
        return iter(self._sets)

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
