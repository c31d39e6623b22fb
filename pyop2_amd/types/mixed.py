"""Mixed (multi-block) data: MixedDataSet, MixedMap, MixedDat and MixedMat.

Reference: ``pyop2/types/dataset.py:166-312`` (MixedDataSet),
``types/map.py:297-437`` (MixedMap), ``types/dat.py:646-880`` (MixedDat),
``types/mat.py:86-114, 613-670`` (mixed Sparsity and MatNest blocks).

A mixed object is a tuple of ordinary ones, one per component of a
MixedSet.  In a par_loop a MixedDat argument is one kernel buffer: the packs
of its components, each gathered through its component map, laid out one
after the other (``codegen/builder.py:413-498``).  A MixedMat argument is
one element matrix whose rows are the packs of the row components and whose
columns are the packs of the column components; block (i, j) of it is added
into block Mat (i, j) through row map i and column map j
(``builder.py:609-680``; the Mat side is ``MixedMat`` in ``types/mat.py``).
Every component lives in its own device storage, so a component Dat or
block Mat is an ordinary Dat or Mat.
"""
import numbers

import numpy as np

from pyop2_amd import configuration as conf, exceptions as ex, utils
from pyop2_amd.types.access import Access
from pyop2_amd.types.dataset import DataSet
from pyop2_amd.types.set import MixedSet, Set

__all__ = ["MixedDataSet", "MixedMap", "MixedDat"]


class MixedDataSet:
    """A tuple of DataSets on the components of a MixedSet.

    ``MixedDataSet(mixed_set, dims)`` with ``dims`` one int for every
    component or one per component, or ``MixedDataSet(iterable)`` of
    DataSets / Sets (a Set becomes a DataSet of dim 1); ``mixed_set ** dim``
    is the usual spelling.
    """

    def __init__(self, arg, dims=None):
        if isinstance(arg, MixedSet):
            sets = arg.split
            if dims is None:
                dims = 1
            if isinstance(dims, numbers.Integral):
                dims = (dims,) * len(sets)
            dims = tuple(dims)
            if len(dims) != len(sets):
                raise ValueError("Got MixedSet of %d Sets but %s dims" % (len(sets), len(dims)))
            dsets = tuple(s ** d for s, d in zip(sets, dims))
        else:
            if dims is not None:
                raise ValueError("dims are only accepted together with a MixedSet")
            dsets = []
            for d in arg:
                if isinstance(d, DataSet):
                    dsets.append(d)
                elif isinstance(d, Set):
                    dsets.append(d ** 1)
                else:
                    raise ex.DataSetTypeError("All components of a MixedDataSet must be DataSets or Sets")
            dsets = tuple(dsets)
        if not dsets:
            raise ValueError("A MixedDataSet needs at least one component")
        self._dsets = dsets
        self.comm = dsets[0].comm

    @utils.cached_property
    def split(self):
        return self._dsets

    @utils.cached_property
    def set(self):
        return MixedSet(tuple(d.set for d in self._dsets))

    @utils.cached_property
    def dim(self):
        return tuple(d.dim for d in self._dsets)

    @utils.cached_property
    def cdim(self):
        return sum(d.cdim for d in self._dsets)

    @property
    def halo(self):
        return None

    def __getitem__(self, idx):
        return self._dsets[idx]

    def __iter__(self):
        return iter(self._dsets)

    def __len__(self):
        return len(self._dsets)

    def __eq__(self, other):
        return type(self) is type(other) and self._dsets == other._dsets

    def __hash__(self):
        return hash(self._dsets)

    def __str__(self):
        return "OP2 MixedDataSet composed of DataSets: %s" % (self._dsets,)

    def __repr__(self):
        return "MixedDataSet(%r)" % (self._dsets,)


class MixedMap:
    """A tuple of Maps from one iteration set to the components of a MixedSet."""

    def __init__(self, maps):
        from pyop2_amd.types.map import Map
        maps = tuple(maps)
        if not maps or not all(isinstance(m, Map) for m in maps):
            raise ex.MapTypeError("All components of a MixedMap must be Maps")
        if len({m.iterset for m in maps}) != 1:
            raise ex.MapTypeError("All components of a MixedMap need to have the same iterset")
        self._maps = maps
        self.comm = maps[0].comm

    @utils.cached_property
    def split(self):
        return self._maps

    @utils.cached_property
    def iterset(self):
        return self._maps[0].iterset

    @utils.cached_property
    def toset(self):
        return MixedSet(tuple(m.toset for m in self._maps))

    @utils.cached_property
    def arity(self):
        return sum(m.arity for m in self._maps)

    @utils.cached_property
    def arities(self):
        return tuple(m.arity for m in self._maps)

    @utils.cached_property
    def arange(self):
        return (0,) + tuple(np.cumsum(self.arities))

    @utils.cached_property
    def values(self):
        return np.hstack([m.values for m in self._maps])

    @utils.cached_property
    def values_with_halo(self):
        return np.hstack([m.values_with_halo for m in self._maps])

    @utils.cached_property
    def name(self):
        return tuple(m.name for m in self._maps)

    @utils.cached_property
    def flattened_maps(self):
        out = []
        for m in self._maps:
            out.extend(m.flattened_maps)
        return tuple(out)

    @property
    def _kernel_args_(self):
        return sum((m._kernel_args_ for m in self._maps), ())

    def __getitem__(self, idx):
        return self._maps[idx]

    def __iter__(self):
        return iter(self._maps)

    def __len__(self):
        return len(self._maps)

    def __le__(self, o):
        return self == o

    def __str__(self):
        return "OP2 MixedMap composed of Maps: %s" % (self._maps,)

    def __repr__(self):
        return "MixedMap(%r)" % (self._maps,)


class MixedDat:
    """A tuple of Dats on the components of a MixedDataSet.

    ``MixedDat(mixed_dataset)`` or ``MixedDat(mixed_set)`` allocates zeroed
    component Dats; ``MixedDat(iterable)`` takes Dats (or DataSets / Sets,
    which are allocated).
    """

    _modes = [Access.READ, Access.WRITE, Access.RW, Access.INC, Access.MIN, Access.MAX]

    def __init__(self, mdset_or_dats, name=None):
        from pyop2_amd.types.dat import Dat
        if isinstance(mdset_or_dats, MixedSet):
            mdset_or_dats = MixedDataSet(mdset_or_dats)
        if isinstance(mdset_or_dats, MixedDataSet):
            dats = tuple(Dat(d) for d in mdset_or_dats.split)
        else:
            dats = tuple(d if isinstance(d, Dat) else Dat(d) for d in mdset_or_dats)
        if not dats:
            raise ValueError("A MixedDat needs at least one component")
        if len({d.dtype for d in dats}) != 1:
            raise ex.DataValueError("MixedDat with different dtypes is not supported")
        self._dats = dats
        self.comm = dats[0].comm
        self._name = name or "mdat_#x%x" % id(self)

    @utils.cached_property
    def split(self):
        return self._dats

    @utils.cached_property
    def dataset(self):
        return MixedDataSet(tuple(d.dataset for d in self._dats))

    @property
    def dtype(self):
        return self._dats[0].dtype

    @property
    def name(self):
        return self._name

    @property
    def data(self):
        return tuple(d.data for d in self._dats)

    @property
    def data_ro(self):
        return tuple(d.data_ro for d in self._dats)

    @property
    def data_with_halos(self):
        return tuple(d.data_with_halos for d in self._dats)

    @property
    def data_ro_with_halos(self):
        return tuple(d.data_ro_with_halos for d in self._dats)

    @property
    def nbytes(self):
        return sum(d.nbytes for d in self._dats)

    @property
    def halo_valid(self):
        return all(d.halo_valid for d in self._dats)

    @halo_valid.setter
    def halo_valid(self, val):
        for d in self._dats:
            d.halo_valid = val

    @property
    def _kernel_args_(self):
        return sum((d._kernel_args_ for d in self._dats), ())

    def _device_modified(self):
        for d in self._dats:
            d._device_modified()

    def increment_dat_version(self):
        for d in self._dats:
            d.increment_dat_version()

    @property
    def dat_version(self):
        return sum(d.dat_version for d in self._dats)

    def zero(self, subset=None):
        for d in self._dats:
            d.zero(subset)

    def inner(self, other):
        return sum(a.inner(b) for a, b in zip(self._dats, other.split))

    @property
    def norm(self):
        return self.inner(self) ** 0.5

    def _has_halo(self):
        return any(d._has_halo() for d in self._dats)

    def global_to_local_begin(self, access_mode):
        for d in self._dats:
            d.global_to_local_begin(access_mode)

    def global_to_local_end(self, access_mode):
        for d in self._dats:
            d.global_to_local_end(access_mode)

    def local_to_global_begin(self, insert_mode):
        for d in self._dats:
            d.local_to_global_begin(insert_mode)

    def local_to_global_end(self, insert_mode):
        for d in self._dats:
            d.local_to_global_end(insert_mode)

    @utils.validate_in(('access', _modes, ex.ModeValueError))
    def __call__(self, access, path=None):
        from pyop2_amd.parloop import MixedDatLegacyArg
        if path is None:
            raise NotImplementedError("direct loops over a MixedDat are not supported: give a MixedMap")
        if not isinstance(path, MixedMap):
            raise ex.MapTypeError("A MixedDat is accessed through a MixedMap")
        if conf.configuration["type_check"]:
            if len(path) != len(self._dats) or any(m.toset != d.dataset.set
                                                   for m, d in zip(path, self._dats)):
                raise ex.MapValueError("To Set of MixedMap does not match Set of MixedDat.")
        return MixedDatLegacyArg(self, path, access)

    def __getitem__(self, idx):
        return self._dats[idx]

    def __iter__(self):
        return iter(self._dats)

    def __len__(self):
        return len(self._dats)

    def __str__(self):
        return "OP2 MixedDat composed of Dats: %s" % (self._dats,)

    def __repr__(self):
        return "MixedDat(%r)" % (self._dats,)
