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
        # extruded maps: entry i steps by offset[i] per layer (map.py:33-60)
        if offset is None or len(offset) == 0:
            self._offset = None
        else:
            off = np.asarray(offset, dtype=IntType).reshape(-1)
            self._offset = np.broadcast_to(off, (self._arity,)).copy() if off.size == 1 else off
            if self._offset.shape != (self._arity,):
                raise ex.DataValueError("Map offset must have one entry per map entry")
        # periodic extrusion: entry i of the layer l cells above the bottom
        # is offset by offset[i] * (rem(l + q[i]) - rem(q[i])) (map.py:49-52)
        if offset_quotient is None or len(offset_quotient) == 0:
            self._offset_quotient = None
        else:
            self._offset_quotient = utils.verify_reshape(offset_quotient, IntType, (self._arity,))
        self._cache = {}
        self._dev = None

    @utils.cached_property
    def _global_kernel_arg(self):
        from pyop2_amd.global_kernel import MapKernelArg
        off = tuple(int(x) for x in self._offset) if self._offset is not None else None
        quo = tuple(int(x) for x in self._offset_quotient) if self._offset_quotient is not None else None
        return MapKernelArg(self.arity, off, quo)

    def _layer_offset(self, steps, ncell, periodic):
        """Layer offsets of the entries: ``steps`` (..., ) the cells above
        the column's bottom (the cell above counted for interior facets),
        ``ncell`` broadcast alike; returns (..., arity).  Periodic columns
        wrap with the generated loop's remainder ``a < n ? a : a - n``
        (``codegen/builder.py:27-30, 103-121``)."""
        steps = steps[..., None]
        if not periodic:
            return np.broadcast_to(steps, steps.shape[:-1] + (self._arity,))
        n = ncell[..., None]

        def rem(a):
            return np.where(a < n, a, a - n)
        if self._offset_quotient is None:
            return np.broadcast_to(rem(steps), steps.shape[:-1] + (self._arity,))
        q = self._offset_quotient.astype(np.int64)
        return rem(steps + q) - rem(np.broadcast_to(q, steps.shape[:-1] + (self._arity,)))

    def _column_entries(self, iterset, cols, region):
        """Entries and liveness of ``cols`` over ``region``: (len(cols),
        nsteps, nf, arity) int64 entries and bools (a step the column runs
        and a non-negative base)."""
        from pyop2_amd.types import region as reg
        periodic = bool(getattr(iterset, "_extruded_periodic", False))
        lay = iterset.column_layers(cols)
        ncell = (lay[:, 1] - lay[:, 0] - 1) if len(cols) else np.zeros(0, dtype=np.int64)
        lo, hi = reg.layer_steps(ncell, region, periodic)
        nsteps = hi - lo
        ns = int(nsteps.max()) if len(cols) else 0
        nf = reg.facet_factor(region)
        base = self._table[cols].astype(np.int64)
        off = self._offset if self._offset is not None else np.zeros(self._arity, dtype=IntType)
        step = lo[:, None, None] + np.arange(ns, dtype=np.int64)[None, :, None] + np.arange(nf)[None, None, :]
        nc = np.broadcast_to(ncell.astype(np.int64)[:, None, None], step.shape)
        out = base[:, None, None, :] + self._layer_offset(step, nc, periodic) * off.astype(np.int64)
        live = (np.arange(ns)[None, :] < nsteps[:, None])[:, :, None, None] & (base[:, None, None, :] >= 0)
        return out, np.broadcast_to(live, out.shape)

    def column_table(self, iterset, columns, region=None):
        """Entries of ``columns`` (host ids) of an extruded loop over the
        layers ``region`` runs (all cells by default): ``(len(columns),
        nsteps * nf * arity)``; entry (c, s, f, i) = ``table[c][i] +
        offset[i] * (lo_c + s + f)`` for the column's steps s (wrapped on a
        periodic set), -1 past them (``nf`` = 2 for interior facets: the
        cell below, then the one above).  These are the targets a column
        touches, for colouring and patterns."""
        cols = np.asarray(columns, dtype=np.int64)
        out, live = self._column_entries(iterset, cols, region)
        out = np.where(live, out, -1)
        return np.ascontiguousarray(out.reshape(len(cols), -1), dtype=IntType)

    def cell_table(self, iterset, columns, region=None):
        """Entries of every iteration of ``columns`` of an extruded loop over
        ``region``, one row per (column, layer), columns in order, layers
        bottom to top (``nf * arity`` entries a row)."""
        from pyop2_amd.types import region as reg
        nf = reg.facet_factor(region)
        t = self.column_table(iterset, columns, region).reshape(len(columns), -1, nf * self._arity)
        return np.ascontiguousarray(t[self.column_cells_mask(iterset, columns, region)], dtype=IntType)

    def check_extruded(self, iterset, region=None):
        """Raise MapValueError unless every entry an extruded loop over
        ``region`` of ``iterset`` reaches through this map lies in the
        target set.  A Subset's loop reaches only its chosen columns, so only
        those are checked; the result is kept per iteration set contents and
        region, and columns are checked in chunks."""
        from pyop2_amd.types import region as reg
        from pyop2_amd.types.set import Subset
        key = ("extruded-checked", iterset._plan_key(), int(reg.normalise(region)))
        if self._offset is None or key in self._cache:
            return
        n = self._toset.total_size
        if isinstance(iterset, Subset):
            columns_of, cols = iterset.superset, np.asarray(iterset.indices, dtype=np.int64)
        else:
            columns_of, cols = iterset, np.arange(iterset.total_size, dtype=np.int64)
        lay = columns_of.column_layers(cols)
        per_col = max(1, int((lay[:, 1] - lay[:, 0]).max(initial=1)) * 2 * self._arity)
        chunk = max(1, (1 << 22) // per_col)
        for c0 in range(0, len(cols), chunk):
            out, live = self._column_entries(columns_of, cols[c0:c0 + chunk], region)
            e = out[live]
            if e.size and (e.min() < 0 or e.max() >= n):
                raise ex.MapValueError(f"{self} with offsets reaches entries outside [0, {n}) "
                                       f"in the layers of {iterset}")
        self._cache[key] = True

    @staticmethod
    def column_cells_mask(iterset, columns, region=None):
        from pyop2_amd.types import region as reg
        cols = np.asarray(columns, dtype=np.int64)
        lay = iterset.column_layers(cols)
        lo, hi = reg.layer_steps(lay[:, 1] - lay[:, 0] - 1, region,
                                 bool(getattr(iterset, "_extruded_periodic", False)))
        nsteps = hi - lo
        ns = int(nsteps.max()) if len(cols) else 0
        return np.arange(ns)[None, :] < nsteps[:, None]

    @property
    def _table(self):
        """The index table the loops gather through: ``(iterset.total_size,
        arity)`` int32 (a permuted or composed map computes its own); no
        rows for a map built without values."""
        v = self._values
        return v.reshape(self.shape) if v.size == self.shape[0] * self.shape[1] else v.reshape(0, self._arity)

    @property
    def device_values(self):
        """Device copy (int32) of the index table, range-checked once:
        entries must be below ``toset.total_size`` (negative entries mean
        "no target" and are skipped by the loops)."""
        if self._dev is None:
            from pyop2_amd import device as dev
            v = np.ascontiguousarray(self._table, dtype=IntType)
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


class PermutedMap(Map):
    """A :class:`Map` with a constant permutation of each element's entries
    (``pyop2/types/map.py:171-213``): element data is staged as
    ``local[i] = global[map[permutation[i]]]``.  ``values`` are the base
    map's, as in the reference; the loops gather through the permuted table,
    which is built once (one gather per entry, no permutation in the loop)."""

    def __init__(self, map_, permutation):
        if not isinstance(map_, Map):
            raise TypeError("map_ must be a Map instance")
        if isinstance(map_, ComposedMap):
            raise NotImplementedError("PermutedMap of ComposedMap not implemented: simply permute before composing")
        perm = np.asarray(permutation, dtype=Map.dtype)
        if not np.array_equal(np.sort(perm), np.arange(map_.arity, dtype=Map.dtype)):
            raise ValueError(f"permutation {list(perm)} is not a permutation of range({map_.arity})")
        self.map_ = map_
        self.permutation = perm
        self._iterset = map_.iterset
        self._toset = map_.toset
        self.comm = map_.comm
        self._arity = map_.arity
        self._values = map_._values
        self.shape = map_.shape
        self._name = "pmap_#x%x" % id(self)
        # the permuted entries keep their own per-layer offsets; the offset
        # quotients stay as the base map's (builder.py:148-172)
        self._offset = map_._offset[perm] if map_._offset is not None else None
        self._offset_quotient = map_._offset_quotient
        self._cache = {}
        self._dev = None

    @utils.cached_property
    def _table(self):
        return np.ascontiguousarray(self.map_._table[:, self.permutation])

    def __str__(self):
        return "OP2 PermutedMap of %s with permutation %s" % (self.map_, list(self.permutation))

    def __repr__(self):
        return "PermutedMap(%r, %r)" % (self.map_, list(self.permutation))


class ComposedMap(Map):
    """Composition of maps (``pyop2/types/map.py:216-287``): element data is
    staged as ``local[i] = global[maps_[0][maps_[1][...[i]]]]``; every map but
    the first has arity 1.  The composed table is built once; a negative entry
    anywhere along the chain makes the composed entry negative (no target)."""

    def __init__(self, *maps_, name=None):
        if not all(isinstance(m, Map) for m in maps_):
            raise TypeError("All maps must be Map instances")
        for tomap, frommap in zip(maps_[:-1], maps_[1:]):
            if tomap.iterset is not frommap.toset:
                raise ex.MapTypeError("tomap.iterset must match frommap.toset")
            if tomap.comm is not frommap.comm:
                raise ex.MapTypeError("All maps needs to share a communicator")
            if frommap.arity != 1:
                raise ex.MapTypeError("frommap.arity must be 1")
        self._iterset = maps_[-1].iterset
        self._toset = maps_[0].toset
        self.comm = self._toset.comm
        self._arity = maps_[0].arity
        self._values = None
        self.shape = (self._iterset.total_size, self._arity)
        self._name = name or "cmap_#x%x" % id(self)
        # the first map carries the layer stepping (builder.py:175-193)
        self._offset = maps_[0]._offset
        self._offset_quotient = maps_[0]._offset_quotient
        self._cache = {}
        self._dev = None
        self.maps_ = tuple(maps_)

    @utils.cached_property
    def _table(self):
        if any(len(m._table) != m.iterset.total_size for m in self.maps_):
            return np.zeros((0, self._arity), dtype=Map.dtype)
        idx = np.asarray(self.maps_[-1]._table[:, 0], dtype=np.int64)
        for m in reversed(self.maps_[1:-1]):
            t = m._table[:, 0]
            idx = np.where(idx >= 0, t[np.clip(idx, 0, None)], -1)
        first = self.maps_[0]._table
        out = first[np.clip(idx, 0, None)].copy() if len(first) else np.zeros(self.shape, Map.dtype)
        out[idx < 0] = -1
        return np.ascontiguousarray(out, dtype=Map.dtype).reshape(self.shape)

    @utils.cached_property
    def values(self):
        raise RuntimeError("ComposedMap does not store values directly")

    @utils.cached_property
    def values_with_halo(self):
        raise RuntimeError("ComposedMap does not store values directly")

    @utils.cached_property
    def flattened_maps(self):
        out = []
        for m in self.maps_:
            out.extend(m.flattened_maps)
        return tuple(out)

    def __str__(self):
        return "OP2 ComposedMap of Maps: [%s]" % ",".join([str(m) for m in self.maps_])

    def __repr__(self):
        return "ComposedMap(%s)" % ",".join([repr(m) for m in self.maps_])

    def __le__(self, o):
        raise NotImplementedError("__le__ not implemented for ComposedMap")
