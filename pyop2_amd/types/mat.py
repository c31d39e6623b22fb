"""Sparsity and Mat on the device (``pyop2/types/mat.py``).

``Sparsity`` keeps the reference's argument processing, identity cache and
``nnz``/``onnz`` semantics (``mat.py:26-291``), but the pattern is built on
the device by the op2hip sparsity builder (the reference fills a PETSc
preallocator, ``sparsity.pyx:121-284``): the sorted unique union of every
``rmap[e, :] x cmap[e, :]`` (owned e) plus the diagonal.

``Mat`` is a device CSR matrix, or BSR with ``dim x dim`` blocks when the
reference would create a PETSc BAIJ (``rdim == cdim > 1`` and
``block_sparsity``, ``mat.py:685-699``), pre-filled with zeros on the exact
pattern.  Insertions outside the pattern are counted on the device and
raised as :class:`PatternError` by ``assemble()`` (PETSc's
``NEW_NONZERO_LOCATION_ERR``).
"""
import ctypes
import itertools
from collections.abc import Sequence

import numpy as np

from pyop2_amd import caching, configuration as conf, exceptions as ex, utils
from pyop2_amd.types.access import Access
from pyop2_amd.types.data_carrier import DataCarrier
from pyop2_amd.types.dataset import DataSet
from pyop2_amd.types import region as reg
from pyop2_amd.types.map import Map
from pyop2_amd.types.region import ALL
from pyop2_amd.types.mixed import MixedDataSet, MixedMap
from pyop2_amd.types.set import MixedSet, Set, Subset


def _cells_map(m, region=None):
    """A map of an extruded loop over ``region`` as a plain map over its
    iterations (owned columns in order, layers bottom to top; twice the
    arity for interior facets), cached on ``m``: the sparsity of an extruded
    pair is the union over the iterations (``sparsity.pyx:285-393``)."""
    from pyop2_amd.types import region as reg
    from pyop2_amd.types.set import Set
    region = reg.normalise(region)
    key = ("cells_map", int(region))
    f = m._cache.get(key)
    if f is None:
        it = m.iterset
        tab = m.cell_table(it, np.arange(it.size), region)
        f = Map(Set(len(tab)), m.toset, reg.facet_factor(region) * m.arity, tab)
        m._cache[key] = f
    return f


def _dof_space(n):
    """A plain Set of ``n`` dofs (the target of a dof-level map)."""
    from pyop2_amd.types.set import Set
    return Set(int(n))


def _dof_map(m, dim, offset):
    """``m`` unwound to dofs: entry (e, k * dim + a) = offset + m[e][k] * dim
    + a, or -1 where m[e][k] is negative (no target)."""
    t = m._table.astype(np.int64)
    d = offset + t[:, :, None] * dim + np.arange(dim)[None, None, :]
    d = np.where(t[:, :, None] >= 0, d, -1).reshape(t.shape[0], -1)
    return Map(m.iterset, _dof_space(offset + m.toset.total_size * dim), m.arity * dim,
               np.ascontiguousarray(d, dtype=np.int32))


class Sparsity(caching.ObjectCached):
    """Non-zero structure of a matrix from pairs of Maps.

    ``Sparsity((row_dset, col_dset), [(rmap, cmap, None), ...])`` or the
    dict form ``{(0, 0): [(rmap, cmap, None)]}``.
    """

    def __init__(self, dsets, maps_and_regions, name=None, nest=None, block_sparse=None,
                 diagonal_block=True):
        if self._initialized:
            return
        self._dsets = dsets
        self._maps_and_regions = maps_and_regions
        self._block_sparse = block_sparse
        self._diagonal_block = diagonal_block
        rset, cset = self.dsets
        self.comm = rset.comm
        self._has_diagonal = (rset == cset) and diagonal_block
        self._name = name or "sparsity_#x%x" % id(self)
        self._pattern = None
        self._dof_pattern = None
        if isinstance(rset, MixedDataSet):
            # one block Sparsity per (row, column) component (mat.py:86-114);
            # a monolithic Sparsity (nest=False) keeps them for their block
            # patterns and builds one dof-level pattern over the whole space
            blocks = [[Sparsity((rd, cd), sorted(maps_and_regions.get((i, j), ()),
                                                 key=lambda t: (id(t[0]), id(t[1]))),
                                name="%s_%d_%d" % (self._name, i, j), block_sparse=block_sparse,
                                diagonal_block=diagonal_block and i == j)
                       for j, cd in enumerate(cset.split)] for i, rd in enumerate(rset.split)]
            self._dims = tuple(tuple((rd.cdim, cd.cdim) for cd in cset.split) for rd in rset.split)
            self._nested = bool(nest)
            if self._nested:
                self._blocks = blocks
            else:
                self._blocks = [[self]]
                self._component_blocks = blocks
        else:
            self._nested = False
            self._blocks = [[self]]
            self._dims = (((rset.cdim, cset.cdim),),)
        self._initialized = True

    _cache = {}

    @classmethod
    @utils.validate_type(('name', str, ex.NameTypeError))
    def _process_args(cls, dsets, maps_and_regions, name=None, nest=None, block_sparse=None,
                      diagonal_block=True):
        if len(dsets) != 2:
            raise RuntimeError(f"dsets must be a tuple of two DataSets: got {dsets}")
        mixed = any(isinstance(d, MixedDataSet) for d in dsets)
        if mixed:
            dsets = tuple(d if isinstance(d, MixedDataSet) else MixedDataSet((d,)) for d in dsets)
        for dset in dsets:
            if not isinstance(dset, (DataSet, MixedDataSet)) and dset is not None:
                raise ex.DataSetTypeError("All data sets must be of type DataSet, not type %r" % type(dset))
        if isinstance(maps_and_regions, Sequence):
            maps_and_regions = {(0, 0): maps_and_regions}
        elif not isinstance(maps_and_regions, dict):
            raise TypeError(f"maps_and_regions must be dict or Sequence: got {type(maps_and_regions)}")
        nr, nc = (len(dsets[0]), len(dsets[1])) if mixed else (1, 1)
        processed = {(0, 0): frozenset()}
        for key, val in maps_and_regions.items():
            i, j = key
            if i >= nr or j >= nc:
                raise RuntimeError(f"(i, j) must be < {(nr, nc)}: got {(i, j)}")
            rds, cds = (dsets[0][i], dsets[1][j]) if mixed else dsets
            pv = set()
            for entry in set(val):
                rmap, cmap = entry[0], entry[1]
                iteration_regions = entry[2] if len(entry) > 2 else None
                if len(entry) > 3 and entry[3] is not None:
                    raise NotImplementedError("extra sparsity map data is not on the device path")
                for m in (rmap, cmap):
                    if not isinstance(m, Map):
                        raise ex.MapTypeError("All maps must be of type map, not type %r" % type(m))
                    if len(m._table) == 0 and m.iterset.total_size > 0:
                        raise ex.MapValueError("Unpopulated map values when trying to build sparsity.")
                if rmap.toset is not rds.set or cmap.toset is not cds.set:
                    raise RuntimeError("Map toset must be the same as DataSet set")
                if rmap.iterset is not cmap.iterset:
                    raise RuntimeError("Iterset of both maps in a pair must be the same")
                if iteration_regions is None:
                    iteration_regions = (ALL,)
                else:
                    iteration_regions = tuple(sorted(reg.normalise(r) for r in iteration_regions))
                pv.add((rmap, cmap, iteration_regions))
            if pv:
                processed[key] = frozenset(pv)
        cache = dsets[0].set
        if nest is None:
            nest = conf.configuration["matnest"]
        if block_sparse is None:
            block_sparse = conf.configuration["block_sparsity"]
        kwargs = {"name": name, "nest": nest, "block_sparse": block_sparse,
                  "diagonal_block": diagonal_block}
        return (cache,) + (tuple(dsets), processed), kwargs

    @classmethod
    def _cache_key(cls, dsets, maps_and_regions, name, nest, block_sparse, diagonal_block, *args, **kwargs):
        return (dsets, tuple(maps_and_regions.items()), nest, block_sparse)

    def __getitem__(self, idx):
        try:
            i, j = idx
            return self._blocks[i][j]
        except TypeError:
            return self._blocks[idx]

    @property
    def dsets(self):
        return self._dsets

    @utils.cached_property
    def rcmaps(self):
        return {key: [(r, c) for r, c, _ in sorted(val, key=lambda t: (id(t[0]), id(t[1])))]
                for key, val in self._maps_and_regions.items()}

    @utils.cached_property
    def iteration_regions(self):
        return {key: [regions for _, _, regions in val] for key, val in self._maps_and_regions.items()}

    @property
    def dims(self):
        return self._dims

    @utils.cached_property
    def shape(self):
        if self.monolithic:
            return (len(self._component_blocks), len(self._component_blocks[0]))
        return (len(self._blocks), len(self._blocks[0]))

    @property
    def monolithic(self):
        """A Sparsity of mixed DataSets stored as one matrix (nest=False)."""
        return hasattr(self, "_component_blocks")

    def component(self, i, j):
        """The Sparsity of block (i, j) of a mixed Sparsity, nested or not."""
        return self._component_blocks[i][j] if self.monolithic else self._blocks[i][j]

    @utils.cached_property
    def offsets(self):
        """Monolithic: the first dof row of each row component and the first
        dof column of each column component (the reference's local_ises,
        ``dataset.py:166-180``), and the totals."""
        rset, cset = self._dsets
        roff = np.concatenate([[0], np.cumsum([d.total_size * d.cdim for d in rset])]).astype(np.int64)
        coff = np.concatenate([[0], np.cumsum([d.total_size * d.cdim for d in cset])]).astype(np.int64)
        return roff, coff

    @property
    def nested(self):
        return self._nested

    @property
    def name(self):
        return self._name

    def __iter__(self):
        for row in self._blocks:
            for s in row:
                yield s

    def __str__(self):
        return "OP2 Sparsity: dsets %s, maps_and_regions %s, name %s" % (self._dsets, self._maps_and_regions, self._name)

    def __repr__(self):
        return "Sparsity(%r, %r, name=%r)" % (self.dsets, self._maps_and_regions, self.name)

    def __contains__(self, other):
        rm, cm = other[0], other[1]
        return any(r is rm and c is cm for r, c in self.rcmaps.get((0, 0), ()))

    # -- the pattern -----------------------------------------------------------
    @property
    def nrows(self):
        """Block rows (the row Set's elements); dof rows when monolithic."""
        if self.monolithic:
            return int(self.offsets[0][-1])
        return self._dsets[0].set.total_size

    @property
    def ncols(self):
        if self.monolithic:
            return int(self.offsets[1][-1])
        return self._dsets[1].set.total_size

    def _pair_maps(self, key):
        """The plain (rmap, cmap) pairs of block ``key``: an extruded pair
        becomes its per-cell maps over each of its regions."""
        pairs = []
        for rmap, cmap, regions in sorted(self._maps_and_regions.get(key, ()),
                                          key=lambda t: (id(t[0]), id(t[1]))):
            if getattr(rmap.iterset, "_extruded", False):
                pairs += [(_cells_map(rmap, r), _cells_map(cmap, r)) for r in regions]
            else:
                pairs.append((rmap, cmap))
        return pairs

    def _monolithic_pairs(self):
        """Dof-level map pairs of a monolithic Sparsity: every map entry of
        block (i, j) unwound to its ``dim`` dofs and shifted to its
        component's place, plus the diagonal node blocks of the diagonal
        blocks (``sparsity.pyx:141-160, 212-219``)."""
        rset, cset = self._dsets
        roff, coff = self.offsets
        pairs = []
        for (i, j) in sorted(self._maps_and_regions):
            for rmap, cmap in self._pair_maps((i, j)):
                pairs.append((_dof_map(rmap, rset[i].cdim, roff[i]), _dof_map(cmap, cset[j].cdim, coff[j])))
        if self._has_diagonal:
            for i in range(min(len(rset), len(cset))):
                n = min(rset[i].total_size, cset[i].total_size)
                nodes = np.arange(n, dtype=np.int64)[:, None]
                it = Set(n)
                rows = roff[i] + nodes * rset[i].cdim + np.arange(rset[i].cdim)[None, :]
                cols = coff[i] + nodes * cset[i].cdim + np.arange(cset[i].cdim)[None, :]
                pairs.append((Map(it, _dof_space(roff[-1]), rset[i].cdim, rows.astype(np.int32)),
                              Map(it, _dof_space(coff[-1]), cset[i].cdim, cols.astype(np.int32))))
        return pairs

    def _build(self):
        if self._pattern is not None:
            return self._pattern
        from pyop2_amd import device as dev, plans
        L = dev.lib()
        s = dev.stream()
        if self.monolithic:
            pairs = self._monolithic_pairs()
            set_diag = False
        else:
            pairs = self._pair_maps((0, 0))
            set_diag = self._has_diagonal
        nrows, ncols = self.nrows, self.ncols
        arr = (dev.MapPair * max(len(pairs), 1))()
        incp, ince = [], []
        for k, (rmap, cmap) in enumerate(pairs):
            size = rmap.iterset.size
            tr = plans.transpose(rmap, nrows, size)
            incp.append(tr.inc_ptr)
            ince.append(tr.inc_ent)
            arr[k] = dev.MapPair(dev.ptr(rmap.device_values), dev.ptr(cmap.device_values),
                                 rmap.arity, cmap.arity, size)
        ip = (ctypes.c_void_p * max(len(pairs), 1))(*[dev.ptr(t) for t in incp])
        ie = (ctypes.c_void_p * max(len(pairs), 1))(*[dev.ptr(t) for t in ince])
        rowptr = dev.empty((nrows + 1,), np.int32)
        need = ctypes.c_size_t()
        dev.check(L.op2hip_scan_workspace(nrows, ctypes.byref(need)), "scan_workspace")
        work = dev.workspace(need.value + 256 + 4 * (nrows + 64))
        nnz = ctypes.c_int64()
        dev.check(L.op2hip_sparsity_count(nrows, ncols, arr, len(pairs), ip, ie,
                                          int(set_diag), dev.ptr(rowptr),
                                          ctypes.byref(nnz), dev.ptr(work), work.numel(), s),
                  "sparsity_count")
        colidx = dev.empty((max(nnz.value, 1),), np.int32)
        dev.check(L.op2hip_sparsity_fill(nrows, ncols, arr, len(pairs), ip, ie,
                                         int(set_diag), dev.ptr(rowptr),
                                         dev.ptr(colidx), s), "sparsity_fill")
        self._pattern = (rowptr, colidx[:nnz.value], int(nnz.value))
        return self._pattern

    @property
    def block_pattern(self):
        """Device ``(rowptr, colidx)`` of the block pattern (int32)."""
        rowptr, colidx, _ = self._build()
        return rowptr, colidx

    @property
    def block_nnz(self):
        return self._build()[2]

    def dof_pattern(self):
        """Device ``(rowptr, colidx)`` of the dof (scalar) pattern."""
        rdim, cdim = self._dims[0][0]
        if (rdim == 1 and cdim == 1) or self.monolithic:
            return self.block_pattern
        if self._dof_pattern is None:
            from pyop2_amd import device as dev
            rowptr, colidx, nnz = self._build()
            rp = dev.empty((self.nrows * rdim + 1,), np.int32)
            ci = dev.empty((max(nnz * rdim * cdim, 1),), np.int32)
            dev.check(dev.lib().op2hip_pattern_expand(self.nrows, dev.ptr(rowptr), dev.ptr(colidx),
                                                      rdim, cdim, dev.ptr(rp), dev.ptr(ci),
                                                      dev.stream()), "pattern_expand")
            self._dof_pattern = (rp, ci[:nnz * rdim * cdim])
        return self._dof_pattern

    @utils.cached_property
    def nnz(self):
        """Non-zeros per row as the reference reports them (block rows when
        block-sparse with square blocks, else dof rows; a monolithic
        Sparsity's pattern is at dof level already)."""
        from pyop2_amd import device as dev
        rowptr, _, _ = self._build()
        nnz = np.diff(dev.to_host(rowptr)).astype(np.int32)
        rdim, cdim = self._dims[0][0]
        if not self.monolithic and not (self._block_sparse and rdim == cdim):
            nnz = np.repeat(nnz * cdim, rdim)
        return nnz

    @utils.cached_property
    def onnz(self):
        return np.zeros_like(self.nnz)

    @utils.cached_property
    def nz(self):
        return int(self.nnz.sum())

    @utils.cached_property
    def onz(self):
        return 0


class Mat(DataCarrier):
    """Device matrix on a Sparsity (CSR, or BSR for square vector blocks)."""

    ASSEMBLED = "ASSEMBLED"
    INSERT_VALUES = "INSERT_VALUES"
    ADD_VALUES = "ADD_VALUES"

    _modes = [Access.WRITE, Access.INC]

    def __new__(cls, sparsity=None, *args, **kwargs):
        if cls is Mat and getattr(sparsity, "nested", False):
            cls = MixedMat
        elif cls is Mat and getattr(sparsity, "monolithic", False):
            cls = MonolithicMat
        return super().__new__(cls)

    @utils.validate_type(('sparsity', Sparsity, ex.SparsityTypeError),
                         ('name', str, ex.NameTypeError))
    def __init__(self, sparsity, dtype=None, name=None, mat_type=None):
        self._sparsity = sparsity
        self.comm = sparsity.comm
        self._datatype = np.dtype(dtype or np.float64)
        if self._datatype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise RuntimeError("Can only create a matrix of type float64 (or float32), %s is not supported"
                               % self._datatype)
        self._name = name or "mat_#x%x" % id(self)
        self.mat_type = mat_type
        rdim, cdim = self.dims[0][0]
        self._block = bool(rdim == cdim and rdim > 1 and sparsity._block_sparse)
        self._values = None
        self._status = None
        self._desc = None
        self.assembly_state = Mat.ASSEMBLED

    def _init(self):
        if self._values is not None:
            return
        from pyop2_amd import device as dev
        rdim, cdim = self.dims[0][0]
        if self._block:
            rowptr, colidx = self._sparsity.block_pattern
            bs = rdim * cdim
            nrows, ncols, brd, bcd = self._sparsity.nrows, self._sparsity.ncols, rdim, cdim
        else:
            rowptr, colidx = self._sparsity.dof_pattern()
            bs = 1
            nrows, ncols, brd, bcd = self._sparsity.nrows * rdim, self._sparsity.ncols * cdim, 1, 1
        nnz = colidx.numel()
        self._rowptr, self._colidx = rowptr, colidx
        self._values = dev.zeros((max(nnz * bs, 1),), self._datatype)
        self._status = dev.zeros((1,), np.int32)
        self._nnz = nnz
        self._bs = bs
        d = dev.MatDesc()
        d.rowptr, d.colidx, d.values = dev.ptr(rowptr), dev.ptr(colidx), dev.ptr(self._values)
        d.nrows, d.ncols, d.rdim, d.cdim = nrows, ncols, brd, bcd
        d.status = dev.ptr(self._status)
        d.inc_ptr = d.inc_ent = None
        self._desc = d

    # -- device views -------------------------------------------------------------
    @property
    def rowptr(self):
        self._init()
        return self._rowptr

    @property
    def colidx(self):
        self._init()
        return self._colidx

    @property
    def device_values(self):
        """Device values tensor: nnz blocks of ``bs`` entries (row-major)."""
        self._init()
        return self._values

    @property
    def block_shape(self):
        rdim, cdim = self.dims[0][0]
        return (rdim, cdim) if self._block else (1, 1)

    def descriptor(self, inc_ptr=None, inc_ent=None):
        """``op2hip_mat`` for the C ABI (optionally with row incidences)."""
        from pyop2_amd import device as dev
        self._init()
        d = dev.MatDesc.from_buffer_copy(self._desc)
        d.inc_ptr = dev.ptr(inc_ptr) if inc_ptr is not None else None
        d.inc_ent = dev.ptr(inc_ent) if inc_ent is not None else None
        return d

    @property
    def _kernel_args_(self):
        self._init()
        return (ctypes.addressof(self._desc),)

    # -- API ------------------------------------------------------------------------
    @utils.validate_in(('access', _modes, ex.ModeValueError))
    def __call__(self, access, path, lgmaps=None, unroll_map=False):
        from pyop2_amd.parloop import MatLegacyArg
        path_maps = utils.as_tuple(path, Map, 2)
        if conf.configuration["type_check"] and tuple(path_maps) not in self.sparsity:
            raise ex.MapValueError("Path maps not in sparsity maps")
        if lgmaps is not None:
            raise NotImplementedError("lgmaps are not supported on the device path")
        return MatLegacyArg(self, path_maps, access, lgmaps, unroll_map)

    @property
    def sparsity(self):
        return self._sparsity

    @property
    def dims(self):
        return self._sparsity._dims

    @property
    def dtype(self):
        return self._datatype

    @property
    def nrows(self):
        return self._sparsity.nrows * self.dims[0][0][0]

    @property
    def ncols(self):
        return self._sparsity.ncols * self.dims[0][0][1]

    @property
    def nblock_rows(self):
        return self._sparsity.nrows

    @property
    def nblock_cols(self):
        return self._sparsity.ncols

    @property
    def nblocks(self):
        return 1

    @property
    def is_mixed(self):
        return False

    @property
    def nbytes(self):
        rdim, cdim = self.dims[0][0]
        mult = rdim * cdim if self._sparsity._block_sparse else 1
        return (self._sparsity.nz + self._sparsity.onz) * self.dtype.itemsize * mult

    def __iter__(self):
        yield self

    def __getitem__(self, idx):
        return self

    @property
    def blocks(self):
        return [[self]]

    def change_assembly_state(self, new_state):
        if new_state == Mat.ASSEMBLED or self.assembly_state == Mat.ASSEMBLED:
            self.assembly_state = new_state
        elif new_state != self.assembly_state:
            self._flush_assembly()
            self.assembly_state = new_state

    def _flush_assembly(self):
        """Device writes are ordered on the stream: nothing to flush."""

    def _check_status(self):
        from pyop2_amd import device as dev
        if self._status is None:
            return
        n = int(self._status.item())
        if n:
            self._status.zero_()
            raise ex.PatternError(f"{n} Mat insertions fell outside the sparsity pattern of {self.name}")

    def assemble(self):
        """Finalise the matrix (checks that every insertion hit the pattern)."""
        self._init()
        self._check_status()
        self.change_assembly_state(Mat.ASSEMBLED)

    def zero(self):
        """Zero every stored value (the pattern is kept)."""
        self._init()
        self._values.zero_()

    # -- entries set from the host (the reference's MatSetValues calls) ----------
    # A few entries change on the device; only their positions are found on
    # the host, from a host copy of the pattern kept after the first use.
    def _host_pattern(self):
        if getattr(self, "_hpattern", None) is None:
            from pyop2_amd import device as dev
            self._init()
            self._hpattern = (dev.to_host(self._rowptr).astype(np.int64),
                              dev.to_host(self._colidx).astype(np.int64))
        return self._hpattern

    def _positions(self, rows, cols):
        """Flat positions in the values array of the scalar (dof) entries
        ``(rows[k], cols[k])``; PatternError for an entry off the pattern."""
        rp, ci = self._host_pattern()
        rd, cd = self.block_shape
        rows = np.asarray(rows, dtype=np.int64).ravel()
        cols = np.asarray(cols, dtype=np.int64).ravel()
        br, ri = np.divmod(rows, rd)
        bc, cj = np.divmod(cols, cd)
        out = np.empty(len(rows), dtype=np.int64)
        for k in range(len(rows)):
            lo, hi = rp[br[k]], rp[br[k] + 1]
            s = lo + int(np.searchsorted(ci[lo:hi], bc[k]))
            if s >= hi or ci[s] != bc[k]:
                raise ex.PatternError(f"entry ({rows[k]}, {cols[k]}) is not in the sparsity pattern")
            out[k] = (s * rd + ri[k]) * cd + cj[k]
        return out

    def _write_entries(self, rows, cols, vals, add):
        """values[(r, c)] += v (add) or = v; among repeated entries an insert
        keeps the last, as successive MatSetValues calls do."""
        from pyop2_amd import device as dev
        vals = np.asarray(vals, dtype=self.dtype).ravel()
        if not len(vals):
            return
        pos = self._positions(rows, cols)
        if not add:
            last = len(pos) - 1 - np.unique(pos[::-1], return_index=True)[1]
            pos, vals = pos[last], vals[last]
        idx = dev.to_device(pos)
        self._values.index_put_((idx,), dev.to_device(vals), accumulate=add)

    def zero_rows(self, rows, diag_val=1.0):
        """Zero the dof rows ``rows`` (an iterable or a Subset), then set
        their diagonal entry to ``diag_val``."""
        from pyop2_amd import device as dev
        self.assemble()
        rows = rows.indices if isinstance(rows, Subset) else rows
        rows = np.unique(np.asarray(rows, dtype=np.int64).ravel())
        if not len(rows):
            return
        rp, ci = self._host_pattern()
        rd, cd = self.block_shape
        br, ri = np.divmod(rows, rd)
        spans = [np.arange(rp[b], rp[b + 1]) for b in br]
        pos = np.concatenate([((s * rd + r)[:, None] * cd + np.arange(cd)[None, :]).ravel()
                              for s, r in zip(spans, ri)])
        self._values.index_fill_(0, dev.to_device(pos), 0)
        on_pattern = [r for r in rows if r < self.ncols and self._has_entry(r, r)]
        self._write_entries(on_pattern, on_pattern, np.full(len(on_pattern), diag_val), add=False)

    def _has_entry(self, r, c):
        rp, ci = self._host_pattern()
        rd, cd = self.block_shape
        b, bc = r // rd, c // cd
        lo, hi = rp[b], rp[b + 1]
        s = lo + int(np.searchsorted(ci[lo:hi], bc))
        return s < hi and ci[s] == bc

    def set_local_diagonal_entries(self, rows, diag_val=1.0, idx=None):
        """Insert ``diag_val`` on the diagonal of the block rows ``rows``
        (every component, or component ``idx``)."""
        rows = np.asarray(rows, dtype=np.int64).ravel()
        rbs, _ = self.dims[0][0]
        if rbs > 1:
            comps = [idx] if idx is not None else range(rbs)
            rows = (rbs * rows[:, None] + np.asarray(list(comps))[None, :]).ravel()
        self.change_assembly_state(Mat.INSERT_VALUES)
        self._write_entries(rows, rows, np.full(len(rows), diag_val), add=False)

    def _dense_block_entries(self, rows, cols, values):
        """The dof entries of the dense block ``values`` at block ``rows`` x
        ``cols`` (MatSetValuesBlockedLocal's layout)."""
        rd, cd = self.dims[0][0]
        rows = np.atleast_1d(np.asarray(rows, dtype=np.int64))
        cols = np.atleast_1d(np.asarray(cols, dtype=np.int64))
        vals = np.asarray(values, dtype=self.dtype).reshape(len(rows) * rd, len(cols) * cd)
        drows = (rows[:, None] * rd + np.arange(rd)[None, :]).ravel()
        dcols = (cols[:, None] * cd + np.arange(cd)[None, :]).ravel()
        rr, cc = np.meshgrid(drows, dcols, indexing="ij")
        return rr.ravel(), cc.ravel(), vals.ravel()

    def addto_values(self, rows, cols, values):
        """Add the dense block ``values`` at block ``rows`` x ``cols``."""
        self.change_assembly_state(Mat.ADD_VALUES)
        if len(np.atleast_1d(values)):
            self._write_entries(*self._dense_block_entries(rows, cols, values), add=True)

    def set_values(self, rows, cols, values):
        """Insert the dense block ``values`` at block ``rows`` x ``cols``."""
        self.change_assembly_state(Mat.INSERT_VALUES)
        if len(np.atleast_1d(values)):
            self._write_entries(*self._dense_block_entries(rows, cols, values), add=False)

    def _host_csr(self):
        """Host ``(rowptr, colidx, values[nnz, rd, cd])`` copies."""
        from pyop2_amd import device as dev
        self._init()
        rd, cd = self.block_shape
        rp = dev.to_host(self._rowptr)
        ci = dev.to_host(self._colidx)
        v = dev.to_host(self._values)[:self._nnz * rd * cd].reshape(self._nnz, rd, cd).copy()
        return rp, ci, v

    def csr(self):
        """Host scalar CSR ``(rowptr, colidx, values)`` (BSR expanded)."""
        rp, ci, v = self._host_csr()
        rd, cd = self.block_shape
        if rd == 1 and cd == 1:
            return rp, ci, v.reshape(-1)
        nb = len(rp) - 1
        rows, cols, vals = [], [], []
        for r in range(nb):
            for s in range(rp[r], rp[r + 1]):
                for a in range(rd):
                    for b in range(cd):
                        rows.append(r * rd + a)
                        cols.append(ci[s] * cd + b)
                        vals.append(v[s, a, b])
        order = np.lexsort((np.array(cols), np.array(rows)))
        rows, cols, vals = np.array(rows)[order], np.array(cols)[order], np.array(vals)[order]
        srp = np.zeros(nb * rd + 1, dtype=np.int64)
        np.cumsum(np.bincount(rows, minlength=nb * rd), out=srp[1:])
        return srp, cols, vals

    @property
    def values(self):
        """Dense numpy copy (refused above 1e6 entries, as in the reference)."""
        self.assemble()
        if self.nrows * self.ncols > 1000000:
            raise ValueError("Printing dense matrix with more than 1 million entries not allowed.\n"
                             "Are you sure you wanted to do this?")
        rp, ci, v = self._host_csr()
        rd, cd = self.block_shape
        out = np.zeros((self.nrows, self.ncols), dtype=self.dtype)
        for r in range(len(rp) - 1):
            for s in range(rp[r], rp[r + 1]):
                out[r * rd:(r + 1) * rd, ci[s] * cd:(ci[s] + 1) * cd] += v[s]
        return out

    @property
    def dat_version(self):
        if self.assembly_state != Mat.ASSEMBLED:
            raise RuntimeError("Should not ask for state counter if the matrix is not assembled.")
        return 0

    def __str__(self):
        return "OP2 Mat: %s, sparsity (%s), datatype %s" % (self._name, self._sparsity, self._datatype.name)

    def __repr__(self):
        return "Mat(%r, %r, %r)" % (self._sparsity, self._datatype, self._name)


class MixedMat(Mat):
    """Matrix on a mixed Sparsity: one block Mat per (row, column)
    component (the reference's MatNest)."""

    def __init__(self, sparsity, dtype=None, name=None, mat_type=None):
        self._sparsity = sparsity
        self.comm = sparsity.comm
        self._datatype = np.dtype(dtype or np.float64)
        self._name = name or "mat_#x%x" % id(self)
        self.mat_type = mat_type
        nr, nc = sparsity.shape
        self._blocks = [[Mat(sparsity[i, j], self._datatype, "%s_%d_%d" % (self._name, i, j))
                         for j in range(nc)] for i in range(nr)]
        self._block = False
        self._values = None
        self._status = None
        self._desc = None
        self.assembly_state = Mat.ASSEMBLED

    def _init(self):
        for b in self:
            b._init()

    @property
    def _kernel_args_(self):
        return sum((b._kernel_args_ for b in self), ())

    @utils.validate_in(('access', Mat._modes, ex.ModeValueError))
    def __call__(self, access, path, lgmaps=None, unroll_map=False):
        from pyop2_amd.parloop import MixedMatLegacyArg
        if lgmaps is not None:
            raise NotImplementedError("lgmaps are not supported on the device path")
        rmap, cmap = path
        if not (isinstance(rmap, MixedMap) and isinstance(cmap, MixedMap)):
            raise ex.MapTypeError("A MixedMat is accessed through a pair of MixedMaps")
        nr, nc = self._sparsity.shape
        if len(rmap) != nr or len(cmap) != nc:
            raise ex.MapValueError("Path maps do not match the blocks of the Mat")
        if conf.configuration["type_check"]:
            for i in range(nr):
                for j in range(nc):
                    if (rmap[i], cmap[j]) not in self._sparsity.component(i, j):
                        raise ex.MapValueError("Path maps not in sparsity maps")
        return MixedMatLegacyArg(self, (rmap, cmap), access)

    @property
    def nblocks(self):
        nr, nc = self._sparsity.shape
        return nr * nc

    @property
    def is_mixed(self):
        return True

    @property
    def nrows(self):
        return sum(row[0].nrows for row in self._blocks)

    @property
    def ncols(self):
        return sum(b.ncols for b in self._blocks[0])

    @property
    def nbytes(self):
        return sum(b.nbytes for b in self)

    def __iter__(self):
        for row in self._blocks:
            for b in row:
                yield b

    def __getitem__(self, idx):
        try:
            i, j = idx
            return self._blocks[i][j]
        except TypeError:
            return self._blocks[idx]

    @property
    def blocks(self):
        return self._blocks

    def change_assembly_state(self, new_state):
        for b in self:
            b.change_assembly_state(new_state)
        self.assembly_state = new_state

    def assemble(self):
        for b in self:
            b.assemble()
        self.assembly_state = Mat.ASSEMBLED

    def zero(self):
        for b in self:
            b.zero()

    @property
    def values(self):
        """Dense values of the whole matrix (blocks stacked)."""
        return np.block([[b.values for b in row] for row in self._blocks])

    @property
    def dat_version(self):
        return sum(b.dat_version for b in self)

    def __str__(self):
        return "OP2 MixedMat %s on %s" % (self._name, self._sparsity)

    def __repr__(self):
        return "MixedMat(%r, %r, %r)" % (self._sparsity, self._datatype, self._name)


class MonolithicMat(Mat):
    """A Matrix on a monolithic mixed Sparsity (``nest=False``): one scalar
    CSR over the whole mixed dof space, component i's rows and columns
    contiguous at their offsets (``mat.py:613-654``).  ``mat[i, j]`` is a
    :class:`MatBlock` view through which a par_loop adds the element blocks
    of row component i and column component j."""

    def __init__(self, sparsity, dtype=None, name=None, mat_type=None):
        self._sparsity = sparsity
        self.comm = sparsity.comm
        self._datatype = np.dtype(dtype or np.float64)
        if self._datatype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise RuntimeError("Can only create a matrix of type float64 (or float32), %s is not supported"
                               % self._datatype)
        self._name = name or "mat_#x%x" % id(self)
        self.mat_type = mat_type
        self._block = False
        self._values = None
        self._status = None
        self._desc = None
        self.assembly_state = Mat.ASSEMBLED
        nr, nc = sparsity.shape
        self._blocks = [[MatBlock(self, i, j) for j in range(nc)] for i in range(nr)]

    def _init(self):
        if self._values is not None:
            return
        from pyop2_amd import device as dev
        rowptr, colidx = self._sparsity.dof_pattern()
        nnz = colidx.numel()
        self._rowptr, self._colidx = rowptr, colidx
        self._values = dev.zeros((max(nnz, 1),), self._datatype)
        self._status = dev.zeros((1,), np.int32)
        self._nnz, self._bs = nnz, 1
        d = dev.MatDesc()
        d.rowptr, d.colidx, d.values = dev.ptr(rowptr), dev.ptr(colidx), dev.ptr(self._values)
        d.nrows, d.ncols, d.rdim, d.cdim = self._sparsity.nrows, self._sparsity.ncols, 1, 1
        d.status = dev.ptr(self._status)
        d.inc_ptr = d.inc_ent = None
        self._desc = d

    @property
    def block_shape(self):
        return (1, 1)

    @property
    def nrows(self):
        return self._sparsity.nrows

    @property
    def ncols(self):
        return self._sparsity.ncols

    @property
    def nblock_rows(self):
        return self._sparsity.nrows

    @property
    def nblock_cols(self):
        return self._sparsity.ncols

    @property
    def nblocks(self):
        nr, nc = self._sparsity.shape
        return nr * nc

    @property
    def is_mixed(self):
        return True

    @property
    def nbytes(self):
        return self._sparsity.nz * self.dtype.itemsize

    def set_local_diagonal_entries(self, rows, diag_val=1.0, idx=None):
        """Insert ``diag_val`` on the diagonal of the dof rows ``rows``."""
        rows = np.asarray(rows, dtype=np.int64).ravel()
        self.change_assembly_state(Mat.INSERT_VALUES)
        self._write_entries(rows, rows, np.full(len(rows), diag_val), add=False)

    @utils.validate_in(('access', Mat._modes, ex.ModeValueError))
    def __call__(self, access, path, lgmaps=None, unroll_map=False):
        return MixedMat.__call__(self, access, path, lgmaps, unroll_map)

    def __iter__(self):
        for row in self._blocks:
            for b in row:
                yield b

    def __getitem__(self, idx):
        try:
            i, j = idx
            return self._blocks[i][j]
        except TypeError:
            return self._blocks[idx]

    @property
    def blocks(self):
        return self._blocks

    def __str__(self):
        return "OP2 monolithic Mat %s on %s" % (self._name, self._sparsity)

    def __repr__(self):
        return "Mat(%r, %r, %r)" % (self._sparsity, self._datatype, self._name)


class MatBlock:
    """Block (i, j) of a monolithic Mat (``mat.py:868-960``): the rows of
    row component i and the columns of column component j of its parent's
    storage.  Its assembly state is the parent's; assembling a block alone
    is an error, as in the reference."""

    _block = False

    def __init__(self, parent, i, j):
        self._parent = parent
        self._i, self._j = i, j
        roff, coff = parent.sparsity.offsets
        self.row_offset, self.col_offset = int(roff[i]), int(coff[j])
        self._sparsity = parent.sparsity.component(i, j)
        self.comm = parent.comm

    # storage and launch descriptors are the parent's
    @property
    def rowptr(self):
        return self._parent.rowptr

    @property
    def colidx(self):
        return self._parent.colidx

    @property
    def device_values(self):
        return self._parent.device_values

    @property
    def _status(self):
        self._parent._init()
        return self._parent._status

    @property
    def _gen_flag(self):
        """How a generated loop scatters into this block: the parent's scalar
        CSR, rows and columns shifted by the block's offsets."""
        return ("offset", self.row_offset, self.col_offset)

    @property
    def sparsity(self):
        return self._sparsity

    @property
    def dims(self):
        return self._sparsity.dims

    @property
    def dtype(self):
        return self._parent.dtype

    @property
    def block_shape(self):
        return (1, 1)

    @property
    def nrows(self):
        rd, _ = self.dims[0][0]
        return self._sparsity.nrows * rd

    @property
    def ncols(self):
        _, cd = self.dims[0][0]
        return self._sparsity.ncols * cd

    @property
    def assembly_state(self):
        return self._parent.assembly_state

    @assembly_state.setter
    def assembly_state(self, value):
        self._parent.assembly_state = value

    def change_assembly_state(self, new_state):
        self._parent.change_assembly_state(new_state)

    def _flush_assembly(self):
        self._parent._flush_assembly()

    def assemble(self):
        raise RuntimeError("Should never call assemble on MatBlock")

    def __iter__(self):
        yield self

    def __getitem__(self, idx):
        return self

    def _shifted(self, rows, cols):
        return (np.asarray(rows, dtype=np.int64).ravel() + self.row_offset,
                np.asarray(cols, dtype=np.int64).ravel() + self.col_offset)

    def _dense_block_entries(self, rows, cols, values):
        rr, cc, vv = Mat._dense_block_entries(self, rows, cols, values)
        return (*self._shifted(rr, cc), vv)

    def addto_values(self, rows, cols, values):
        self.change_assembly_state(Mat.ADD_VALUES)
        if len(np.atleast_1d(values)):
            self._parent._write_entries(*self._dense_block_entries(rows, cols, values), add=True)

    def set_values(self, rows, cols, values):
        self.change_assembly_state(Mat.INSERT_VALUES)
        if len(np.atleast_1d(values)):
            self._parent._write_entries(*self._dense_block_entries(rows, cols, values), add=False)

    def set_local_diagonal_entries(self, rows, diag_val=1.0, idx=None):
        rows = np.asarray(rows, dtype=np.int64).ravel()
        rbs, _ = self.dims[0][0]
        if rbs > 1:
            comps = [idx] if idx is not None else range(rbs)
            rows = (rbs * rows[:, None] + np.asarray(list(comps))[None, :]).ravel()
        r, c = self._shifted(rows, rows)
        self.change_assembly_state(Mat.INSERT_VALUES)
        self._parent._write_entries(r, c, np.full(len(rows), diag_val), add=False)

    @property
    def values(self):
        """Dense copy of the block (assembles the parent)."""
        full = self._parent.values
        return full[self.row_offset:self.row_offset + self.nrows,
                    self.col_offset:self.col_offset + self.ncols]

    def __str__(self):
        return "OP2 MatBlock (%d, %d) of %s" % (self._i, self._j, self._parent.name)

    def __repr__(self):
        return "MatBlock(%r, %d, %d)" % (self._parent, self._i, self._j)
