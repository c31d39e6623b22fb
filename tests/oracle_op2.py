"""A host-only stand-in for the ``op2`` namespace that executes par_loops with
the oracle (``oracle.seqwrap``): same call forms as ``pyop2_amd.op2`` for the
subset the known-answer cases use, numpy storage, reference semantics
(core part then owned part, zeroed temporaries for INC Globals)."""
import numpy as np

from oracle import seqwrap, sparsity as osp

READ, WRITE, RW, INC, MIN, MAX = "READ", "WRITE", "RW", "INC", "MIN", "MAX"
ON_BOTTOM, ON_TOP, ON_INTERIOR_FACETS, ALL = "bottom", "top", "interior", "all"


class MapTypeError(TypeError):
    pass


class Set:
    def __init__(self, size, name=None):
        if isinstance(size, (int, np.integer)):
            size = (size, size, size)
        self.core_size, self.size, self.total_size = (int(s) for s in size)
        self.name = name

    def __pow__(self, dim):
        return DataSet(self, dim)

    def __call__(self, indices):
        return Subset(self, indices)


class Subset(Set):
    """Elements ``indices`` of ``superset``; part sizes count the indices
    below the superset's part boundaries (``pyop2/types/set.py:394-450``).
    A subset of a subset is a subset of the outer set."""

    def __init__(self, superset, indices):
        indices = np.unique(np.asarray(indices, dtype=np.int32))
        if isinstance(superset, Subset):
            indices, superset = superset.indices[indices], superset.superset
        self.superset = superset
        self.indices = indices
        super().__init__(tuple(int((self.indices < b).sum())
                               for b in (superset.core_size, superset.size, superset.total_size)))


class MixedSet:
    def __init__(self, sets):
        self.sets = tuple(sets)

    def __iter__(self):
        return iter(self.sets)

    def __getitem__(self, i):
        return self.sets[i]

    def __len__(self):
        return len(self.sets)

    def __pow__(self, dim):
        return MixedDataSet([s ** dim for s in self.sets])


class MixedDataSet:
    def __init__(self, dsets):
        self.dsets = tuple(dsets)
        self.set = MixedSet(d.set for d in self.dsets)

    def __iter__(self):
        return iter(self.dsets)

    def __getitem__(self, i):
        return self.dsets[i]

    def __len__(self):
        return len(self.dsets)


class DataSet:
    def __init__(self, set_, dim, name=None):
        self.set = set_
        self.dim = tuple(dim) if isinstance(dim, (tuple, list)) else (int(dim),)
        self.cdim = int(np.prod(self.dim))


class ExtrudedSet(Set):
    """Columns of cells over ``parent`` (constant ``layers`` or a
    (parent.total_size, 2) array of [bottom, top) layers)."""

    def __init__(self, parent, layers, extruded_periodic=False):
        super().__init__((parent.core_size, parent.size, parent.total_size))
        self.parent = parent
        self.periodic = bool(extruded_periodic)
        arr = np.asarray(layers, dtype=np.int32)
        self.layers_array = arr if arr.ndim == 2 else np.array([[0, int(layers)]], dtype=np.int32)

    def cells(self, n):
        a, b = self.layers_array[0 if len(self.layers_array) == 1 else n]
        return range(b - a - 1)

    def layer_range(self, n, region=ALL):
        """Layers column ``n`` runs in ``region`` (builder.py:771-795,
        sparsity.pyx:345-356); a column with no cells runs no top cell."""
        bottom, top = (int(v) for v in self.layers_array[0 if len(self.layers_array) == 1 else n])
        cend = top - 1
        if region == ON_BOTTOM:
            return range(bottom, bottom + 1)
        if region == ON_TOP:
            return range(max(cend - 1, bottom), cend)
        if region == ON_INTERIOR_FACETS:
            return range(bottom, cend if self.periodic else cend - 1)
        return range(bottom, cend)

    def bottom(self, n):
        return int(self.layers_array[0 if len(self.layers_array) == 1 else n][0])


class Map:
    """``table`` is what a loop gathers through: entry ``k`` of element ``n``
    names the target ``table[n, k]``; an extruded map's entry k of layer l
    above the column's bottom is ``table[n, k] + offset[k] * l``."""

    def __init__(self, iterset, toset, arity, values=None, name=None, offset=None, offset_quotient=None):
        self.iterset, self.toset, self.arity = iterset, toset, arity
        self.offset_quotient = None if offset_quotient is None else np.asarray(offset_quotient, dtype=np.int32)
        if values is None:
            values = np.zeros((iterset.total_size, arity), dtype=np.int32)
        self.values = np.ascontiguousarray(np.asarray(values, dtype=np.int32).reshape(iterset.total_size, arity))
        self.offset = None if offset is None else np.broadcast_to(np.asarray(offset, dtype=np.int32),
                                                                  (arity,)).copy()

    def cell_rows(self, region=ALL):
        """One row of entries per iteration of an extruded loop over
        ``region`` (owned columns); interior facets list the cell below,
        then the cell above (sparsity.pyx:357-377, with the layer counted
        from the column's bottom)."""
        nf = 2 if region == ON_INTERIOR_FACETS else 1
        rows = []
        for n in range(self.iterset.size):
            b = self.iterset.bottom(n)
            ncell = len(self.iterset.cells(n))
            for layer in self.iterset.layer_range(n, region):
                row = []
                for f in range(nf):
                    for k, v in enumerate(self.table[n]):
                        step = layer - b + f
                        if self.iterset.periodic and ncell > 0:
                            # sparsity.pyx:369-377 (layers from the bottom)
                            q = 0 if self.offset_quotient is None else int(self.offset_quotient[k])
                            step = (step + q) % ncell - q % ncell
                        row.append(int(v) + int(self.offset[k]) * step if v >= 0 else -1)
                rows.append(row)
        return np.array(rows, dtype=np.int32).reshape(-1, nf * self.arity)

    @property
    def table(self):
        return self.values


class PermutedMap(Map):
    """Entry ``k`` of the permuted map is entry ``permutation[k]`` of the
    base map (``pyop2/codegen/builder.py:145-175``, ``PMap._permute``)."""

    def __init__(self, map_, permutation):
        self.map_, self.permutation = map_, np.asarray(permutation, dtype=np.int32)
        self.iterset, self.toset, self.arity = map_.iterset, map_.toset, map_.arity
        self.values = map_.values
        # PMap (builder.py:148-172): offsets permuted, quotients not
        self.offset = None if map_.offset is None else map_.offset[self.permutation]
        self.offset_quotient = map_.offset_quotient
        self._table = None

    @property
    def table(self):
        if self._table is None:
            base = self.map_.table
            self._table = np.array([[base[n][self.permutation[k]] for k in range(self.arity)]
                                    for n in range(len(base))], dtype=np.int32).reshape(-1, self.arity)
        return self._table


class ComposedMap(Map):
    """Element ``n`` is carried through ``maps_[-1]``, ..., ``maps_[1]`` (entry
    0 each) and then indexes ``maps_[0]`` (``pyop2/codegen/builder.py:178-194``,
    ``CMap.indexed``; the checks follow ``pyop2/types/map.py:216-260``)."""

    def __init__(self, *maps_, name=None):
        for tomap, frommap in zip(maps_[:-1], maps_[1:]):
            if tomap.iterset is not frommap.toset:
                raise MapTypeError("tomap.iterset must match frommap.toset")
            if frommap.arity != 1:
                raise MapTypeError("frommap.arity must be 1")
        self.maps_ = maps_
        self.iterset, self.toset, self.arity = maps_[-1].iterset, maps_[0].toset, maps_[0].arity
        # CMap (builder.py:175-193): the first map steps with the layers
        self.offset, self.offset_quotient = maps_[0].offset, maps_[0].offset_quotient
        self._table = None

    @property
    def values(self):
        raise RuntimeError("ComposedMap does not store values directly")

    @property
    def table(self):
        if self._table is not None:
            return self._table
        tables = [m.table for m in self.maps_]
        out = []
        for n in range(self.iterset.total_size):
            e = n
            for t in reversed(tables[1:]):
                e = int(t[e][0]) if e >= 0 else -1
            # a negative entry along the chain names no target (sparsity.pyx:329-340)
            out.append([int(v) for v in tables[0][e]] if e >= 0 else [-1] * self.arity)
        self._table = np.array(out, dtype=np.int32).reshape(-1, self.arity)
        return self._table


class MixedMap:
    def __init__(self, maps):
        self.maps = tuple(maps)
        self.iterset = self.maps[0].iterset

    def __iter__(self):
        return iter(self.maps)

    def __getitem__(self, i):
        return self.maps[i]

    def __len__(self):
        return len(self.maps)


class MixedDat:
    def __init__(self, arg):
        if isinstance(arg, MixedSet):
            arg = arg ** 1
        if isinstance(arg, MixedDataSet):
            self.dats = tuple(Dat(d) for d in arg)
        else:
            self.dats = tuple(arg)

    def __getitem__(self, i):
        return self.dats[i]

    def __iter__(self):
        return iter(self.dats)

    def __call__(self, access, path):
        return ("mdat", self, access, path)


class Dat:
    def __init__(self, dataset, data=None, dtype=None, name=None):
        if isinstance(dataset, Set):
            dataset = dataset ** 1
        self.dataset = dataset
        shape = (dataset.set.total_size,) + (() if dataset.cdim == 1 else dataset.dim)
        dtype = np.dtype(dtype) if dtype is not None else (np.asarray(data).dtype if data is not None else np.float64)
        self._a = np.zeros(shape, dtype=dtype) if data is None else np.ascontiguousarray(
            np.asarray(data, dtype=dtype).reshape(shape)).copy()

    @property
    def data(self):
        return self._a[:self.dataset.set.size]

    data_ro = data

    def zero(self):
        self._a[...] = 0

    def __call__(self, access, path=None):
        return ("dat", self, access, path)


class DatView:
    """Component ``index`` of ``dat``: a par_loop passes only that value."""

    def __init__(self, dat, index):
        self.dat = dat
        self.index = tuple(index) if isinstance(index, (tuple, list)) else (int(index),)
        self.comp = int(np.ravel_multi_index(self.index, dat.dataset.dim))

    @property
    def data(self):
        return self.dat._a.reshape(len(self.dat._a), -1)[:self.dat.dataset.set.size, self.comp]

    def __call__(self, access, path=None):
        return ("view", self, access, path)


class Global:
    def __init__(self, dim, data=None, dtype=None, name=None, comm=None):
        self.dim = (dim,) if isinstance(dim, int) else tuple(dim)
        self._a = np.zeros(self.dim, dtype=dtype or np.float64) if data is None else \
            np.ascontiguousarray(np.asarray(data, dtype=dtype).reshape(self.dim)).copy()

    @property
    def data(self):
        return self._a

    @data.setter
    def data(self, value):
        self._a[...] = value

    data_ro = data

    def zero(self):
        self._a[...] = 0

    def __call__(self, access):
        return ("glob", self, access, None)


class Sparsity:
    def __new__(cls, dsets, pairs, name=None, block_sparse=True):
        if isinstance(dsets[0], MixedDataSet):
            return MixedSparsity(dsets, pairs, block_sparse)
        return super().__new__(cls)

    def __init__(self, dsets, pairs, name=None, block_sparse=True):
        self.dsets = dsets
        rset, cset = dsets
        self.rdim, self.cdim = rset.cdim, cset.cdim
        self.pairs = [(p[0], p[1]) for p in pairs]
        regions = [tuple(p[2]) if len(p) > 2 and p[2] is not None else (ALL,) for p in pairs]
        diag = rset is cset or (rset.set is cset.set and rset.dim == cset.dim)
        tabs = []
        for (r, c), regs in zip(self.pairs, regions):
            if isinstance(r.iterset, ExtrudedSet):
                for region in regs:
                    rt, ct = r.cell_rows(region), c.cell_rows(region)
                    tabs.append((rt, ct, len(rt)))
            else:
                tabs.append((r.table, c.table, r.iterset.size))
        self.rowptr, self.colidx = osp.pattern(rset.set.total_size, cset.set.total_size, tabs, set_diag=diag)
        self.block = block_sparse and self.rdim == self.cdim and self.rdim > 1
        self.nnz = osp.dof_nnz(self.rowptr, self.rdim, self.cdim, block_sparse)


class MixedSparsity:
    """One block Sparsity per (row, column) component; blocks not given
    hold only their diagonal when on the diagonal."""

    def __init__(self, dsets, blocks, block_sparse=True):
        self.dsets = dsets
        rd, cd = dsets
        self.shape = (len(rd), len(cd))
        self._blocks = [[Sparsity((rd[i], cd[j]), blocks.get((i, j), []), block_sparse=block_sparse)
                         if (i, j) in blocks or i == j else None
                         for j in range(len(cd))] for i in range(len(rd))]

    def __getitem__(self, idx):
        i, j = idx
        return self._blocks[i][j]


class MixedMat:
    def __init__(self, sparsity, dtype=np.float64, name=None):
        self.sparsity = sparsity
        nr, nc = sparsity.shape
        self.blocks = [[Mat(sparsity[i, j], dtype) for j in range(nc)] for i in range(nr)]

    def __getitem__(self, idx):
        i, j = idx
        return self.blocks[i][j]

    def __call__(self, access, maps):
        return ("mmat", self, access, maps)

    def assemble(self):
        pass

    @property
    def values(self):
        return np.block([[b.values for b in row] for row in self.blocks])


class Mat:
    def __new__(cls, sparsity, dtype=np.float64, name=None):
        if isinstance(sparsity, MixedSparsity):
            return MixedMat(sparsity, dtype, name)
        return super().__new__(cls)

    def __init__(self, sparsity, dtype=np.float64, name=None):
        self.sparsity = sparsity
        s = sparsity
        if s.block:
            self.csr = seqwrap.CSRMat(s.rowptr, s.colidx, np.zeros((len(s.colidx), s.rdim, s.cdim)),
                                      s.rdim, s.cdim, block=True)
        else:
            rp, ci = osp.expand_block_pattern(s.rowptr, s.colidx, s.rdim, s.cdim)
            self.csr = seqwrap.CSRMat(rp, ci, np.zeros(len(ci)))

    def __call__(self, access, maps):
        return ("mat", self, access, maps)

    def zero(self):
        self.csr.values[...] = 0

    def assemble(self):
        pass

    @property
    def values(self):
        s = self.sparsity
        n, m = s.dsets[0].set.total_size, s.dsets[1].set.total_size
        if self.csr.block:
            return self.csr.to_dense(n, m)
        return self.csr.to_dense(n * s.rdim, m * s.cdim)


class Kernel:
    def __init__(self, code, name):
        self.code, self.name = code, name


class ParLoop:
    def __init__(self, kernel, iterset, *args):
        self.kernel, self.iterset, self.args = kernel, iterset, args

    def compute(self):
        par_loop(self.kernel, self.iterset, *self.args)


def par_loop(kernel, iterset, *args, pass_layer_arg=False, iteration_region=None):
    sargs = []
    temps = []
    for kind, obj, access, path in args:
        if kind == "dat":
            sargs.append(seqwrap.DatArg(obj._a, access, path.table if path is not None else None))
        elif kind == "view":
            sargs.append(seqwrap.DatArg(obj.dat._a, access, path.table if path is not None else None,
                                        comp=obj.comp))
        elif kind == "mdat":
            sargs.append(seqwrap.MixedDatArg([seqwrap.DatArg(d._a, access, m.table) for d, m in zip(obj, path)],
                                             access))
        elif kind == "mmat":
            rmaps, cmaps = path
            blocks = []
            for i, r in enumerate(rmaps):
                for j, c in enumerate(cmaps):
                    b = obj[i, j]
                    blocks.append(seqwrap.MatArg(b.csr, access, r.table, c.table, b.sparsity.rdim,
                                                 b.sparsity.cdim))
            sargs.append(seqwrap.MixedMatArg(blocks, (len(rmaps), len(cmaps)), access))
        elif kind == "glob":
            if access == INC:
                tmp = np.zeros_like(obj._a)
                temps.append((obj, tmp))
                sargs.append(seqwrap.GlobalArg(tmp, access))
            else:
                sargs.append(seqwrap.GlobalArg(obj._a, access))
        else:
            r, c = path
            s = obj.sparsity
            rd, cd = s.rdim, s.cdim
            sargs.append(seqwrap.MatArg(obj.csr, access, r.table, c.table, rd, cd))
    parts = [(0, iterset.core_size), (iterset.core_size, iterset.size)]
    subset = iterset.indices if isinstance(iterset, Subset) else None
    ext = {}
    columns = iterset.superset if isinstance(iterset, Subset) else iterset
    if isinstance(columns, ExtrudedSet):
        offs, quos = [], []
        for kind, obj, access, path in args:
            for m in (path if isinstance(path, tuple) else (path,)):
                if isinstance(m, Map) and m.offset is not None:
                    offs.append((m.table, m.offset))
                    if m.offset_quotient is not None:
                        quos.append((m.table, m.offset_quotient))
        ext = dict(layers=columns.layers_array, offsets=offs, pass_layer=pass_layer_arg,
                   region=iteration_region or ALL, periodic=columns.periodic, quotients=quos)
    err = seqwrap.run(kernel.code, kernel.name, sargs, parts, subset=subset, **ext)
    assert err == 0, "insertion outside the sparsity pattern"
    for g, tmp in temps:
        g._a += tmp
