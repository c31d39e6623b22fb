"""A host-only stand-in for the ``op2`` namespace that executes par_loops with
the oracle (``oracle.seqwrap``): same call forms as ``pyop2_amd.op2`` for the
subset the known-answer cases use, numpy storage, reference semantics
(core part then owned part, zeroed temporaries for INC Globals)."""
import numpy as np

from oracle import seqwrap, sparsity as osp

READ, WRITE, RW, INC, MIN, MAX = "READ", "WRITE", "RW", "INC", "MIN", "MAX"


class Set:
    def __init__(self, size, name=None):
        if isinstance(size, (int, np.integer)):
            size = (size, size, size)
        self.core_size, self.size, self.total_size = (int(s) for s in size)
        self.name = name

    def __pow__(self, dim):
        return DataSet(self, dim)


class DataSet:
    def __init__(self, set_, dim):
        self.set = set_
        self.dim = tuple(dim) if isinstance(dim, (tuple, list)) else (int(dim),)
        self.cdim = int(np.prod(self.dim))


class Map:
    def __init__(self, iterset, toset, arity, values=None, name=None):
        self.iterset, self.toset, self.arity = iterset, toset, arity
        self.values = np.ascontiguousarray(np.asarray(values, dtype=np.int32).reshape(iterset.total_size, arity))


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

    def __call__(self, access, path=None):
        return ("dat", self, access, path)


class Global:
    def __init__(self, dim, data=None, dtype=None, name=None, comm=None):
        self.dim = (dim,) if isinstance(dim, int) else tuple(dim)
        self._a = np.zeros(self.dim, dtype=dtype or np.float64) if data is None else \
            np.ascontiguousarray(np.asarray(data, dtype=dtype).reshape(self.dim)).copy()

    @property
    def data(self):
        return self._a

    data_ro = data

    def __call__(self, access):
        return ("glob", self, access, None)


class Sparsity:
    def __init__(self, dsets, pairs, name=None, block_sparse=True):
        self.dsets = dsets
        rset, cset = dsets
        self.rdim, self.cdim = rset.cdim, cset.cdim
        self.pairs = [(r, c) for r, c, _ in pairs]
        diag = rset is cset or (rset.set is cset.set and rset.dim == cset.dim)
        self.rowptr, self.colidx = osp.pattern(rset.set.total_size, cset.set.total_size,
                                               [(r.values, c.values, r.iterset.size) for r, c in self.pairs],
                                               set_diag=diag)
        self.block = block_sparse and self.rdim == self.cdim and self.rdim > 1
        self.nnz = osp.dof_nnz(self.rowptr, self.rdim, self.cdim, block_sparse)


class Mat:
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


def par_loop(kernel, iterset, *args):
    sargs = []
    temps = []
    for kind, obj, access, path in args:
        if kind == "dat":
            sargs.append(seqwrap.DatArg(obj._a, access, path.values if path is not None else None))
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
            sargs.append(seqwrap.MatArg(obj.csr, access, r.values, c.values, rd, cd))
    parts = [(0, iterset.core_size), (iterset.core_size, iterset.size)]
    err = seqwrap.run(kernel.code, kernel.name, sargs, parts)
    assert err == 0, "insertion outside the sparsity pattern"
    for g, tmp in temps:
        g._a += tmp
