"""DataSet: a Set with a per-element shape (``pyop2/types/dataset.py:16-108``).

Identity-cached on its Set by ``(set, dim)`` like the reference.  The PETSc
layout objects of the reference (lgmaps, layout vectors) have no role on a
single device rank and are not provided.
"""
import numbers

import numpy as np

from pyop2_amd import caching, exceptions as ex, utils
from pyop2_amd.types.set import Set, Subset


class DataSet(caching.ObjectCached):

    @utils.validate_type(('iter_set', Set, ex.SetTypeError),
                         ('dim', (numbers.Integral, tuple, list), ex.DimTypeError),
                         ('name', str, ex.NameTypeError))
    def __init__(self, iter_set, dim=1, name=None):
        if self._initialized:
            return
        if isinstance(iter_set, Subset):
            raise NotImplementedError("Deriving a DataSet from a Subset is unsupported")
        self.comm = iter_set.comm
        self._set = iter_set
        self._dim = utils.as_tuple(dim, numbers.Integral)
        self._cdim = int(np.prod(self._dim))
        self._name = name or "dset_#x%x" % id(self)
        self._initialized = True

    @classmethod
    def _process_args(cls, *args, **kwargs):
        return (args[0],) + args, kwargs

    @classmethod
    def _cache_key(cls, iter_set, dim=1, name=None):
        try:
            return (iter_set, utils.as_tuple(dim, numbers.Integral))
        except TypeError:
            raise ex.DimTypeError("DataSet dim must be an integer or a tuple of integers, not %r" % (dim,))

    def __getattr__(self, name):
        if name.startswith("__") or name == "_set":
            raise AttributeError(name)
        return getattr(self._set, name)

    def __getitem__(self, idx):
        assert idx == 0
        return self

    @property
    def dim(self):
        return self._dim

    @property
    def cdim(self):
        return self._cdim

    @property
    def name(self):
        return self._name

    @property
    def set(self):
        return self._set

    def __iter__(self):
        yield self

    def __len__(self):
        return 1

    def __str__(self):
        return "OP2 DataSet: %s on set %s, with dim %s" % (self._name, self._set, self._dim)

    def __repr__(self):
        return "DataSet(%r, %r, %r)" % (self._set, self._dim, self._name)

    def __contains__(self, dat):
        return dat.dataset == self
