"""Dat: values on every element of a DataSet (``pyop2/types/dat.py``).

Storage is ``(total_size,)`` when ``cdim == 1`` else ``(total_size, *dim)``,
C-contiguous (``dat.py:76``), held on the device between par_loops (see
:mod:`pyop2_amd.types.data_carrier`).  ``data``/``data_ro``/``data_wo`` and
the ``*_with_halos`` variants keep the reference's semantics and return
numpy views of the host copy; ``device_data`` exposes the torch tensor.
"""
import contextlib

import numpy as np

from pyop2_amd import configuration as conf, exceptions as ex, utils
from pyop2_amd.types.access import Access
from pyop2_amd.types.data_carrier import DataCarrier, DeviceStorage
from pyop2_amd.types.dataset import DataSet
from pyop2_amd.types.set import Set


_READING = frozenset((Access.READ, Access.RW))
_REDUCING = frozenset((Access.INC, Access.MIN, Access.MAX))


class AbstractDat(DataCarrier, DeviceStorage):

    _modes = [Access.READ, Access.WRITE, Access.RW, Access.INC, Access.MIN, Access.MAX]

    @utils.validate_type(('dataset', (DataCarrier, DataSet, Set), ex.DataSetTypeError),
                         ('name', str, ex.NameTypeError))
    @utils.validate_dtype(('dtype', None, ex.DataTypeError))
    def __init__(self, dataset, data=None, dtype=None, name=None):
        if isinstance(dataset, Dat):
            self.__init__(dataset.dataset, None, dtype=dataset.dtype, name="copy_of_%s" % dataset.name)
            dataset.copy(self)
            return
        if type(dataset) is Set:
            dataset = dataset ** 1
        shape = (dataset.total_size,) + (() if dataset.cdim == 1 else dataset.dim)
        self._storage_init(data, dtype, shape)
        self._dataset = dataset
        self.comm = dataset.comm
        self.halo_valid = True
        self._name = name or "dat_#x%x" % id(self)
        self._halo_frozen = False
        self._frozen_access_mode = None

    @property
    def _kernel_args_(self):
        from pyop2_amd import device as dev
        return (dev.ptr(self._device_array()),)

    @utils.validate_in(('access', _modes, ex.ModeValueError))
    def __call__(self, access, path=None):
        from pyop2_amd.parloop import DatLegacyArg
        if conf.configuration["type_check"] and path and path.toset != self.dataset.set:
            raise ex.MapValueError("To Set of Map does not match Set of Dat.")
        return DatLegacyArg(self, path, access)

    def __getitem__(self, idx):
        if idx != 0:
            raise ex.IndexValueError("Can only extract component 0 from %r" % self)
        return self

    @utils.cached_property
    def split(self):
        return (self,)

    @property
    def dataset(self):
        return self._dataset

    @property
    def dim(self):
        return self.dataset.dim

    @property
    def cdim(self):
        return self.dataset.cdim

    @property
    def shape(self):
        return self._shape

    @property
    def data(self):
        """Writable numpy view of the owned values (marks the Dat modified)."""
        self.increment_dat_version()
        if self.dataset.total_size > 0 and self._dtype.itemsize == 0:
            raise RuntimeError("Illegal access: no data associated with this Dat!")
        self.halo_valid = False
        a = self._host_array()
        self._host_modified()
        v = a[:self.dataset.size].view()
        v.setflags(write=True)
        return v

    @property
    def data_with_halos(self):
        self.increment_dat_version()
        self.global_to_local_begin(Access.RW)
        self.global_to_local_end(Access.RW)
        self.halo_valid = False
        a = self._host_array()
        self._host_modified()
        v = a.view()
        v.setflags(write=True)
        return v

    @property
    def data_ro(self):
        """Read-only numpy view of the owned values."""
        v = self._host_array()[:self.dataset.size].view()
        v.setflags(write=False)
        return v

    @property
    def data_ro_with_halos(self):
        self.global_to_local_begin(Access.READ)
        self.global_to_local_end(Access.READ)
        v = self._host_array().view()
        v.setflags(write=False)
        return v

    @property
    def data_wo(self):
        return self.data

    @property
    def data_wo_with_halos(self):
        self.increment_dat_version()
        self.halo_valid = False
        a = self._host_array()
        self._host_modified()
        v = a.view()
        v.setflags(write=True)
        return v

    def save(self, filename):
        """The owned values as a ``.npy`` file (numpy's format, no pickles)."""
        np.save(filename, np.asarray(self.data_ro), allow_pickle=False)

    def load(self, filename):
        """Owned values from a ``.npy`` file written by ``save`` (the suffix
        may be left out)."""
        path = filename if filename.endswith(".npy") else filename + ".npy"
        self.data[:] = np.load(path, allow_pickle=False)

    @property
    def nbytes(self):
        return self.dtype.itemsize * self.dataset.total_size * self.dataset.cdim

    def _owned_rows_of(self, subset):
        """The owned rows a Subset of this Dat's set selects."""
        if subset.superset != self.dataset.set:
            raise ex.MapValueError("The subset and dataset are incompatible")
        return subset.owned_indices

    def _device_current(self):
        return self._where in (2, 3)

    def zero(self, subset=None):
        """Zero every value, halo included (the halo is then valid), on the
        side that holds the current copy; or the owned rows of a Subset."""
        self.increment_dat_version()
        if subset is not None:
            self.data[self._owned_rows_of(subset)] = 0
            return
        if self._device_current():
            self._device_array().zero_()
            self._device_modified()
        else:
            self._host_array().fill(0)
            self._host_modified()
        self.halo_valid = True

    def copy(self, other, subset=None):
        """Copies the values into the Dat ``other``: all of them (halo
        included when this Dat's halo is valid), or the owned rows of a
        Subset."""
        if other is self:
            return
        if subset is not None:
            rows = self._owned_rows_of(subset)
            other.data[rows] = self.data_ro[rows]
            return
        if not self.halo_valid:
            other.data[:] = self.data_ro
            return
        if self._device_current():
            other._device_array().copy_(self._device_array())
            other._device_modified()
        else:
            np.copyto(other._host_array(), self._host_array())
            other._host_modified()
        other.halo_valid = True

    def __iter__(self):
        yield self

    def __len__(self):
        return 1

    def __str__(self):
        return "OP2 Dat: %s on (%s) with datatype %s" % (self._name, self._dataset, self.dtype.name)

    def __repr__(self):
        return "Dat(%r, None, %r, %r)" % (self._dataset, self.dtype, self._name)

    # -- pointwise algebra on the device (torch) ------------------------------
    def _check_shape(self, other):
        if other.dataset.dim != self.dataset.dim:
            raise ValueError('Mismatched shapes in operands %s and %s' % (self.dataset.dim, other.dataset.dim))

    def _operand(self, other):
        from pyop2_amd.types.glob import Global
        if isinstance(other, AbstractDat):
            self._check_shape(other)
            return other._device_array()[:self.dataset.size]
        if isinstance(other, Global):
            return other._device_array()
        return other

    def _iop(self, other, op):
        # the result keeps this Dat's type, as the reference's generated
        # ``self[i] = self[i] op other[i]`` does (dat.py:466-491): computed in
        # the promoted type, then converted (integers truncate toward zero)
        a = self._device_array()[:self.dataset.size]
        a.copy_(getattr(a, op.rstrip("_"))(self._operand(other)))
        self._device_modified()
        self.increment_dat_version()
        return self

    def _op(self, other, op):
        ret = type(self)(self.dataset, None, self.dtype)
        ret._device_array()[:self.dataset.size] = getattr(self._device_array()[:self.dataset.size], op)(self._operand(other))
        ret._device_modified()
        return ret

    def __iadd__(self, other):
        return self._iop(other, "add_")

    def __isub__(self, other):
        return self._iop(other, "sub_")

    def __imul__(self, other):
        return self._iop(other, "mul_")

    def __itruediv__(self, other):
        return self._iop(other, "div_")

    def __add__(self, other):
        return self._op(other, "add")

    def __radd__(self, other):
        return self + other

    def __sub__(self, other):
        return self._op(other, "sub")

    def __mul__(self, other):
        return self._op(other, "mul")

    def __rmul__(self, other):
        return self * other

    def __truediv__(self, other):
        return self._op(other, "div")

    def __neg__(self):
        return self._op(-1, "mul")

    def __pos__(self):
        """A copy (``dat.py:495-497``)."""
        return Dat(self)

    def __rsub__(self, other):
        """``other - self`` (``dat.py:538-544``)."""
        ret = -self
        ret += other
        return ret

    def inner(self, other):
        """Inner product of the owned values (a float)."""
        self._check_shape(other)
        a = self._device_array()[:self.dataset.size]
        b = other._device_array()[:self.dataset.size]
        return float((a * b).sum().item())

    @property
    def norm(self):
        """Euclidean norm of the owned values."""
        return self.inner(self) ** 0.5

    # -- halo exchanges -------------------------------------------------------------
    # The DataSet's halo object (the reference's Halo interface, with the
    # forward and reverse exchange halves) moves the values; on one rank
    # there is none and every hook returns at once.  Before a read of stale
    # halo values the forward exchange refreshes them.  Before an INC, MIN or
    # MAX loop the halo entries are reset to the operation's identity, so
    # that after the loop they hold only what it contributed, which the
    # reverse exchange sends to the owners.
    def _live_halo(self):
        halo = self.dataset.halo
        return None if (halo is None or self._halo_frozen) else halo

    def _has_halo(self):
        """Whether the Dat's set has a halo object (exchanges can do work)."""
        return self.dataset.halo is not None

    def _reset_halo_entries(self, access_mode):
        from pyop2_amd.datatypes import dtype_limits
        lowest, highest = dtype_limits(self.dtype)
        identity = {Access.INC: 0, Access.MIN: highest, Access.MAX: lowest}[access_mode]
        self._device_array()[self.dataset.size:] = identity
        self._device_modified()

    def global_to_local_begin(self, access_mode):
        halo = self._live_halo()
        if halo is None:
            return
        if access_mode in _REDUCING:
            self._reset_halo_entries(access_mode)
        elif access_mode in _READING and not self.halo_valid:
            halo.global_to_local_begin(self, Access.WRITE)

    def global_to_local_end(self, access_mode):
        halo = self._live_halo()
        if halo is None:
            return
        if access_mode in _REDUCING:
            self.halo_valid = False
        elif access_mode in _READING and not self.halo_valid:
            halo.global_to_local_end(self, Access.WRITE)
            self.halo_valid = True

    def local_to_global_begin(self, insert_mode):
        halo = self._live_halo()
        if halo is not None:
            halo.local_to_global_begin(self, insert_mode)

    def local_to_global_end(self, insert_mode):
        halo = self._live_halo()
        if halo is not None:
            halo.local_to_global_end(self, insert_mode)
            self.halo_valid = False

    def frozen_halo(self, access_mode):
        return frozen_halo(self, access_mode)

    def freeze_halo(self, access_mode):
        if self._halo_frozen:
            raise RuntimeError("Expected an unfrozen halo")
        self._halo_frozen = True
        self._frozen_access_mode = access_mode

    def unfreeze_halo(self):
        if not self._halo_frozen:
            raise RuntimeError("Expected a frozen halo")
        self._halo_frozen = False
        self._frozen_access_mode = None


class Dat(AbstractDat):
    """OP2 vector data (see :class:`AbstractDat`)."""

    @contextlib.contextmanager
    def vec_context(self, access):
        raise NotImplementedError("PETSc Vec views are not part of the device path")


class DatView(AbstractDat):
    """One component of a Dat (``pyop2/types/dat.py:668-760``).  A par_loop
    hands the kernel that component only (one value per map entry, or a
    pointer to the element's component for a direct argument); the data
    accessors are views of the parent's arrays, and versions and halo
    validity are the parent's."""

    def __init__(self, dat, index):
        index = utils.as_tuple(index, int)
        if len(index) != len(dat.dim):
            raise ex.IndexValueError("DatView index %s has the wrong rank for shape %s" % (index, dat.dim))
        for i, d in zip(index, dat.dim):
            if not (0 <= i < d):
                raise ex.IndexValueError("Can't create DatView with index %s for Dat with shape %s"
                                         % (index, dat.dim))
        self.index = index
        self._idx = (slice(None), *index) if dat.cdim > 1 else (slice(None),)
        self._parent = dat
        self._dataset = dat.dataset
        self.comm = dat.comm
        self._name = "view[%s](%s)" % (index, dat.name)
        self._halo_frozen = False
        self._frozen_access_mode = None

    @property
    def dtype(self):
        return self._parent.dtype

    @property
    def cdim(self):
        return 1

    @property
    def dim(self):
        return (1,)

    @property
    def shape(self):
        return (self.dataset.total_size,)

    @property
    def _kernel_args_(self):
        return self._parent._kernel_args_

    def _device_array(self):
        return self._parent._device_array()

    def _device_modified(self):
        self._parent._device_modified()

    @property
    def halo_valid(self):
        return self._parent.halo_valid

    @halo_valid.setter
    def halo_valid(self, value):
        self._parent.halo_valid = value

    @property
    def dat_version(self):
        return self._parent.dat_version

    def increment_dat_version(self):
        self._parent.increment_dat_version()

    @property
    def _data(self):
        return self._parent._data[self._idx]

    @property
    def data(self):
        return self._parent.data[self._idx]

    @property
    def data_ro(self):
        return self._parent.data_ro[self._idx]

    @property
    def data_wo(self):
        return self._parent.data_wo[self._idx]

    @property
    def data_with_halos(self):
        return self._parent.data_with_halos[self._idx]

    @property
    def data_ro_with_halos(self):
        return self._parent.data_ro_with_halos[self._idx]

    @property
    def data_wo_with_halos(self):
        return self._parent.data_wo_with_halos[self._idx]

    def zero(self, subset=None):
        """Zero this component (``dat.py:576-590``): all of it, halo
        included, which makes the halo valid; or a Subset's owned entries."""
        self._parent.increment_dat_version()
        if subset is None:
            self._parent._data[self._idx] = 0
            self._parent.halo_valid = True
        elif subset.superset != self.dataset.set:
            raise ex.MapValueError("The subset and dataset are incompatible")
        else:
            self.data[subset.owned_indices] = 0

    def __str__(self):
        return "OP2 DatView: %s" % self._name

    def __repr__(self):
        return "DatView(%r, %r)" % (self._parent, self.index)


@contextlib.contextmanager
def frozen_halo(dat, access_mode):
    """Temporarily disable halo exchanges of ``dat``."""
    dat.freeze_halo(access_mode)
    try:
        yield
    finally:
        dat.unfreeze_halo()
