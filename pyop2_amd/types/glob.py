"""Global and Constant (``pyop2/types/glob.py``).

A Global holds a small array on the host and the device.  INC/MIN/MAX
reductions happen on the device; on one rank the reference's allreduce is
the identity (``parloop.py:409-452``).
"""
import numbers

import numpy as np

from pyop2_amd import exceptions as ex, utils
from pyop2_amd.types.access import Access
from pyop2_amd.types.data_carrier import DataCarrier, DeviceStorage


class SetFreeDataCarrier(DataCarrier, DeviceStorage):

    @utils.validate_type(('name', str, ex.NameTypeError))
    def __init__(self, dim, data=None, dtype=None, name=None):
        self._dim = utils.as_tuple(dim, int)
        self._cdim = int(np.prod(self._dim))
        self._storage_init(data, dtype, self._dim)
        self._buf = np.empty(self._dim, dtype=self._dtype)
        self._name = name or "%s_#x%x" % (type(self).__name__.lower(), id(self))

    @property
    def _kernel_args_(self):
        from pyop2_amd import device as dev
        return (dev.ptr(self._device_array()),)

    def __iter__(self):
        yield self

    def __len__(self):
        return 1

    def __getitem__(self, idx):
        if idx != 0:
            raise ex.IndexValueError("Can only extract component 0 from %r" % self)
        return self

    @property
    def shape(self):
        return self._dim

    @property
    def data(self):
        self.increment_dat_version()
        if self._dtype.itemsize == 0:
            raise RuntimeError("Illegal access: No data associated with this Global!")
        a = self._host_array()
        self._host_modified()
        return a

    @data.setter
    def data(self, value):
        self.increment_dat_version()
        a = self._host_array()
        a[...] = utils.verify_reshape(value, self.dtype, self.dim)
        self._host_modified()

    @property
    def data_ro(self):
        view = self._host_array().view()
        view.setflags(write=False)
        return view

    @property
    def data_wo(self):
        return self.data

    @property
    def data_with_halos(self):
        return self.data

    @property
    def data_ro_with_halos(self):
        return self.data_ro

    @property
    def data_wo_with_halos(self):
        return self.data_wo

    @property
    def halo_valid(self):
        return True

    @halo_valid.setter
    def halo_valid(self, value):
        pass

    def copy(self, other, subset=None):
        other.data = np.copy(self.data_ro)

    @utils.cached_property
    def split(self):
        return (self,)

    @property
    def nbytes(self):
        return self.dtype.itemsize * self._cdim

    def global_to_local_begin(self, access_mode):
        pass

    def global_to_local_end(self, access_mode):
        pass

    def local_to_global_begin(self, insert_mode):
        pass

    def local_to_global_end(self, insert_mode):
        pass

    def frozen_halo(self, access_mode):
        import contextlib
        return contextlib.nullcontext()

    def freeze_halo(self, access_mode):
        pass

    def unfreeze_halo(self):
        pass

    _halo_frozen = False

    def _scalar(self, other):
        return other.data_ro if isinstance(other, SetFreeDataCarrier) else other

    def __add__(self, other):
        return type(self)(self.dim, data=self.data_ro + self._scalar(other), dtype=self.dtype, name=self.name)

    def __sub__(self, other):
        return type(self)(self.dim, data=self.data_ro - self._scalar(other), dtype=self.dtype, name=self.name)

    def __mul__(self, other):
        return type(self)(self.dim, data=self.data_ro * self._scalar(other), dtype=self.dtype, name=self.name)

    def __truediv__(self, other):
        return type(self)(self.dim, data=self.data_ro / self._scalar(other), dtype=self.dtype, name=self.name)

    def __radd__(self, other):
        return self + other

    def __rsub__(self, other):
        """``other - self`` (``glob.py:164-170``)."""
        ret = -self
        ret += other
        return ret

    def __rmul__(self, other):
        return self * other

    def __pos__(self):
        """A copy (``glob.py:147-148``)."""
        return self.duplicate()

    def __iadd__(self, other):
        self.data = self.data_ro + self._scalar(other)
        return self

    def __isub__(self, other):
        self.data = self.data_ro - self._scalar(other)
        return self

    def __imul__(self, other):
        self.data = self.data_ro * self._scalar(other)
        return self

    def __itruediv__(self, other):
        self.data = self.data_ro / self._scalar(other)
        return self

    def inner(self, other):
        assert isinstance(other, SetFreeDataCarrier)
        return np.dot(self.data_ro, np.conj(other.data_ro))


class Global(SetFreeDataCarrier):
    """OP2 global value; ``G(op2.READ)`` etc. makes a par_loop argument."""

    _modes = [Access.READ, Access.INC, Access.MIN, Access.MAX]

    def __init__(self, dim, data=None, dtype=None, name=None, comm=None):
        if isinstance(dim, (Global, Constant)):
            other = dim
            self.__init__(other._dim, None, dtype=other.dtype, name="copy_of_%s" % other.name,
                          comm=comm or other.comm)
            other.copy(self)
            return
        super().__init__(dim, data, dtype, name)
        self.comm = comm

    def __str__(self):
        return "OP2 Global Argument: %s with dim %s and value %s" % (self._name, self._dim, self.data_ro)

    def __repr__(self):
        return "Global(%r, %r, %r, %r)" % (self._dim, self.data_ro, self.dtype, self._name)

    @utils.validate_in(('access', _modes, ex.ModeValueError))
    def __call__(self, access, map_=None):
        from pyop2_amd.parloop import GlobalLegacyArg
        assert map_ is None
        return GlobalLegacyArg(self, access)

    def __neg__(self):
        return type(self)(self.dim, data=-np.copy(self.data_ro), dtype=self.dtype, name=self.name, comm=self.comm)

    def duplicate(self):
        return type(self)(self.dim, data=np.copy(self.data_ro), dtype=self.dtype, name=self.name, comm=self.comm)

    def zero(self, subset=None):
        assert subset is None
        self.increment_dat_version()
        if self._where in (2, 3):
            self._device_array().zero_()
            self._device_modified()
        else:
            self._host_array()[...] = 0
            self._host_modified()


class Constant(SetFreeDataCarrier):
    """A Global without a communicator that is only ever READ."""

    _modes = [Access.READ]

    def __init__(self, dim, data=None, dtype=None, name=None, comm=None):
        if isinstance(dim, (Global, Constant)):
            other = dim
            self.__init__(other._dim, None, dtype=other.dtype, name="copy_of_%s" % other.name)
            other.copy(self)
            return
        super().__init__(dim, data, dtype, name)
        if comm is not None:
            raise ValueError("Constants should not have communicators")
        self.comm = None

    def __str__(self):
        return "OP2 Constant Argument: %s with dim %s and value %s" % (self._name, self._dim, self.data_ro)

    def __repr__(self):
        return "Constant(%r, %r, %r, %r)" % (self._dim, self.data_ro, self.dtype, self._name)

    @utils.validate_in(('access', _modes, ex.ModeValueError))
    def __call__(self, access, map_=None):
        from pyop2_amd.parloop import GlobalLegacyArg
        assert map_ is None
        return GlobalLegacyArg(self, access)

    def duplicate(self):
        return type(self)(self.dim, data=np.copy(self.data_ro), dtype=self.dtype, name=self.name)
