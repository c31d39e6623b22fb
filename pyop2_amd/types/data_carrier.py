"""Host/device residency of Dat and Global storage.

The reference stores every Dat as a numpy array (``pyop2/types/
data_carrier.py:51-77``, lazily zero-allocated).  Here the authoritative copy
of an array may be the host numpy array, the device tensor (HBM), or both:

* ``_host_array()`` returns the full host array, downloading it when the
  device copy is newer (``data_ro`` and friends);
* ``_device_array()`` returns the device tensor, uploading the host array
  when it is newer (every par_loop argument);
* writers mark their side as the only valid copy.

Arrays that were never touched are allocated as zeros on whichever side
asks first, as the reference's lazy ``_data`` does.
"""
import numpy as np

from pyop2_amd import device as dev

_NONE, _HOST, _DEVICE, _BOTH = 0, 1, 2, 3


class DataCarrier:
    """Base of Dat, Global and Mat (``pyop2/types/data_carrier.py:13-48``)."""

    @property
    def dtype(self):
        return self._dtype

    @property
    def ctype(self):
        from pyop2_amd.datatypes import as_cstr
        return as_cstr(self.dtype)

    @property
    def name(self):
        return self._name

    @property
    def dim(self):
        return self._dim

    @property
    def cdim(self):
        return self._cdim

    def increment_dat_version(self):
        """Count a modification (``data_carrier.py:42-48``): carriers with
        storage keep a version; a Mat has none."""
        if isinstance(self, DeviceStorage):
            DeviceStorage.increment_dat_version(self)


class DeviceStorage:
    """Mixin holding ``_shape``/``_dtype`` data on the host and the device."""

    def _storage_init(self, data, dtype, shape):
        from pyop2_amd.utils import verify_reshape
        self._shape = tuple(shape)
        self._host = None
        self._dev = None
        self._where = _NONE
        self._version = 0
        if data is None:
            self._dtype = np.dtype(dtype if dtype is not None else np.float64)
        else:
            self._host = verify_reshape(data, dtype, self._shape, allow_none=True)
            self._dtype = self._host.dtype
            self._where = _HOST

    # -- versioning (the reference's dat_version counter) ------------------
    @property
    def dat_version(self):
        return self._version

    def increment_dat_version(self):
        self._version += 1

    # -- residency ------------------------------------------------------------
    @property
    def _is_allocated(self):
        return self._where != _NONE

    def _host_array(self):
        if self._where == _NONE:
            self._host = np.zeros(self._shape, dtype=self._dtype)
            self._where = _HOST
        elif self._where == _DEVICE:
            if self._host is None or self._host.shape != self._shape:
                self._host = np.empty(self._shape, dtype=self._dtype)
            dev.to_host(self._dev, out=self._host)
            self._where = _BOTH
        return self._host

    def _device_array(self):
        if self._where == _NONE:
            self._dev = dev.zeros(self._shape, self._dtype)
            self._where = _DEVICE
        elif self._where == _HOST:
            if self._dev is None:
                self._dev = dev.to_device(self._host)
            else:
                self._dev.copy_(dev.to_device(self._host))
            self._where = _BOTH
        return self._dev

    def _host_modified(self):
        """The host array is now the only valid copy."""
        self._host_array()
        self._where = _HOST

    def _device_modified(self):
        """The device tensor is now the only valid copy."""
        self._device_array()
        self._where = _DEVICE

    @property
    def _data(self):
        """Full host array (the reference's ``_data``), treated as modified."""
        a = self._host_array()
        self._where = _HOST
        return a

    # -- device accessors -------------------------------------------------------
    @property
    def device_data(self):
        """Device tensor of all values (incl. halo); marked as modified."""
        t = self._device_array()
        self.increment_dat_version()
        self._where = _DEVICE
        return t

    @property
    def device_data_ro(self):
        """Device tensor of all values (incl. halo); not to be modified."""
        return self._device_array()
