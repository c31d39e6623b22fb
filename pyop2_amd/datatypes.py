"""Index and scalar types (``pyop2/datatypes.py:6-10``: PETSc int = int32,
scalar = float64)."""
import numpy as np

IntType = np.dtype(np.int32)
RealType = np.dtype(np.float64)
ScalarType = np.dtype(np.float64)

_CTYPES = {
    np.dtype(np.float64): "double",
    np.dtype(np.float32): "float",
    np.dtype(np.int32): "int32_t",
    np.dtype(np.uint32): "uint32_t",
    np.dtype(np.int64): "int64_t",
    np.dtype(np.uint64): "uint64_t",
    np.dtype(np.int16): "int16_t",
    np.dtype(np.uint16): "uint16_t",
    np.dtype(np.int8): "int8_t",
    np.dtype(np.uint8): "uint8_t",
}


def as_cstr(dtype):
    """C type name of a numpy dtype."""
    try:
        return _CTYPES[np.dtype(dtype)]
    except KeyError:
        raise TypeError(f"unsupported dtype {dtype}") from None


def dtype_limits(dtype):
    """(min, max) representable values of ``dtype``."""
    dtype = np.dtype(dtype)
    if dtype.kind == "f":
        info = np.finfo(dtype)
    else:
        info = np.iinfo(dtype)
    return info.min, info.max
