"""Exception types of the op2 API (same names and bases as
``pyop2/exceptions.py``, so ``except MapValueError`` code keeps working)."""


class DataTypeError(TypeError):
    """Invalid type for data."""


class DimTypeError(TypeError):
    """Invalid type for dimension."""


class ArityTypeError(TypeError):
    """Invalid type for arity."""


class IndexTypeError(TypeError):
    """Invalid type for index."""


class NameTypeError(TypeError):
    """Invalid type for name."""


class SetTypeError(TypeError):
    """Invalid type for Set."""


class SizeTypeError(TypeError):
    """Invalid type for size."""


class SubsetIndexOutOfBounds(TypeError):
    """Out of bound index."""


class SparsityTypeError(TypeError):
    """Invalid type for Sparsity."""


class MapTypeError(TypeError):
    """Invalid type for Map."""


class DataSetTypeError(TypeError):
    """Invalid type for DataSet."""


class MatTypeError(TypeError):
    """Invalid type for Mat."""


class DatTypeError(TypeError):
    """Invalid type for Dat."""


class KernelTypeError(TypeError):
    """Invalid type for Kernel."""


class DataValueError(ValueError):
    """Illegal value for data."""


class IndexValueError(ValueError):
    """Illegal value for index."""


class ModeValueError(ValueError):
    """Illegal value for mode."""


class IterateValueError(ValueError):
    """Illegal value for iterate."""


class SetValueError(ValueError):
    """Illegal value for Set."""


class MapValueError(ValueError):
    """Illegal value for Map."""


class ConfigurationError(RuntimeError):
    """Illegal configuration value or type."""


class CompilationError(RuntimeError):
    """Error during runtime compilation of a generated device kernel."""


class SparsityFormatError(ValueError):
    """Unable to produce a sparsity for this matrix format."""


class CachingError(ValueError):
    """A caching error."""


class HashError(CachingError):
    """Something is wrong with the hash."""


class DeviceError(RuntimeError):
    """The HIP device or the op2hip library is unavailable or failed."""


class PatternError(RuntimeError):
    """A Mat insertion fell outside the sparsity pattern
    (PETSc's NEW_NONZERO_LOCATION_ERR in the reference)."""
