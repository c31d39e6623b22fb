"""Typed configuration dictionary (mirrors ``pyop2/configuration.py``).

Keys kept from the reference: ``type_check``, ``block_sparsity``, ``matnest``,
``compute_kernel_flops``, ``simd_width``, ``debug``, ``cache_dir``.
Device keys:

* ``strategy`` -- INC conflict resolution: ``auto`` (autotuned per loop),
  ``atomic``, ``coloured``, ``staged`` or ``rowgather`` (Mat only).
* ``autotune`` -- time the candidate strategies on first use (``auto`` only).
* ``device`` -- HIP device ordinal.
"""
import os
from tempfile import gettempdir

from pyop2_amd.exceptions import ConfigurationError


class Configuration(dict):
    cache_dir = os.path.join(gettempdir(), "pyop2_amd-cache-uid%s" % os.getuid())
    DEFAULTS = {
        "simd_width": ("PYOP2_SIMD_WIDTH", int, 4),
        "debug": ("PYOP2_DEBUG", bool, False),
        "compute_kernel_flops": ("PYOP2_COMPUTE_KERNEL_FLOPS", bool, False),
        "type_check": ("PYOP2_TYPE_CHECK", bool, True),
        "cache_dir": ("PYOP2_CACHE_DIR", str, cache_dir),
        "matnest": ("PYOP2_MATNEST", bool, True),
        "block_sparsity": ("PYOP2_BLOCK_SPARSITY", bool, True),
        "strategy": ("PYOP2_AMD_STRATEGY", str, "auto"),
        "autotune": ("PYOP2_AMD_AUTOTUNE", bool, True),
        "device": ("PYOP2_AMD_DEVICE", int, 0),
        # staged Dat INC loops: nodes shared between tiles through partial
        # sums and a second pass (False: device atomic adds)
        "staged_partials": ("PYOP2_AMD_STAGED_PARTIALS", bool, True),
    }

    def __init__(self):
        def convert(env, typ, v):
            if not isinstance(typ, type):
                typ = typ[0]
            try:
                if typ is bool:
                    return bool(int(os.environ.get(env, v)))
                return typ(os.environ.get(env, v))
            except ValueError:
                raise ValueError("Cannot convert value of environment variable %s to %r" % (env, typ))
        defaults = dict((k, convert(env, typ, v)) for k, (env, typ, v) in Configuration.DEFAULTS.items())
        super().__init__(**defaults)
        self._set = set()
        self._defaults = defaults

    def reset(self):
        """Reset the configuration parameters to the default values."""
        self.update(self._defaults)
        self._set = set()

    def reconfigure(self, **kwargs):
        """Update the configuration parameters with new values."""
        for k, v in kwargs.items():
            self[k] = v

    def unsafe_reconfigure(self, **kwargs):
        """Update the configuration without type checks."""
        self.update(kwargs)

    def __setitem__(self, key, value):
        if key in Configuration.DEFAULTS:
            valid_type = Configuration.DEFAULTS[key][1]
            if not isinstance(value, valid_type):
                raise ConfigurationError("Values for configuration key %s must be of type %r, not %r"
                                         % (key, valid_type, type(value)))
        if key == "strategy" and value not in ("auto", "atomic", "coloured", "staged", "rowgather",
                                               "rowgather_thread", "rowgather_slots",
                                               "rowgather_staged", "rowgather_cells",
                                               "rowgather_rows", "split"):
            raise ConfigurationError(f"unknown strategy {value!r}")
        self._set.add(key)
        super().__setitem__(key, value)


configuration = Configuration()
