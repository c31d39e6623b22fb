"""Run-time options of the device par_loop.

The option names and their environment variables are the reference's
interface (``pyop2/configuration.py:83-119``) for the options that mean
something on this path: ``type_check``, ``block_sparsity``, ``matnest``,
``compute_kernel_flops``, ``simd_width``, ``debug``, ``cache_dir``.  The
device adds:

* ``strategy`` -- how indirect INC conflicts are resolved: ``auto`` (the
  autotuner's choice per loop), ``atomic``, ``coloured``, ``staged``,
  ``split`` or one of the Mat row gathers (``rowgather*``);
* ``autotune`` -- whether ``auto`` times its candidates on first use;
* ``device`` -- the HIP device ordinal;
* ``staged_partials`` -- staged Dat INC loops send the nodes their tiles
  share through per-tile partial sums and a second pass (False: device
  atomic adds).

Every option is read once from its environment variable at import.  A
value assigned later must have the option's type (``ConfigurationError``
otherwise); ``unsafe_reconfigure`` skips that check.
"""
import os
from collections import namedtuple
from tempfile import gettempdir

from pyop2_amd.exceptions import ConfigurationError

_Option = namedtuple("_Option", "env type default")

_STRATEGIES = frozenset(("auto", "atomic", "coloured", "staged", "staged_coloured", "split", "rowgather",
                         "rowgather_thread", "rowgather_slots", "rowgather_staged",
                         "rowgather_cells", "rowgather_rows"))


def _from_env(opt):
    raw = os.environ.get(opt.env)
    if raw is None:
        return opt.default
    try:
        return bool(int(raw)) if opt.type is bool else opt.type(raw)
    except ValueError:
        raise ValueError("environment variable %s=%r is not a valid %s"
                         % (opt.env, raw, opt.type.__name__))


class Configuration(dict):
    """The options as a dict whose assignments are type-checked."""

    OPTIONS = {
        "simd_width": _Option("PYOP2_SIMD_WIDTH", int, 4),
        "debug": _Option("PYOP2_DEBUG", bool, False),
        "compute_kernel_flops": _Option("PYOP2_COMPUTE_KERNEL_FLOPS", bool, False),
        "type_check": _Option("PYOP2_TYPE_CHECK", bool, True),
        "cache_dir": _Option("PYOP2_CACHE_DIR", str,
                             os.path.join(gettempdir(), "pyop2_amd-cache-uid%d" % os.getuid())),
        "matnest": _Option("PYOP2_MATNEST", bool, True),
        "block_sparsity": _Option("PYOP2_BLOCK_SPARSITY", bool, True),
        "strategy": _Option("PYOP2_AMD_STRATEGY", str, "auto"),
        "autotune": _Option("PYOP2_AMD_AUTOTUNE", bool, True),
        "device": _Option("PYOP2_AMD_DEVICE", int, 0),
        "staged_partials": _Option("PYOP2_AMD_STAGED_PARTIALS", bool, True),
    }

    def __init__(self):
        self._initial = {name: _from_env(opt) for name, opt in self.OPTIONS.items()}
        super().__init__(self._initial)

    @staticmethod
    def _validated(key, value):
        opt = Configuration.OPTIONS.get(key)
        if opt is not None and not isinstance(value, opt.type):
            raise ConfigurationError("option %r takes %s values, not %r"
                                     % (key, opt.type.__name__, type(value).__name__))
        if key == "strategy" and value not in _STRATEGIES:
            raise ConfigurationError("unknown strategy %r (one of %s)" % (value, sorted(_STRATEGIES)))
        return value

    def __setitem__(self, key, value):
        super().__setitem__(key, self._validated(key, value))

    def reconfigure(self, **kwargs):
        """Assign several options, each checked."""
        for key, value in kwargs.items():
            self[key] = value

    def unsafe_reconfigure(self, **kwargs):
        """Assign several options without checks (restoring saved values)."""
        dict.update(self, kwargs)

    def reset(self):
        """Back to the values read at import."""
        dict.clear(self)
        dict.update(self, self._initial)


configuration = Configuration()
