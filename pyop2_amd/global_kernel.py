"""Global kernel: the drop-in boundary of the par_loop hot path.

The reference's ``GlobalKernel`` (``pyop2/global_kernel.py:250-429``) turns a
local kernel plus argument descriptors into a JIT-compiled C function
``wrap_<name>(start, end, ptrs...)`` that ``Parloop._compute`` calls once per
partition (``parloop.py:224-232``).  Here ``compile()`` returns a device
executor with the same ``(start, end, *ptrs)`` contract:

* a hand-written HIP entry point of ``libop2hip.so`` when the local kernel
  is a registered element kernel and the argument signature matches
  (``pyop2_amd.registry``);
* otherwise a gather/compute/scatter kernel generated around the user's C
  code and compiled for gfx950 at run time (``pyop2_amd.codegen``).

The descriptor classes keep the reference's fields and cache keys
(``global_kernel.py:23-247``).
"""
import collections
import ctypes
import itertools
from dataclasses import dataclass
from typing import Optional, Tuple

import numpy as np

from pyop2_amd.configuration import configuration
from pyop2_amd.types.access import READ
from pyop2_amd.types.region import IterationRegion
from pyop2_amd.utils import cached_property


@dataclass(eq=False, frozen=True)
class MapKernelArg:
    """A map argument: its arity and, on extruded loops, its per-layer
    offsets and (periodic) offset quotients."""
    arity: int
    offset: Optional[Tuple[int, ...]] = None
    offset_quotient: Optional[Tuple[int, ...]] = None

    @property
    def cache_key(self):
        return type(self), self.arity, self.offset, self.offset_quotient


@dataclass(frozen=True)
class GlobalKernelArg:
    dim: Tuple[int, ...]

    @property
    def cache_key(self):
        return type(self), self.dim

    @property
    def maps(self):
        return ()


@dataclass(frozen=True)
class DatKernelArg:
    dim: Tuple[int, ...]
    map_: MapKernelArg = None
    index: Optional[Tuple[int, ...]] = None

    @property
    def is_direct(self):
        return self.map_ is None

    @property
    def is_indirect(self):
        return not self.is_direct

    @property
    def cache_key(self):
        map_key = self.map_.cache_key if self.map_ is not None else None
        return type(self), self.dim, map_key, self.index

    @property
    def maps(self):
        return (self.map_,) if self.map_ is not None else ()


@dataclass(frozen=True)
class MatKernelArg:
    dims: Tuple[Tuple[int, ...], Tuple[int, ...]]
    maps: Tuple[MapKernelArg, MapKernelArg]
    unroll: bool = False

    @property
    def cache_key(self):
        return type(self), self.dims, tuple(m.cache_key for m in self.maps), self.unroll


@dataclass(frozen=True)
class MixedDatKernelArg:
    """A MixedDat argument: one DatKernelArg per component, packed one after
    the other into the kernel's buffer (``global_kernel.py:141-180``)."""
    arguments: Tuple[DatKernelArg, ...]

    @property
    def cache_key(self):
        return tuple(a.cache_key for a in self.arguments)

    @property
    def maps(self):
        return tuple(m for a in self.arguments for m in a.maps)

    @property
    def is_direct(self):
        return False

    @property
    def is_indirect(self):
        return True

    @property
    def split(self):
        return self.arguments


@dataclass(frozen=True)
class MixedMatKernelArg:
    """A MixedMat argument: the MatKernelArg of every block, row-major over
    ``shape`` = (row components, column components) (``global_kernel.py:
    209-247``).  ``maps`` lists the row maps, then the column maps."""
    arguments: Tuple[MatKernelArg, ...]
    shape: Tuple[int, int]

    @property
    def cache_key(self):
        return type(self), tuple(a.cache_key for a in self.arguments), self.shape

    @property
    def maps(self):
        nr, nc = self.shape
        rows = tuple(self.arguments[i * nc].maps[0] for i in range(nr))
        cols = tuple(self.arguments[j].maps[1] for j in range(nc))
        return rows + cols

    @property
    def split(self):
        return self.arguments


class PassthroughKernelArg:
    @property
    def cache_key(self):
        return type(self)

    @property
    def maps(self):
        return ()


class GlobalKernel:
    """Compiled form of a local kernel over a given argument signature."""

    def __init__(self, local_kernel, arguments, *, extruded=False, extruded_periodic=False,
                 constant_layers=False, subset=False, iteration_region=None, pass_layer_arg=False):
        if not len(local_kernel.accesses) == len(arguments):
            raise ValueError("Number of arguments passed to the local and global kernels do not match")
        from pyop2_amd.types import region as reg
        if iteration_region is not None:
            iteration_region = reg.normalise(iteration_region)
            if iteration_region == reg.ON_INTERIOR_FACETS and not extruded:
                raise ValueError("ON_INTERIOR_FACETS needs an extruded iteration set")
        if pass_layer_arg and not extruded:
            raise ValueError("Cannot request layer argument for non-extruded iteration")
        from pyop2_amd.types.glob import Constant
        if any(isinstance(garg, Constant) and larg.access is not READ
               for larg, garg in zip(local_kernel.arguments, arguments)):
            raise ValueError("Constants can only ever be read in a parloop, not modified")
        counter = itertools.count()
        seen_maps = collections.defaultdict(lambda: next(counter))
        self.cache_key = (
            local_kernel.cache_key,
            *[a.cache_key for a in arguments],
            *[seen_maps[m] for a in arguments for m in a.maps],
            extruded, extruded_periodic, constant_layers, subset,
            iteration_region, pass_layer_arg, configuration["simd_width"],
        )
        self.local_kernel = local_kernel
        self.arguments = arguments
        self._subset = subset
        self._extruded = bool(extruded)
        self._constant_layers = bool(constant_layers)
        self._pass_layer_arg = bool(pass_layer_arg)
        # the layers a column runs (codegen/builder.py:771-795); no effect on
        # a loop that is not extruded
        self._iteration_region = iteration_region if extruded else None
        self._extruded_periodic = bool(extruded_periodic) and bool(extruded)
        self._executor = None

    def __call__(self, comm, *args, parloop=None):
        """Execute over ``args = (start, end, *ptrs)`` on the current stream."""
        fn = self.compile(comm)
        start, end, ptrs = args[0], args[1], args[2:]
        fn(start, end, ptrs, parloop)

    @cached_property
    def name(self):
        return f"wrap_{self.local_kernel.name}"

    @cached_property
    def zipped_arguments(self):
        return tuple(zip(self.local_kernel.arguments, self.arguments))

    def compile(self, comm=None):
        """The executor: a hand-written entry point or a generated kernel."""
        if self._executor is None:
            from pyop2_amd import registry
            ex = registry.lookup(self)
            if ex is None:
                from pyop2_amd.codegen import GeneratedKernel
                ex = GeneratedKernel(self)
            self._executor = ex
        return self._executor

    @property
    def is_hand_written(self):
        from pyop2_amd.codegen import GeneratedKernel
        return not isinstance(self.compile(), GeneratedKernel)

    @property
    def _wrapper_name(self):
        import warnings
        warnings.warn("GlobalKernel._wrapper_name is a deprecated alias for GlobalKernel.name",
                      DeprecationWarning)
        return self.name

    @cached_property
    def argtypes(self):
        """ctypes types of ``__call__``'s ``(start, end, *ptrs)``
        (``global_kernel.py:410-417``): two int32, then one pointer for an
        extruded set's layers, one for a Subset's indices, one per data
        pointer of each argument (a Mat is one descriptor, a mixed argument
        one per component or block), and one per distinct map."""
        n = int(self._extruded) + int(bool(self._subset))
        for a in self.arguments:
            n += len(a.arguments) if isinstance(a, (MixedDatKernelArg, MixedMatKernelArg)) else 1
        n += len({id(m) for a in self.arguments for m in a.maps})
        return (ctypes.c_int32, ctypes.c_int32) + (ctypes.c_void_p,) * n

    def num_flops(self, iterset):
        """Flops of one execution over ``iterset``: the local kernel's, times
        the layers an extruded loop runs per column (``global_kernel.py:419-429``)."""
        size = 1
        if getattr(iterset, "_extruded", False):
            region = self._iteration_region
            layers = np.mean(iterset.layers_array[:, 1] - iterset.layers_array[:, 0])
            if region == IterationRegion.INTERIOR_FACETS:
                size = layers - 2
            elif region not in {IterationRegion.TOP, IterationRegion.BOTTOM}:
                size = layers - 1
        return size * self.local_kernel.num_flops
