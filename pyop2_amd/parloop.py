"""Parallel loops (``pyop2/parloop.py``) on one device rank.

Same construction, checks and call sequence as the reference
(``Parloop.__call__``, ``parloop.py:243-260``): version bump, zeroed
temporaries for INC Globals (``:514-530``), halo hooks, ``_compute`` over
the core part and then the owned part, Global finalisation and Mat state
updates.  ``_compute`` hands ``(start, end, *ptrs)`` to the GlobalKernel,
whose executor launches HIP kernels on the current torch stream.
"""
import abc
from dataclasses import dataclass
from typing import Any, Optional, Tuple

import numpy as np

from pyop2_amd.configuration import configuration
from pyop2_amd.exceptions import KernelTypeError, MapValueError, SetTypeError
from pyop2_amd.global_kernel import (DatKernelArg, GlobalKernel, GlobalKernelArg, MatKernelArg,
                                     MixedDatKernelArg, MixedMatKernelArg, PassthroughKernelArg)
from pyop2_amd.local_kernel import CStringLocalKernel, LocalKernel
from pyop2_amd.types.access import Access
from pyop2_amd.types.dat import AbstractDat, Dat
from pyop2_amd.types.data_carrier import DataCarrier
from pyop2_amd.types.glob import Constant, Global
from pyop2_amd.types.map import Map
from pyop2_amd.types.mat import Mat
from pyop2_amd.types.set import ExtrudedSet, MixedSet, Set, Subset
from pyop2_amd.utils import cached_property


class ParloopArg(abc.ABC):

    @staticmethod
    def check_map(m):
        if configuration["type_check"]:
            if m.iterset.total_size > 0 and len(m._table) == 0:
                raise MapValueError(f"{m} is not initialized")


@dataclass
class GlobalParloopArg(ParloopArg):
    data: Global

    @property
    def _kernel_args_(self):
        return self.data._kernel_args_

    @property
    def map_kernel_args(self):
        return ()

    @property
    def maps(self):
        return ()


@dataclass
class DatParloopArg(ParloopArg):
    data: Dat
    map_: Optional[Map] = None

    def __post_init__(self):
        if self.map_ is not None:
            self.check_map(self.map_)

    @property
    def _kernel_args_(self):
        return self.data._kernel_args_

    @property
    def map_kernel_args(self):
        return self.map_._kernel_args_ if self.map_ else ()

    @property
    def maps(self):
        return (self.map_,) if self.map_ is not None else ()


@dataclass
class MatParloopArg(ParloopArg):
    data: Mat
    maps: Tuple[Map, Map]
    lgmaps: Optional[Any] = None

    def __post_init__(self):
        for m in self.maps:
            self.check_map(m)

    @property
    def _kernel_args_(self):
        return self.data._kernel_args_

    @property
    def map_kernel_args(self):
        rmap, cmap = self.maps
        return tuple(rmap._kernel_args_) + tuple(cmap._kernel_args_)


@dataclass
class MixedDatParloopArg(ParloopArg):
    """A MixedDat through a MixedMap: component data pointers, then the
    component maps (``parloop.py:90-117``)."""
    data: Any
    map_: Any

    def __post_init__(self):
        for m in self.map_.split:
            self.check_map(m)

    @property
    def _kernel_args_(self):
        return self.data._kernel_args_

    @property
    def map_kernel_args(self):
        return self.map_._kernel_args_

    @property
    def maps(self):
        return tuple(self.map_.split)


@dataclass
class MixedMatParloopArg(ParloopArg):
    """A MixedMat through a pair of MixedMaps; ``maps`` lists the row
    component maps, then the column component maps."""
    data: Any
    mixed_maps: Tuple[Any, Any]

    def __post_init__(self):
        for m in self.maps:
            self.check_map(m)

    @property
    def _kernel_args_(self):
        return self.data._kernel_args_

    @property
    def map_kernel_args(self):
        return sum((m._kernel_args_ for m in self.maps), ())

    @property
    def maps(self):
        rmap, cmap = self.mixed_maps
        return tuple(rmap.split) + tuple(cmap.split)


@dataclass
class PassthroughParloopArg(ParloopArg):
    data: int

    @property
    def _kernel_args_(self):
        return (self.data,)

    @property
    def map_kernel_args(self):
        return ()

    @property
    def maps(self):
        return ()


class Parloop:
    """A parallel loop invocation of ``global_knl`` over ``iterset``."""

    def __init__(self, global_knl, iterset, arguments):
        if len(global_knl.arguments) != len(arguments):
            raise ValueError("You are trying to pass in a different number of "
                             "arguments than the kernel is expecting")
        self.check_iterset(iterset, global_knl, arguments)
        self._check_frozen_access_modes(global_knl.local_kernel, arguments)
        self.global_kernel = global_knl
        self.iterset = iterset
        self.comm = iterset.comm
        self.arguments, self.reduced_globals = self.prepare_reduced_globals(arguments, global_knl)

    @property
    def local_kernel(self):
        return self.global_kernel.local_kernel

    @property
    def accesses(self):
        return self.local_kernel.accesses

    @property
    def arglist(self):
        """Pointers in the reference's order (``parloop.py:203-212``): the
        iteration set's (an extruded set's layers, a Subset's indices), the
        data of each argument, then the maps' pointers without repeats, in
        first-use order."""
        arglist = self.iterset._kernel_args_
        for d in self.arguments:
            arglist += d._kernel_args_
        maps = {p: None for d in self.arguments for p in d.map_kernel_args}
        return arglist + tuple(maps)

    @property
    def zipped_arguments(self):
        return self.zip_arguments(self.global_kernel, self.arguments)

    def replace_data(self, index, new_argument):
        self.arguments[index].data = new_argument

    def _compute(self, part):
        """Execute the kernel over ``[part.offset, part.offset + part.size)``."""
        if part.size == 0:
            return
        self.global_kernel(self.comm, part.offset, part.offset + part.size, *self.arglist,
                           parloop=self)

    @cached_property
    def num_flops(self):
        return self.global_kernel.num_flops(self.iterset)

    def compute(self):
        self()

    def __call__(self):
        """Execute the kernel over all members of the iteration space."""
        self.increment_dat_version()
        self.zero_global_increments()
        self.replace_lgmaps()
        self.global_to_local_begin()
        self._compute(self.iterset.core_part)
        self.global_to_local_end()
        self._compute(self.iterset.owned_part)
        self.local_to_global_begin()
        self.update_arg_data_state()
        self.finalize_global_increments()
        self.local_to_global_end()

    def increment_dat_version(self):
        for lk_arg, gk_arg, pl_arg in self.zipped_arguments:
            if isinstance(pl_arg, PassthroughParloopArg):
                continue
            if lk_arg.access is not Access.READ:
                if pl_arg.data in self.reduced_globals:
                    self.reduced_globals[pl_arg.data].data.increment_dat_version()
                else:
                    pl_arg.data.increment_dat_version()

    def zero_global_increments(self):
        for g in self.reduced_globals.keys():
            g._device_array().zero_()
            g._device_modified()

    def replace_lgmaps(self):
        for lk_arg, gk_arg, pl_arg in self.zipped_arguments:
            if isinstance(gk_arg, (MatKernelArg, MixedMatKernelArg)):
                new_state = {Access.INC: Mat.ADD_VALUES, Access.WRITE: Mat.INSERT_VALUES}[lk_arg.access]
                pl_arg.data.change_assembly_state(new_state)

    def _indirect_dat_idxs(self, modes):
        seen = set()
        out = []
        for i, (lk, gk, pl) in enumerate(self.zipped_arguments):
            if isinstance(gk, (DatKernelArg, MixedDatKernelArg)) and gk.is_indirect and pl.data not in seen \
                    and lk.access in modes:
                out.append(i)
                seen.add(pl.data)
        return out

    def global_to_local_begin(self):
        for i in self._indirect_dat_idxs({Access.READ, Access.RW, Access.INC, Access.MIN, Access.MAX}):
            self.arguments[i].data.global_to_local_begin(self.accesses[i])

    def global_to_local_end(self):
        for i in self._indirect_dat_idxs({Access.READ, Access.RW, Access.INC, Access.MIN, Access.MAX}):
            self.arguments[i].data.global_to_local_end(self.accesses[i])

    def local_to_global_begin(self):
        for i in self._indirect_dat_idxs({Access.INC, Access.MIN, Access.MAX}):
            self.arguments[i].data.local_to_global_begin(self.accesses[i])

    def local_to_global_end(self):
        for i in self._indirect_dat_idxs({Access.INC, Access.MIN, Access.MAX}):
            self.arguments[i].data.local_to_global_end(self.accesses[i])

    def finalize_global_increments(self):
        """Add the loop's increments into the user's Globals (single rank:
        the reference's allreduce of ``parloop.py:409-452`` is the identity)."""
        torch = None
        for tmp, glob in self.reduced_globals.items():
            g = glob.data
            dst, src = g._device_array(), tmp._device_array()
            if dst.dtype != src.dtype:
                src = src.to(dst.dtype)
            if not dst.dtype.is_floating_point and not dst.dtype.is_signed:
                # torch has no unsigned add: two's-complement add of the
                # same-width signed view gives the same (wrapping) bits
                from pyop2_amd import device as dev
                torch = torch or dev.torch()
                signed = {1: torch.int8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[dst.element_size()]
                dst.view(signed).add_(src.view(signed))
            else:
                dst.add_(src)
            g._device_modified()

    def update_arg_data_state(self):
        for i, (wrapper_arg, d) in enumerate(zip(self.global_kernel.arguments, self.arguments)):
            access = self.accesses[i]
            if access is Access.READ:
                continue
            if isinstance(wrapper_arg, (DatKernelArg, MixedDatKernelArg)):
                d.data.halo_valid = False
                d.data._device_modified()
            elif isinstance(wrapper_arg, GlobalKernelArg):
                d.data._device_modified()
            elif isinstance(wrapper_arg, MatKernelArg):
                d.data.assembly_state = {Access.WRITE: Mat.INSERT_VALUES, Access.INC: Mat.ADD_VALUES}[access]
            elif isinstance(wrapper_arg, MixedMatKernelArg):
                d.data.change_assembly_state({Access.WRITE: Mat.INSERT_VALUES, Access.INC: Mat.ADD_VALUES}[access])

    @classmethod
    def check_iterset(cls, iterset, global_knl, arguments):
        if not configuration["type_check"]:
            return
        if not isinstance(iterset, Set):
            raise SetTypeError("Iteration set is of the wrong type")
        if isinstance(iterset, MixedSet):
            raise SetTypeError("Cannot iterate over mixed sets")
        if isinstance(iterset, Subset):
            iterset = iterset.superset
        for i, (lk_arg, gk_arg, pl_arg) in enumerate(cls.zip_arguments(global_knl, arguments)):
            if isinstance(gk_arg, DatKernelArg) and gk_arg.is_direct:
                _iterset = iterset.parent if isinstance(iterset, ExtrudedSet) else iterset
                if pl_arg.data.dataset.set != _iterset:
                    raise MapValueError(f"Iterset of direct arg {i} does not match parloop iterset")
            for j, m in enumerate(pl_arg.maps):
                if m.iterset != iterset and m.iterset not in iterset:
                    raise MapValueError(f"Iterset of arg {i} map {j} does not match parloop iterset")

    @classmethod
    def _check_frozen_access_modes(cls, local_knl, arguments):
        for lknl_arg, pl_arg in zip(local_knl.arguments, arguments):
            if isinstance(pl_arg.data, AbstractDat):
                if any(d._halo_frozen and d._frozen_access_mode != lknl_arg.access for d in pl_arg.data):
                    raise RuntimeError("Dats with frozen halos must always be accessed with the same access mode")

    def prepare_reduced_globals(self, arguments, global_knl):
        """Swap INC Globals for zeroed temporaries (``parloop.py:514-530``)."""
        arguments = list(arguments)
        reduced_globals = {}
        for i, (lk_arg, gk_arg, pl_arg) in enumerate(self.zip_arguments(global_knl, arguments)):
            if isinstance(gk_arg, GlobalKernelArg) and lk_arg.access == Access.INC:
                tmp = Global(gk_arg.dim, data=np.zeros(pl_arg.data.dim, dtype=lk_arg.dtype),
                             dtype=lk_arg.dtype, comm=self.comm)
                reduced_globals[tmp] = pl_arg
                arguments[i] = GlobalParloopArg(tmp)
        return arguments, reduced_globals

    @staticmethod
    def zip_arguments(global_knl, arguments):
        return tuple(zip(global_knl.local_kernel.arguments, global_knl.arguments, arguments))


class LegacyArg(abc.ABC):
    """Old-style par_loop input carrying data and access together."""

    @property
    @abc.abstractmethod
    def global_kernel_arg(self):
        """The corresponding kernel argument descriptor."""

    @property
    @abc.abstractmethod
    def parloop_arg(self):
        """The corresponding ParloopArg."""


@dataclass
class GlobalLegacyArg(LegacyArg):
    data: Global
    access: Access

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def global_kernel_arg(self):
        return GlobalKernelArg(self.data.dim)

    @property
    def parloop_arg(self):
        return GlobalParloopArg(self.data)


@dataclass
class DatLegacyArg(LegacyArg):
    data: Dat
    map_: Optional[Map]
    access: Access

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def global_kernel_arg(self):
        from pyop2_amd.types.dat import DatView
        map_arg = self.map_._global_kernel_arg if self.map_ is not None else None
        index = self.data.index if isinstance(self.data, DatView) else None
        return DatKernelArg(self.data.dataset.dim, map_arg, index=index)

    @property
    def parloop_arg(self):
        return DatParloopArg(self.data, self.map_)


@dataclass
class MatLegacyArg(LegacyArg):
    data: Mat
    maps: Tuple[Map, Map]
    access: Access
    lgmaps: Optional[Tuple[Any, Any]] = None
    needs_unrolling: Optional[bool] = False

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def global_kernel_arg(self):
        return MatKernelArg(self.data.dims, tuple(m._global_kernel_arg for m in self.maps),
                            unroll=self.needs_unrolling)

    @property
    def parloop_arg(self):
        return MatParloopArg(self.data, self.maps, self.lgmaps)


@dataclass
class MixedDatLegacyArg(LegacyArg):
    data: Any
    map_: Any
    access: Access

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def global_kernel_arg(self):
        return MixedDatKernelArg(tuple(DatKernelArg(d.dataset.dim, m._global_kernel_arg)
                                       for d, m in zip(self.data.split, self.map_.split)))

    @property
    def parloop_arg(self):
        return MixedDatParloopArg(self.data, self.map_)


@dataclass
class MixedMatLegacyArg(LegacyArg):
    data: Any
    maps: Tuple[Any, Any]
    access: Access

    @property
    def dtype(self):
        return self.data.dtype

    @property
    def global_kernel_arg(self):
        rmap, cmap = self.maps
        nr, nc = len(rmap), len(cmap)
        blocks = tuple(MatKernelArg(self.data[i, j].dims, (rmap[i]._global_kernel_arg, cmap[j]._global_kernel_arg))
                       for i in range(nr) for j in range(nc))
        return MixedMatKernelArg(blocks, (nr, nc))

    @property
    def parloop_arg(self):
        return MixedMatParloopArg(self.data, self.maps)


@dataclass
class PassthroughArg(LegacyArg):
    access = Access.RW
    dtype: Any
    data: int

    @property
    def global_kernel_arg(self):
        return PassthroughKernelArg()

    @property
    def parloop_arg(self):
        return PassthroughParloopArg(self.data)


_GLOBAL_KERNELS = {}


def _cached_global_kernel(local_knl, global_knl_args, subset, kwargs):
    """GlobalKernels (and their compiled executors) are reused for equal
    cache keys, as the reference reuses compiled wrappers
    (``pyop2/compilation.py:427-489``, memory cache keyed by ``cache_key``)."""
    gk = GlobalKernel(local_knl, global_knl_args, subset=subset, **kwargs)
    try:
        return _GLOBAL_KERNELS.setdefault(gk.cache_key, gk)
    except TypeError:   # unhashable key component: no caching
        return gk


def ParLoop(*args, **kwargs):
    return LegacyParloop(*args, **kwargs)


def LegacyParloop(local_knl, iterset, *args, **kwargs):
    """Create a :class:`Parloop` from :class:`LegacyArg` inputs."""
    if not all(isinstance(a, LegacyArg) for a in args):
        raise ValueError("LegacyParloop only expects LegacyArg arguments")
    if not isinstance(iterset, Set):
        raise SetTypeError("Iteration set is of the wrong type")
    local_knl.accesses = tuple(a.access for a in args)
    if isinstance(local_knl, CStringLocalKernel):
        local_knl.dtypes = tuple(a.dtype for a in args)
    global_knl_args = tuple(a.global_kernel_arg for a in args)
    subset = isinstance(iterset, Subset)
    extruded = getattr(iterset, "_extruded", False)
    kwargs = dict(kwargs, extruded=extruded, constant_layers=extruded and iterset.constant_layers,
                  extruded_periodic=bool(getattr(iterset, "_extruded_periodic", False)))
    global_knl = _cached_global_kernel(local_knl, global_knl_args, subset, kwargs)
    parloop_args = tuple(a.parloop_arg for a in args)
    return Parloop(global_knl, iterset, parloop_args)


def par_loop(*args, **kwargs):
    parloop(*args, **kwargs)


def parloop(knl, *args, **kwargs):
    """Construct and execute a :class:`Parloop`."""
    if isinstance(knl, GlobalKernel):
        Parloop(knl, *args, **kwargs)()
    elif isinstance(knl, LocalKernel):
        LegacyParloop(knl, *args, **kwargs)()
    else:
        raise KernelTypeError
