"""Parallel loops on one device rank (the reference's ``pyop2/parloop.py``
interface).

A loop is built from a ``GlobalKernel`` (the compiled form of an element
kernel over one argument signature), an iteration set and one bound
argument per kernel argument.  Running it follows the reference's order
(``Parloop.__call__``, ``parloop.py:243-260``):

1. written data gets a new version, INC Globals are swapped for zeroed
   temporaries and Mats are put in the insertion state of their access;
2. the core band of the iteration set is computed while the halos of the
   indirectly read Dats would be exchanged, then the owned band;
3. written Dats are marked dirty on the device, the temporaries are added to
   their Globals, and the reverse halo exchange of INC/MIN/MAX Dats would
   finish (on one rank both exchanges have nothing to move).

Each band is one call ``global_kernel(comm, start, end, *pointers)``, the
pointers being the iteration set's, each argument's data, then each map
once, in first-use order (``parloop.py:203-212``); the GlobalKernel's
executor launches HIP kernels on the current stream.

Bound arguments are one class, ``_Bound``, configured per kind (Dat,
MixedDat, Mat, MixedMat, Global, passthrough) by the reference's names
(``DatParloopArg`` ...).  The legacy form -- ``dat(op2.INC, map)`` -- is one
class, ``_Legacy``, that knows its kernel-argument descriptor and its bound
argument, again under the reference's names (``DatLegacyArg`` ...).
"""
import numpy as np

from pyop2_amd.configuration import configuration
from pyop2_amd.exceptions import KernelTypeError, MapValueError, SetTypeError
from pyop2_amd.global_kernel import (DatKernelArg, GlobalKernel, GlobalKernelArg, MatKernelArg,
                                     MixedDatKernelArg, MixedMatKernelArg, PassthroughKernelArg)
from pyop2_amd.local_kernel import CStringLocalKernel, LocalKernel
from pyop2_amd.types.access import Access
from pyop2_amd.types.dat import AbstractDat
from pyop2_amd.types.glob import Global
from pyop2_amd.types.mat import Mat
from pyop2_amd.types.set import ExtrudedSet, MixedSet, Set, Subset

__all__ = ["Parloop", "LegacyParloop", "ParLoop", "par_loop", "parloop",
           "DatParloopArg", "GlobalParloopArg", "MatParloopArg", "MixedDatParloopArg",
           "MixedMatParloopArg", "PassthroughParloopArg",
           "DatLegacyArg", "GlobalLegacyArg", "MatLegacyArg", "MixedDatLegacyArg",
           "MixedMatLegacyArg", "PassthroughArg"]

_INSERT_STATE = {Access.INC: Mat.ADD_VALUES, Access.WRITE: Mat.INSERT_VALUES}
_EXCHANGED_BEFORE = frozenset((Access.READ, Access.RW, Access.INC, Access.MIN, Access.MAX))
_EXCHANGED_AFTER = frozenset((Access.INC, Access.MIN, Access.MAX))


# ---------------------------------------------------------------------------
# Bound arguments

def _require_values(m):
    """A map without values cannot be read by a loop (``MapValueError``)."""
    if configuration["type_check"] and m.iterset.total_size > 0 and len(m._table) == 0:
        raise MapValueError(f"{m} is not initialized")


class _Bound:
    """A kernel argument's data bound to the maps it is read through.

    ``data`` is the object; ``maps`` the plain maps the loop reads, in the
    order their pointers are passed (a MixedMap contributes its components,
    a Mat its row then column maps).
    """

    kind = None

    def __init__(self, data, maps=()):
        self.data = data
        self.maps = tuple(maps)
        for m in self.maps:
            _require_values(m)

    @property
    def _kernel_args_(self):
        return self.data._kernel_args_

    @property
    def map_kernel_args(self):
        return tuple(p for m in self.maps for p in m._kernel_args_)

    def __repr__(self):
        return "%s(%r, maps=%r)" % (type(self).__name__, self.data, self.maps)


class DatParloopArg(_Bound):
    kind = "dat"

    def __init__(self, data, map_=None):
        self.map_ = map_
        super().__init__(data, () if map_ is None else (map_,))


class MixedDatParloopArg(_Bound):
    kind = "mixed_dat"

    def __init__(self, data, map_):
        self.map_ = map_
        super().__init__(data, map_.split)


class MatParloopArg(_Bound):
    kind = "mat"

    def __init__(self, data, maps, lgmaps=None):
        self.lgmaps = lgmaps
        super().__init__(data, maps)


class MixedMatParloopArg(_Bound):
    kind = "mixed_mat"

    def __init__(self, data, mixed_maps):
        self.mixed_maps = tuple(mixed_maps)
        rmap, cmap = self.mixed_maps
        super().__init__(data, tuple(rmap.split) + tuple(cmap.split))


class GlobalParloopArg(_Bound):
    kind = "global"

    def __init__(self, data):
        super().__init__(data)


class PassthroughParloopArg(_Bound):
    kind = "passthrough"

    def __init__(self, data):
        super().__init__(data)

    @property
    def _kernel_args_(self):
        return (self.data,)


# ---------------------------------------------------------------------------
# Checks made when a loop is built

def _direct_set(iterset):
    """The set a direct Dat of a loop over ``iterset`` must live on."""
    base = iterset.superset if isinstance(iterset, Subset) else iterset
    return base.parent if isinstance(base, ExtrudedSet) else base


def _check_iteration(iterset, global_knl, arguments):
    """The iteration set is a plain set, and every direct Dat and every map
    of the arguments lives on it (``MapValueError`` otherwise)."""
    if not configuration["type_check"]:
        return
    if not isinstance(iterset, Set):
        raise SetTypeError("Iteration set is of the wrong type")
    if isinstance(iterset, MixedSet):
        raise SetTypeError("Cannot iterate over mixed sets")
    base = iterset.superset if isinstance(iterset, Subset) else iterset
    for i, (garg, arg) in enumerate(zip(global_knl.arguments, arguments)):
        if isinstance(garg, DatKernelArg) and garg.is_direct and arg.data.dataset.set != _direct_set(iterset):
            raise MapValueError(f"argument {i} is a direct Dat on another set than the loop's")
        for j, m in enumerate(arg.maps):
            if m.iterset != base and m.iterset not in base:
                raise MapValueError(f"map {j} of argument {i} does not start from the loop's set")


def _check_frozen_halos(local_knl, arguments):
    """A Dat whose halo is frozen keeps the access it was frozen with."""
    for larg, arg in zip(local_knl.arguments, arguments):
        if isinstance(arg.data, AbstractDat):
            for d in arg.data:
                if d._halo_frozen and d._frozen_access_mode != larg.access:
                    raise RuntimeError("Dats with frozen halos must always be accessed with the same access mode")


# ---------------------------------------------------------------------------
# The loop

class Parloop:
    """``global_knl`` over ``iterset`` with bound ``arguments``."""

    def __init__(self, global_knl, iterset, arguments):
        if len(global_knl.arguments) != len(arguments):
            raise ValueError(f"the kernel takes {len(global_knl.arguments)} arguments, "
                             f"{len(arguments)} were given")
        _check_iteration(iterset, global_knl, arguments)
        _check_frozen_halos(global_knl.local_kernel, arguments)
        self.global_kernel = global_knl
        self.iterset = iterset
        self.comm = iterset.comm
        self.arguments = list(arguments)
        # INC Globals accumulate into a zeroed temporary of the kernel's dtype
        # per call; the temporary is then added to the Global
        self.reduced_globals = {}
        for k, (larg, garg, arg) in enumerate(self._triples()):
            if isinstance(garg, GlobalKernelArg) and larg.access == Access.INC:
                tmp = Global(garg.dim, data=np.zeros(arg.data.dim, dtype=larg.dtype), dtype=larg.dtype,
                             comm=self.comm)
                self.reduced_globals[tmp] = arg
                self.arguments[k] = GlobalParloopArg(tmp)

    def _triples(self):
        """(local kernel argument, kernel argument descriptor, bound argument)."""
        return zip(self.global_kernel.local_kernel.arguments, self.global_kernel.arguments,
                   self.arguments)

    @property
    def local_kernel(self):
        return self.global_kernel.local_kernel

    @property
    def accesses(self):
        return self.local_kernel.accesses

    @property
    def zipped_arguments(self):
        return tuple(self._triples())

    @property
    def arglist(self):
        """The executor's pointers: the iteration set's (an extruded set's
        layers, a Subset's indices), each argument's data, then each map's
        pointer once, in first-use order."""
        data = [p for arg in self.arguments for p in arg._kernel_args_]
        maps = dict.fromkeys(p for arg in self.arguments for p in arg.map_kernel_args)
        return tuple(self.iterset._kernel_args_) + tuple(data) + tuple(maps)

    @property
    def num_flops(self):
        return self.global_kernel.num_flops(self.iterset)

    def replace_data(self, index, new_argument):
        self.arguments[index].data = new_argument
        self._static = None

    def _static_parts(self):
        """What every call reads but only the bound arguments fix: the written
        arguments, the indirect Dats each exchange half visits (first use
        order, each once), and whether any of them has a halo at all."""
        st = self.__dict__.get("_static")
        if st is None:
            written = [(garg, arg, larg.access) for larg, garg, arg in self._triples()
                       if larg.access != Access.READ and arg.kind != "passthrough"]
            exch = []
            for larg, garg, arg in self._triples():
                if isinstance(garg, (DatKernelArg, MixedDatKernelArg)) and garg.is_indirect:
                    exch.append((larg.access, arg.data))
            halo = any(data._has_halo() for _, data in exch)
            st = self._static = (written, exch, halo)
        return st

    # -- running --------------------------------------------------------------------
    def _compute(self, part):
        """One band ``[part.offset, part.offset + part.size)``."""
        if part.size:
            self.global_kernel(self.comm, part.offset, part.offset + part.size, *self.arglist,
                               parloop=self)

    def compute(self):
        self()

    def __call__(self):
        self._prepare_outputs()
        self._exchange("global_to_local_begin", _EXCHANGED_BEFORE)
        self._compute(self.iterset.core_part)
        self._exchange("global_to_local_end", _EXCHANGED_BEFORE)
        self._compute(self.iterset.owned_part)
        self._exchange("local_to_global_begin", _EXCHANGED_AFTER)
        self._mark_outputs()
        self._add_reductions()
        self._exchange("local_to_global_end", _EXCHANGED_AFTER)

    def _written(self):
        return self._static_parts()[0]

    def _prepare_outputs(self):
        """New versions for written data (a temporary's Global is the one
        written), zeroed reduction temporaries, Mat insertion states."""
        for garg, arg, access in self._written():
            target = self.reduced_globals.get(arg.data, arg).data
            target.increment_dat_version()
        for tmp in self.reduced_globals:
            tmp._device_array().zero_()
            tmp._device_modified()
        for garg, arg, access in self._written():
            if arg.kind in ("mat", "mixed_mat"):
                arg.data.change_assembly_state(_INSERT_STATE[access])

    def _exchange(self, step, modes):
        """Calls ``step`` (a halo exchange half) once per distinct indirect Dat
        argument accessed with one of ``modes`` (nothing to do when none of
        them has a halo, as on one rank)."""
        _, exch, halo = self._static_parts()
        if not halo:
            return
        done = []
        for access, data in exch:
            if access in modes and not any(data is d for d in done):
                done.append(data)
                getattr(data, step)(access)

    def _mark_outputs(self):
        for garg, arg, access in self._written():
            if arg.kind in ("dat", "mixed_dat"):
                arg.data.halo_valid = False
                arg.data._device_modified()
            elif arg.kind == "global":
                arg.data._device_modified()
            elif arg.kind == "mat":
                arg.data.assembly_state = _INSERT_STATE[access]
            elif arg.kind == "mixed_mat":
                arg.data.change_assembly_state(_INSERT_STATE[access])

    def _add_reductions(self):
        """Each temporary added into its Global (one rank: the reference's
        allreduce, ``parloop.py:409-452``, is the identity)."""
        for tmp, arg in self.reduced_globals.items():
            glob = arg.data
            dst, src = glob._device_array(), tmp._device_array()
            src = src.to(dst.dtype) if src.dtype != dst.dtype else src
            if dst.dtype.is_floating_point or dst.dtype.is_signed:
                dst.add_(src)
            else:
                # torch cannot add unsigned integers: the signed view of the
                # same width adds to the same wrapped bits
                dst.view(_signed_of(dst)).add_(src.view(_signed_of(src)))
            glob._device_modified()


def _signed_of(t):
    from pyop2_amd import device as dev
    torch = dev.torch()
    return {1: torch.int8, 2: torch.int16, 4: torch.int32, 8: torch.int64}[t.element_size()]


# ---------------------------------------------------------------------------
# The legacy form: data(access, maps)

class _Legacy:
    """One ``data(access, ...)`` argument of a legacy par_loop: it gives the
    kernel argument's descriptor (``global_kernel_arg``) and its bound
    argument (``parloop_arg``)."""

    def __init__(self, data, access):
        self.data = data
        self.access = access

    @property
    def dtype(self):
        return self.data.dtype


class GlobalLegacyArg(_Legacy):
    @property
    def global_kernel_arg(self):
        return GlobalKernelArg(self.data.dim)

    @property
    def parloop_arg(self):
        return GlobalParloopArg(self.data)


class DatLegacyArg(_Legacy):
    def __init__(self, data, map_, access):
        super().__init__(data, access)
        self.map_ = map_

    @property
    def global_kernel_arg(self):
        from pyop2_amd.types.dat import DatView
        return DatKernelArg(self.data.dataset.dim,
                            None if self.map_ is None else self.map_._global_kernel_arg,
                            index=self.data.index if isinstance(self.data, DatView) else None)

    @property
    def parloop_arg(self):
        return DatParloopArg(self.data, self.map_)


class MatLegacyArg(_Legacy):
    def __init__(self, data, maps, access, lgmaps=None, needs_unrolling=False):
        super().__init__(data, access)
        self.maps = tuple(maps)
        self.lgmaps = lgmaps
        self.needs_unrolling = needs_unrolling

    @property
    def global_kernel_arg(self):
        return MatKernelArg(self.data.dims, tuple(m._global_kernel_arg for m in self.maps),
                            unroll=self.needs_unrolling)

    @property
    def parloop_arg(self):
        return MatParloopArg(self.data, self.maps, self.lgmaps)


class MixedDatLegacyArg(_Legacy):
    def __init__(self, data, map_, access):
        super().__init__(data, access)
        self.map_ = map_

    @property
    def global_kernel_arg(self):
        return MixedDatKernelArg(tuple(DatKernelArg(d.dataset.dim, m._global_kernel_arg)
                                       for d, m in zip(self.data.split, self.map_.split)))

    @property
    def parloop_arg(self):
        return MixedDatParloopArg(self.data, self.map_)


class MixedMatLegacyArg(_Legacy):
    def __init__(self, data, maps, access):
        super().__init__(data, access)
        self.maps = tuple(maps)

    @property
    def global_kernel_arg(self):
        rows, cols = self.maps
        blocks = tuple(MatKernelArg(self.data[i, j].dims, (r._global_kernel_arg, c._global_kernel_arg))
                       for i, r in enumerate(rows) for j, c in enumerate(cols))
        return MixedMatKernelArg(blocks, (len(rows), len(cols)))

    @property
    def parloop_arg(self):
        return MixedMatParloopArg(self.data, self.maps)


class PassthroughArg(_Legacy):
    """A raw value handed to the kernel unchanged (accessed RW)."""

    def __init__(self, dtype, data):
        super().__init__(data, Access.RW)
        self._dtype = dtype

    @property
    def dtype(self):
        return self._dtype

    @property
    def global_kernel_arg(self):
        return PassthroughKernelArg()

    @property
    def parloop_arg(self):
        return PassthroughParloopArg(self.data)


_GLOBAL_KERNELS = {}


def _shared_global_kernel(local_knl, kernel_args, **kwargs):
    """One GlobalKernel (and so one compiled executor) per cache key, as the
    reference reuses its compiled wrappers (``pyop2/compilation.py:427-489``)."""
    gk = GlobalKernel(local_knl, kernel_args, **kwargs)
    try:
        return _GLOBAL_KERNELS.setdefault(gk.cache_key, gk)
    except TypeError:   # a key part that cannot be hashed: not shared
        return gk


def LegacyParloop(local_knl, iterset, *args, **kwargs):
    """A Parloop from legacy ``data(access, ...)`` arguments: the element
    kernel's accesses and dtypes are taken from them."""
    if not all(isinstance(a, _Legacy) for a in args):
        raise ValueError("LegacyParloop only expects LegacyArg arguments")
    if not isinstance(iterset, Set):
        raise SetTypeError("Iteration set is of the wrong type")
    local_knl.accesses = tuple(a.access for a in args)
    if isinstance(local_knl, CStringLocalKernel):
        local_knl.dtypes = tuple(a.dtype for a in args)
    extruded = bool(getattr(iterset, "_extruded", False))
    kwargs.update(subset=isinstance(iterset, Subset), extruded=extruded,
                  constant_layers=extruded and bool(iterset.constant_layers),
                  extruded_periodic=bool(getattr(iterset, "_extruded_periodic", False)))
    global_knl = _shared_global_kernel(local_knl, tuple(a.global_kernel_arg for a in args), **kwargs)
    return Parloop(global_knl, iterset, [a.parloop_arg for a in args])


ParLoop = LegacyParloop


def parloop(knl, *args, **kwargs):
    """Build and run a loop: a GlobalKernel with bound arguments, or an
    element kernel with legacy arguments."""
    if isinstance(knl, GlobalKernel):
        loop = Parloop(knl, *args, **kwargs)
    elif isinstance(knl, LocalKernel):
        loop = LegacyParloop(knl, *args, **kwargs)
    else:
        raise KernelTypeError
    loop()


def par_loop(*args, **kwargs):
    parloop(*args, **kwargs)
