"""Binding of the op2hip C ABI (``include/op2hip.h``) and device state.

The library ``pyop2_amd/_lib/libop2hip.so`` is loaded after torch so that
both share one HIP runtime; device memory and streams come from torch.
There is no CPU fallback: without the library or a gfx950 device every
launch raises :class:`DeviceError`.
"""
import ctypes
import os
import threading

import numpy as np

from pyop2_amd.exceptions import DeviceError, PatternError

__all__ = ["lib", "torch", "device", "stream", "check", "ExecPlan", "MatDesc",
           "MapPair", "ptr", "available", "empty", "zeros", "to_device"]

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libop2hip.so")

OK, ERR_HIP, ERR_ARG, ERR_WORKSPACE, ERR_OVERFLOW, ERR_JIT, ERR_PATTERN = range(7)
(ATOMIC, COLOURED, STAGED, ROWGATHER, ROWGATHER_THREAD, ROWGATHER_SLOTS, ROWGATHER_STAGED,
 ROWGATHER_CELLS, ROWGATHER_ROWLIST) = range(9)
STRATEGIES = {"atomic": ATOMIC, "coloured": COLOURED, "staged": STAGED, "staged_coloured": STAGED,
              "rowgather": ROWGATHER,
              "rowgather_thread": ROWGATHER_THREAD, "rowgather_slots": ROWGATHER_SLOTS,
              "rowgather_staged": ROWGATHER_STAGED, "rowgather_cells": ROWGATHER_CELLS,
              "rowgather_rows": ROWGATHER_ROWLIST}

_i32p = ctypes.POINTER(ctypes.c_int32)


class ExecPlan(ctypes.Structure):
    """``op2hip_exec``."""
    _fields_ = [
        ("strategy", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("colour_elems", ctypes.c_void_p),
        ("colour_offsets", ctypes.c_void_p),
        ("ncolours", ctypes.c_int32),
        ("ntiles", ctypes.c_int32),
        ("tile_ptr", ctypes.c_void_p),
        ("tile_elems", ctypes.c_void_p),
        ("tile_node_ptr", ctypes.c_void_p * 2),
        ("tile_nodes", ctypes.c_void_p * 2),
        ("tile_nexcl", ctypes.c_void_p * 2),
        ("elem_local", ctypes.c_void_p * 2),
        ("max_tile_nodes", ctypes.c_int32 * 2),
        ("tile_shared_off", ctypes.c_void_p * 2),
        ("nshared", ctypes.c_int32 * 2),
        ("shared_nodes", ctypes.c_void_p * 2),
        ("shared_ptr", ctypes.c_void_p * 2),
        ("shared_slots", ctypes.c_void_p * 2),
        ("partials", ctypes.c_void_p * 2),
        ("nbatches", ctypes.c_int32),
        ("batch_rows", ctypes.c_void_p),
        ("batch_max_nnz", ctypes.c_int32),
        ("batch_max_inc", ctypes.c_int32),
        ("batch_max_rows", ctypes.c_int32),
        ("batch_desc", ctypes.c_void_p),
        ("slot_table", ctypes.c_void_p),
        ("gather_count", ctypes.c_void_p),
        ("gather_list", ctypes.c_void_p),
        ("cell_ptr", ctypes.c_void_p),
        ("cells", ctypes.c_void_p),
        ("block_order", ctypes.c_void_p),
        ("contrib_ptr", ctypes.c_void_p),
        ("contrib", ctypes.c_void_p),
        ("max_batch_cells", ctypes.c_int32),
        ("max_batch_contrib", ctypes.c_int32),
        ("row_order", ctypes.c_void_p),
        ("batch_ent", ctypes.c_void_p),
        ("row_base", ctypes.c_void_p),
        ("subset", ctypes.c_void_p),
        ("batch_packed", ctypes.c_void_p),
    ]


class MatDesc(ctypes.Structure):
    """``op2hip_mat``."""
    _fields_ = [
        ("rowptr", ctypes.c_void_p),
        ("colidx", ctypes.c_void_p),
        ("values", ctypes.c_void_p),
        ("nrows", ctypes.c_int32),
        ("ncols", ctypes.c_int32),
        ("rdim", ctypes.c_int32),
        ("cdim", ctypes.c_int32),
        ("status", ctypes.c_void_p),
        ("inc_ptr", ctypes.c_void_p),
        ("inc_ent", ctypes.c_void_p),
    ]


class MapPair(ctypes.Structure):
    """``op2hip_map_pair``."""
    _fields_ = [
        ("rmap", ctypes.c_void_p),
        ("cmap", ctypes.c_void_p),
        ("rarity", ctypes.c_int32),
        ("carity", ctypes.c_int32),
        ("iter_size", ctypes.c_int32),
    ]


_V = ctypes.c_void_p
_I = ctypes.c_int32
_SIGNATURES = {
    "op2hip_abi_version": (ctypes.c_int, []),
    "op2hip_last_error": (ctypes.c_char_p, []),
    "op2hip_init": (ctypes.c_int, [_I]),
    "op2hip_scan_workspace": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(ctypes.c_size_t)]),
    "op2hip_exclusive_scan": (ctypes.c_int, [ctypes.c_int64, _V, _V, ctypes.POINTER(ctypes.c_int64),
                                             _V, ctypes.c_size_t, _V]),
    "op2hip_transpose_workspace": (ctypes.c_int, [_I, ctypes.POINTER(ctypes.c_size_t)]),
    "op2hip_transpose_count": (ctypes.c_int, [_I, _V, _I, _I, _V, ctypes.POINTER(ctypes.c_int64),
                                              _V, ctypes.c_size_t, _V]),
    "op2hip_transpose_fill": (ctypes.c_int, [_I, _V, _I, _I, _V, _V, _V, ctypes.c_size_t, _V]),
    "op2hip_sparsity_count": (ctypes.c_int, [_I, _I, _V, _I, _V, _V, _I, _V,
                                             ctypes.POINTER(ctypes.c_int64), _V, ctypes.c_size_t, _V]),
    "op2hip_sparsity_fill": (ctypes.c_int, [_I, _I, _V, _I, _V, _V, _I, _V, _V, _V]),
    "op2hip_pattern_expand": (ctypes.c_int, [_I, _V, _V, _I, _I, _V, _V, _V]),
    "op2hip_colour_greedy": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _I, _V, _i32p]),
    "op2hip_batched_p1_laplace": (ctypes.c_int, [_V, _V, _I, _I, _I, _I, _V]),
    "op2hip_plan_tiles": (ctypes.c_int, [_I, _I, _I, _V, _V, _V, _I, _I, _i32p, _V, _V, _V, _V, _V,
                                         _V, _V]),
    "op2hip_plan_tile_reduce": (ctypes.c_int, [_I, _V, _V, _V, _I, _V, _i32p, _V, _V, _V]),
    "op2hip_plan_row_batches": (ctypes.c_int, [_I, _V, _V, _I, _I, _V, _i32p, _i32p, _i32p, _i32p]),
    "op2hip_slot_table_bytes": (ctypes.c_int, [_I, ctypes.POINTER(ctypes.c_size_t)]),
    "op2hip_build_slot_table": (ctypes.c_int, [_I, _V, _V, _V, _V, _V, _V, _I, _I, _V, _V, _V, _V]),
    "op2hip_plan_row_list": (ctypes.c_int, [_I, _V, _V, _V, _V, _V, _I, _I, _I, _I, _V, _i32p, _V, _V, _V,
                                            ctypes.POINTER(ctypes.c_int64), _i32p, _i32p, _i32p]),
    "op2hip_plan_cell_gather": (ctypes.c_int, [_I, _V, _V, _I, _I, _V, _V, _V, _V, _V, _V, _V]),
    "op2hip_gather_table_count": (ctypes.c_int, [_I, _I, _V, _V, _I, _V, _V, _V, _V]),
    "op2hip_gather_table_fill": (ctypes.c_int, [_I, _I, _V, _V, _I, _V, _V, _V, _V, _V]),
    "op2hip_wrap_p1_lumped_mass": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _V]),
    "op2hip_wrap_p1_laplace_residual": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _V, _V]),
    "op2hip_wrap_p1_laplace_matrix": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _V]),
    "op2hip_wrap_p2_elasticity_residual": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _V,
                                                          ctypes.c_double, ctypes.c_double, _V, _V]),
    "op2hip_wrap_p2_elasticity_residual_noalloc": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _V,
                                                                  ctypes.c_double, ctypes.c_double, _V,
                                                                  ctypes.c_int64, _V]),
    "op2hip_wrap_p2_elasticity_residual_f32": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _V,
                                                              ctypes.c_double, ctypes.c_double, _V, _V]),
    "op2hip_wrap_p2_elasticity_matrix": (ctypes.c_int, [_I, _I, _V, _V, _V, _V,
                                                        ctypes.c_double, ctypes.c_double, _V, _V]),
    "op2hip_wrap_p2_elasticity_matrix_f32": (ctypes.c_int, [_I, _I, _V, _V, _V, _V,
                                                            ctypes.c_double, ctypes.c_double, _V, _V]),
    "op2hip_wrap_p2_elasticity_fused": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _V, _V,
                                                       ctypes.c_double, ctypes.c_double, _V, _V]),
    "op2hip_wrap_p2_elasticity_fused_f32": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _V, _V,
                                                           ctypes.c_double, ctypes.c_double, _V, _V]),
    "op2hip_wrap_p2_elasticity_adjoint": (ctypes.c_int, [_I, _I, _V, _V, _V, _V, _V, _V, _V,
                                                         ctypes.c_double, ctypes.c_double, _V, _V]),
    "op2hip_jit_compile": (ctypes.c_int, [ctypes.c_char_p, _V, _I, ctypes.POINTER(_V),
                                          ctypes.c_char_p, ctypes.c_size_t]),
    "op2hip_jit_code_object": (ctypes.c_int, [ctypes.c_char_p, _V, _I, _V,
                                              ctypes.POINTER(ctypes.c_size_t), ctypes.c_char_p,
                                              ctypes.c_size_t]),
    "op2hip_jit_load": (ctypes.c_int, [_V, ctypes.POINTER(_V)]),
    "op2hip_jit_unload": (ctypes.c_int, [_V]),
    "op2hip_jit_launch": (ctypes.c_int, [_V, ctypes.c_char_p, ctypes.c_uint32, ctypes.c_uint32,
                                         ctypes.c_uint32, _V, _V]),
}

_state = threading.local()
_lock = threading.Lock()
_lib = None
_torch = None
_device_ready = False


def torch():
    """The torch module (imported lazily: the host API works without it)."""
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


def lib():
    """The loaded op2hip library (no device initialisation)."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                torch()  # share torch's HIP runtime
                if not os.path.exists(_LIB_PATH):
                    raise DeviceError(f"op2hip library not built: {_LIB_PATH} is missing "
                                      "(run `python -m pyop2_amd.build`)")
                L = _Bound(ctypes.CDLL(_LIB_PATH))
                if L.op2hip_abi_version() != 2:
                    raise DeviceError("op2hip ABI version mismatch")
                _lib = L
    return _lib


class _Bound:
    """The library's entry points with their declared argument types only:
    an undeclared name raises instead of passing 64-bit pointers through
    ctypes' default int conversion."""

    def __init__(self, cdll):
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(cdll, name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)
        self._handle = cdll

    def __getattr__(self, name):
        raise AttributeError(f"{name}: no ctypes signature declared for this library entry point")


def available():
    """True if a HIP device and the library are usable."""
    try:
        device()
        return True
    except DeviceError:
        return False


def device():
    """The torch device all op2 data lives on (initialises it once)."""
    global _device_ready
    from pyop2_amd.configuration import configuration
    t = torch()
    idx = configuration["device"]
    if not _device_ready:
        L = lib()
        if not t.cuda.is_available():
            raise DeviceError("no HIP device is visible: pyop2_amd executes par_loops on "
                              "an MI355X (gfx950) only and has no CPU fallback")
        t.cuda.set_device(idx)
        check(L.op2hip_init(idx), "op2hip_init")
        _device_ready = True
    return t.device("cuda", idx)


def stream():
    """Current torch stream of the op2 device, as a void* for the C ABI."""
    t = torch()
    return ctypes.c_void_p(t.cuda.current_stream(device()).cuda_stream)


def check(rc, what=""):
    if rc == OK:
        return
    msg = lib().op2hip_last_error().decode(errors="replace")
    if rc == ERR_PATTERN:
        raise PatternError(f"{what}: {msg}")
    raise DeviceError(f"{what} failed (op2hip error {rc}): {msg}")


def ptr(t):
    """Device address of a torch tensor (or 0 for None)."""
    return 0 if t is None else t.data_ptr()


_TORCH_DTYPES = {}


def torch_dtype(np_dtype):
    t = torch()
    if not _TORCH_DTYPES:
        _TORCH_DTYPES.update({
            np.dtype(np.float64): t.float64, np.dtype(np.float32): t.float32,
            np.dtype(np.int32): t.int32, np.dtype(np.int64): t.int64,
            np.dtype(np.int16): t.int16, np.dtype(np.int8): t.int8,
            np.dtype(np.uint8): t.uint8, np.dtype(np.uint32): t.uint32,
            np.dtype(np.uint64): t.uint64, np.dtype(np.uint16): t.uint16,
            np.dtype(np.bool_): t.bool,
        })
    try:
        return _TORCH_DTYPES[np.dtype(np_dtype)]
    except KeyError:
        raise TypeError(f"dtype {np_dtype} has no device equivalent") from None


def empty(shape, dtype):
    return torch().empty(shape, dtype=torch_dtype(dtype), device=device())


def zeros(shape, dtype):
    return torch().zeros(shape, dtype=torch_dtype(dtype), device=device())


def to_device(a):
    """Copy a host numpy array to a new device tensor (same dtype and shape)."""
    a = np.ascontiguousarray(a)
    t = torch()
    dev = device()
    if a.dtype in (np.dtype(np.uint32), np.dtype(np.uint64), np.dtype(np.uint16)):
        # torch.from_numpy lacks some unsigned types: go through a same-size view
        view = {4: np.int32, 8: np.int64, 2: np.int16}[a.dtype.itemsize]
        src = t.from_numpy(a.view(view))
        return src.to(dev).view(torch_dtype(a.dtype))
    return t.from_numpy(a).to(dev)


def to_host(tensor, out=None):
    """Copy a device tensor into ``out`` (numpy) or a new array."""
    t = torch()
    dt = tensor.dtype
    if dt in (t.uint32, t.uint64, t.uint16):
        view = {t.uint32: t.int32, t.uint64: t.int64, t.uint16: t.int16}[dt]
        host = tensor.view(view).cpu().numpy().view({t.uint32: np.uint32, t.uint64: np.uint64,
                                                     t.uint16: np.uint16}[dt])
    else:
        host = tensor.cpu().numpy()
    if out is None:
        return host
    out[...] = host.reshape(out.shape)
    return out


def workspace(nbytes):
    """A device byte buffer of at least ``nbytes`` (torch allocation)."""
    return torch().empty(max(int(nbytes), 8), dtype=torch().uint8, device=device())


class _LaunchTimer:
    """Records a HIP event pair on the launch stream around every library
    launch while active (``with device.launch_timer() as t: ...``);
    ``t.times(name)`` then gives each launch's duration in ms."""

    def __init__(self):
        self.events = []

    def times(self, name=None):
        out = []
        for n, e0, e1 in self.events:
            if name is None or n == name:
                e1.synchronize()
                out.append(e0.elapsed_time(e1))
        return out


class _TimerState:
    active = None   # process-wide: autograd runs backward on its own thread


_timer = _TimerState()


class launch_timer:
    def __enter__(self):
        self.t = _LaunchTimer()
        _timer.active = self.t
        return self.t

    def __exit__(self, *exc):
        _timer.active = None
        return False


class timed_launch:
    """Context around one launch: event pair when a launch_timer is active."""

    __slots__ = ("name", "e1", "t")

    def __init__(self, name):
        self.name = name
        self.t = _timer.active

    def __enter__(self):
        if self.t is not None:
            tc = torch().cuda
            s = tc.current_stream(device())
            e0 = tc.Event(enable_timing=True)
            self.e1 = tc.Event(enable_timing=True)
            e0.record(s)
            self.t.events.append((self.name, e0, self.e1))
        return self

    def __exit__(self, *exc):
        if self.t is not None:
            self.e1.record(torch().cuda.current_stream(device()))
        return False
