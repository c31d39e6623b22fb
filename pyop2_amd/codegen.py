"""Generated gather/compute/scatter kernels for arbitrary C-string kernels.

For a local kernel with no hand-written implementation, the executor emits
one HIP kernel per GlobalKernel signature: each work-item takes one
iteration element ``n`` and follows the reference wrapper's per-element
semantics (``pyop2/codegen/builder.py``):

* indirect Dat READ/RW/MIN/MAX arguments are gathered into private buffers
  ``buf[i*dim + j] = dat[map[n][i]*dim + j]``; INC/WRITE buffers start at
  zero; direct Dats pass ``&dat[n*dim]``;
* Globals: READ passes the device array; INC uses a zeroed buffer and
  MIN/MAX a copy, reduced over the wavefront before one atomic per wave;
* Mats: a zeroed ``(ar_r*rdim) x (ar_c*cdim)`` buffer, added (INC) or
  stored (WRITE) block by block at the slot found by binary search of the
  row's sorted column indices;
* the user's C function is compiled unchanged as a device function
  (``#pragma clang force_cuda_host_device``) and called with pointer
  adaptors that convert to whatever pointer types it declares;
* unpack: INC adds (device-scope atomics, or plain adds inside a colour),
  MIN/MAX atomic min/max, WRITE/RW stores.

Indirect WRITE/RW arguments and Mat WRITE need exclusive targets, so such
loops run coloured (greedy colouring, one launch per colour); INC-only
loops run with atomics.  Sources are compiled for gfx950 with hiprtc and
the code objects cached in memory and under ``<cache_dir>/hip``.
"""
import ctypes
import hashlib
import os
import re
import threading

import numpy as np

from pyop2_amd import device as dev
from pyop2_amd.configuration import configuration
from pyop2_amd.datatypes import as_cstr
from pyop2_amd.exceptions import CompilationError
from pyop2_amd.global_kernel import (DatKernelArg, GlobalKernelArg, MatKernelArg, MixedDatKernelArg,
                                     MixedMatKernelArg)
from pyop2_amd.types import region as reg
from pyop2_amd.types.access import INC, MAX, MIN, READ, RW, WRITE
from pyop2_amd.types.set import Subset

__all__ = ["GeneratedKernel", "generate_source", "compile_options"]

_PRELUDE = r"""
typedef signed char int8_t;
typedef unsigned char uint8_t;
typedef short int16_t;
typedef unsigned short uint16_t;
typedef int int32_t;
typedef unsigned int uint32_t;
typedef long long int64_t;
typedef unsigned long long uint64_t;
#define restrict __restrict__

/* Converts to any pointer type a user kernel declares for an argument. */
struct op2_ptr_ {
  void *p;
  template <typename U> __host__ __device__ operator U *() const { return static_cast<U *>(p); }
};

template <typename T> __device__ inline void op2_atomic_add(T *p, T v) { atomicAdd(p, v); }
__device__ inline void op2_atomic_add(int64_t *p, int64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(v));
}
__device__ inline void op2_atomic_add(uint64_t *p, uint64_t v) {
  atomicAdd(reinterpret_cast<unsigned long long *>(p), static_cast<unsigned long long>(v));
}
template <typename T> __device__ inline void op2_atomic_min(T *p, T v) { atomicMin(p, v); }
template <typename T> __device__ inline void op2_atomic_max(T *p, T v) { atomicMax(p, v); }
__device__ inline void op2_atomic_min(int64_t *p, int64_t v) { atomicMin(reinterpret_cast<long long *>(p), (long long)v); }
__device__ inline void op2_atomic_max(int64_t *p, int64_t v) { atomicMax(reinterpret_cast<long long *>(p), (long long)v); }
__device__ inline void op2_atomic_min(uint64_t *p, uint64_t v) { atomicMin(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v); }
__device__ inline void op2_atomic_max(uint64_t *p, uint64_t v) { atomicMax(reinterpret_cast<unsigned long long *>(p), (unsigned long long)v); }
__device__ inline void op2_atomic_min(double *p, double v) {
  unsigned long long *a = reinterpret_cast<unsigned long long *>(p);
  unsigned long long old = *a, assumed;
  while (__longlong_as_double(old) > v) {
    assumed = old;
    old = atomicCAS(a, assumed, __double_as_longlong(v));
    if (old == assumed) break;
  }
}
__device__ inline void op2_atomic_max(double *p, double v) {
  unsigned long long *a = reinterpret_cast<unsigned long long *>(p);
  unsigned long long old = *a, assumed;
  while (__longlong_as_double(old) < v) {
    assumed = old;
    old = atomicCAS(a, assumed, __double_as_longlong(v));
    if (old == assumed) break;
  }
}
__device__ inline void op2_atomic_min(float *p, float v) {
  unsigned int *a = reinterpret_cast<unsigned int *>(p);
  unsigned int old = *a, assumed;
  while (__uint_as_float(old) > v) {
    assumed = old;
    old = atomicCAS(a, assumed, __float_as_uint(v));
    if (old == assumed) break;
  }
}
__device__ inline void op2_atomic_max(float *p, float v) {
  unsigned int *a = reinterpret_cast<unsigned int *>(p);
  unsigned int old = *a, assumed;
  while (__uint_as_float(old) < v) {
    assumed = old;
    old = atomicCAS(a, assumed, __float_as_uint(v));
    if (old == assumed) break;
  }
}
template <typename T> __device__ inline T op2_min(T a, T b) { return b < a ? b : a; }
template <typename T> __device__ inline T op2_max(T a, T b) { return b > a ? b : a; }

template <typename T> __device__ inline T op2_wave_sum(T v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <typename T> __device__ inline T op2_wave_min(T v) {
  for (int o = 32; o > 0; o >>= 1) v = op2_min(v, (T)__shfl_xor(v, o, 64));
  return v;
}
template <typename T> __device__ inline T op2_wave_max(T v) {
  for (int o = 32; o > 0; o >>= 1) v = op2_max(v, (T)__shfl_xor(v, o, 64));
  return v;
}

/* Remainder of a periodic column's layer offset, a in [0, 2n). */
__device__ inline int32_t op2_rem(int32_t a, int32_t n) { return a < n ? a : a - n; }

__device__ inline int32_t op2_find_slot(const int32_t *__restrict__ rowptr,
                                         const int32_t *__restrict__ colidx,
                                         int32_t row, int32_t col) {
  int32_t l = rowptr[row], h = rowptr[row + 1] - 1;
  while (l <= h) {
    const int32_t m = l + ((h - l) >> 1);
    const int32_t c = colidx[m];
    if (c == col) return m;
    if (c < col) l = m + 1; else h = m - 1;
  }
  return -1;
}
"""

_STD_HEADER = re.compile(r'^\s*#\s*include\s*<(math|stdlib|stdio|string|stdint|stdbool|float|limits|complex|inttypes|stddef|assert)\.h>\s*$',
                         re.M)

# Integer types without wavefront shuffles / atomics (Global reductions refused).
_SHUFFLE_T = {"int8_t": "int32_t", "uint8_t": "uint32_t", "int16_t": "int32_t",
              "uint16_t": "uint32_t"}


def compile_options():
    return ["--offload-arch=gfx950", "-O3", "-std=c++17", "-munsafe-fp-atomics",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _strip_headers(text):
    return _STD_HEADER.sub("", text)


def _kernel_maps(garg):
    """MapKernelArgs of a kernel argument in the order of its parloop
    argument's ``maps``."""
    if isinstance(garg, DatKernelArg):
        return (garg.map_,) if garg.map_ is not None else ()
    if isinstance(garg, MatKernelArg):
        return tuple(garg.maps)
    if isinstance(garg, (MixedDatKernelArg, MixedMatKernelArg)):
        return tuple(garg.maps)
    return ()


def _dat_store(acc, tgt, val):
    """Statement storing a Dat pack entry ``val`` into ``tgt`` for ``acc``
    (atomic when the launch is not coloured)."""
    if acc == INC:
        return f"if (atomic) op2_atomic_add(&{tgt}, {val}); else {tgt} += {val};"
    if acc == MIN:
        return f"if (atomic) op2_atomic_min(&{tgt}, {val}); else {tgt} = op2_min({tgt}, {val});"
    if acc == MAX:
        return f"if (atomic) op2_atomic_max(&{tgt}, {val}); else {tgt} = op2_max({tgt}, {val});"
    return f"{tgt} = {val};"


def _region_bounds(region, periodic=False):
    """C expressions of the first and one-past-last layer a column runs in
    ``region``, from its ``bottom_`` and cell end ``cend_``
    (``builder.py:771-795``; a column with no cells runs no top cell; a
    periodic column has as many interior facets as cells)."""
    region = reg.normalise(region)
    if region == reg.ON_INTERIOR_FACETS and periodic:
        return "bottom_", "cend_"
    if region == reg.ON_BOTTOM:
        return "bottom_", "bottom_ + 1"
    if region == reg.ON_TOP:
        return "(cend_ - 1 > bottom_ ? cend_ - 1 : bottom_)", "cend_"
    if region == reg.ON_INTERIOR_FACETS:
        return "bottom_", "cend_ - 1"
    return "bottom_", "cend_"


def generate_source(gk, map_index, block_flags=None):
    """HIP source of the wrapper kernel ``op2_wrap`` for GlobalKernel ``gk``.

    ``map_index[(arg, j)]`` is the position of an argument's j-th map in the
    unique-map list; ``block_flags[arg]`` says whether a Mat argument is
    stored as BSR (else scalar CSR over dofs).
    """
    block_flags = block_flags or {}
    lk = gk.local_kernel
    extruded = getattr(gk, "_extruded", False)
    region = getattr(gk, "_iteration_region", None) if extruded else None
    # interior facets: every map names the cell below, then the cell above
    nf = reg.facet_factor(region)
    periodic = extruded and bool(getattr(gk, "_extruded_periodic", False))
    # map offsets per unique map (extruded maps step by offset[i] per layer)
    # and, on a periodic set, their offset quotients
    offsets, quots = {}, {}
    for k, garg in enumerate(gk.arguments):
        for j, mk in enumerate(_kernel_maps(garg)):
            if mk.offset is not None and extruded:
                offsets[map_index[(k, j)]] = mk.offset
                if periodic and mk.offset_quotient is not None:
                    quots[map_index[(k, j)]] = mk.offset_quotient

    def _step(m, ar, i, f):
        """Layer offset of entry ``i`` of map ``m``: the cells above the
        column's bottom (``f`` adds the cell above), wrapped on a periodic
        set (``builder.py:103-121``)."""
        st = "lyr_" if f is None else f"(lyr_ + {f})"
        if not periodic:
            return st
        if m not in quots:
            return f"op2_rem({st}, nl_)"
        return f"(op2_rem({st} + quo{m}[{i}], nl_) - op2_rem(quo{m}[{i}], nl_))"

    def _mv(m, ar, i):
        """Entry ``i`` of unique map ``m`` (arity ``ar``) for element ``n``
        (and layer); with interior facets ``i`` runs over ``2 * ar`` entries,
        the cell below first (``builder.py:81-129``)."""
        if nf == 1:
            base = f"map{m}[(int64_t)n * {ar} + {i}]"
            if m in offsets:
                return f"(({base}) >= 0 ? ({base}) + off{m}[{i}] * {_step(m, ar, i, None)} : -1)"
            return base
        if m not in offsets:
            raise ValueError("ON_INTERIOR_FACETS needs extruded maps (maps with an offset)")
        i_ = f"({i}) % {ar}"
        base = f"map{m}[(int64_t)n * {ar} + {i_}]"
        return f"(({base}) >= 0 ? ({base}) + off{m}[{i_}] * {_step(m, ar, i_, f'({i}) / {ar}')} : -1)"
    headers = "\n".join(_strip_headers(h) for h in lk.headers)
    code = _strip_headers(lk.code)
    params = ["int32_t lo", "int32_t hi", "const int32_t *__restrict__ list", "int32_t atomic"]
    pre, call, post, red = [], [], [], []
    nmaps = len(set(map_index.values()))
    for k, (larg, garg) in enumerate(zip(lk.arguments, gk.arguments)):
        acc = larg.access
        t = as_cstr(np.dtype(larg.dtype))
        if isinstance(garg, DatKernelArg):
            dim = int(np.prod(garg.dim))
            # a DatView passes one component: stride dim, offset comp, width 1
            # (builder.py:303-410, view_index)
            comp = 0 if garg.index is None else int(np.ravel_multi_index(garg.index, garg.dim))
            stride, dim = dim, (dim if garg.index is None else 1)
            params.append(f"{t} *__restrict__ arg{k}")
            if garg.map_ is None:
                call.append(f"op2_ptr_{{(void *)(arg{k} + (int64_t)n * {stride} + {comp})}}")
                continue
            m = map_index[(k, 0)]
            ar = garg.map_.arity
            ear = nf * ar
            pre.append(f"    int32_t mi{k}[{ear}];")
            pre.append(f"    for (int i = 0; i < {ear}; ++i) mi{k}[i] = {_mv(m, ar, 'i')};")
            pre.append(f"    {t} buf{k}[{ear * dim}];")
            if acc in (INC, WRITE):
                pre.append(f"    for (int q = 0; q < {ear * dim}; ++q) buf{k}[q] = ({t})0;")
            else:
                pre.append(f"    for (int i = 0; i < {ear}; ++i) for (int j = 0; j < {dim}; ++j)"
                           f" buf{k}[i * {dim} + j] = mi{k}[i] >= 0 ? arg{k}[(int64_t)mi{k}[i] * {stride} + {comp} + j] : ({t})0;")
            call.append(f"op2_ptr_{{(void *)buf{k}}}")
            if acc == READ:
                continue
            tgt = f"arg{k}[(int64_t)mi{k}[i] * {stride} + {comp} + j]"
            val = f"buf{k}[i * {dim} + j]"
            post.append(f"    for (int i = 0; i < {ear}; ++i) {{ if (mi{k}[i] < 0) continue;"
                        f" for (int j = 0; j < {dim}; ++j) {{ {_dat_store(acc, tgt, val)} }} }}")
        elif isinstance(garg, GlobalKernelArg):
            dim = int(np.prod(garg.dim)) if len(garg.dim) else 1
            params.append(f"{t} *__restrict__ arg{k}")
            if acc == READ:
                call.append(f"op2_ptr_{{(void *)arg{k}}}")
                continue
            # declared before the element guard so every lane takes part in the reduction
            if acc == INC:
                red.append(("decl", f"  {t} buf{k}[{dim}];\n  for (int q = 0; q < {dim}; ++q) buf{k}[q] = ({t})0;"))
            else:
                red.append(("decl", f"  {t} buf{k}[{dim}];\n  for (int q = 0; q < {dim}; ++q) buf{k}[q] = arg{k}[q];"))
            call.append(f"op2_ptr_{{(void *)buf{k}}}")
            fn = {INC: "sum", MIN: "min", MAX: "max"}.get(acc)
            if fn is None:
                raise NotImplementedError(f"Global access {acc!r} in a par_loop")
            if t in _SHUFFLE_T:
                raise NotImplementedError(f"Global reductions of {t} are not supported on the device")
            aop = {INC: "op2_atomic_add", MIN: "op2_atomic_min", MAX: "op2_atomic_max"}[acc]
            red.append(("post", f"  for (int q = 0; q < {dim}; ++q) {{ {t} v = op2_wave_{fn}(buf{k}[q]);"
                                f" if (lane == 0) {aop}(&arg{k}[q], v); }}"))
        elif isinstance(garg, MixedDatKernelArg):
            # one buffer: the packs of the components one after the other
            comps = garg.arguments
            total = sum(nf * c.map_.arity * int(np.prod(c.dim)) for c in comps)
            pre.append(f"    {t} buf{k}[{total}];")
            if acc in (INC, WRITE):
                pre.append(f"    for (int q = 0; q < {total}; ++q) buf{k}[q] = ({t})0;")
            off = 0
            for c, comp in enumerate(comps):
                dim, ar, m = int(np.prod(comp.dim)), comp.map_.arity, map_index[(k, c)]
                ear = nf * ar
                a_, mi_ = f"arg{k}_{c}", f"mi{k}_{c}"
                params.append(f"{t} *__restrict__ {a_}")
                pre.append(f"    int32_t {mi_}[{ear}];")
                pre.append(f"    for (int i = 0; i < {ear}; ++i) {mi_}[i] = {_mv(m, ar, 'i')};")
                if acc not in (INC, WRITE):
                    pre.append(f"    for (int i = 0; i < {ear}; ++i) for (int j = 0; j < {dim}; ++j)"
                               f" buf{k}[{off} + i * {dim} + j] = {mi_}[i] >= 0 ?"
                               f" {a_}[(int64_t){mi_}[i] * {dim} + j] : ({t})0;")
                if acc != READ:
                    tgt = f"{a_}[(int64_t){mi_}[i] * {dim} + j]"
                    val = f"buf{k}[{off} + i * {dim} + j]"
                    post.append(f"    for (int i = 0; i < {ear}; ++i) {{ if ({mi_}[i] < 0) continue;"
                                f" for (int j = 0; j < {dim}; ++j) {{ {_dat_store(acc, tgt, val)} }} }}")
                off += ear * dim
            call.append(f"op2_ptr_{{(void *)buf{k}}}")
        elif isinstance(garg, MixedMatKernelArg):
            # one element matrix; block (i, j) is rows of row component i
            # times columns of column component j, added into block Mat (i, j)
            nr, nc = garg.shape
            blocks = garg.arguments
            rds = [blocks[i * nc].dims[0][0][0] for i in range(nr)]
            cds = [blocks[j].dims[0][0][1] for j in range(nc)]
            rars = [nf * blocks[i * nc].maps[0].arity for i in range(nr)]
            cars = [nf * blocks[j].maps[1].arity for j in range(nc)]
            roff = [sum(rars[q] * rds[q] for q in range(i)) for i in range(nr)]
            coff = [sum(cars[q] * cds[q] for q in range(j)) for j in range(nc)]
            NR, NC = roff[-1] + rars[-1] * rds[-1], coff[-1] + cars[-1] * cds[-1]
            pre.append(f"    {t} buf{k}[{NR * NC}];")
            pre.append(f"    for (int q = 0; q < {NR * NC}; ++q) buf{k}[q] = ({t})0;")
            call.append(f"op2_ptr_{{(void *)buf{k}}}")
            flags = block_flags.get(k) or (False,) * (nr * nc)
            store = ("if (atomic) op2_atomic_add(&{tgt}, {val}); else {tgt} += {val};" if acc == INC
                     else "{tgt} = {val};")
            for i in range(nr):
                for j in range(nc):
                    b = i * nc + j
                    rd, cd, rar, car = rds[i], cds[j], rars[i], cars[j]
                    rm, cm = map_index[(k, i)], map_index[(k, nr + j)]
                    params += [f"const int32_t *__restrict__ rowptr{k}_{b}",
                               f"const int32_t *__restrict__ colidx{k}_{b}", f"{t} *__restrict__ arg{k}_{b}"]
                    val = f"buf{k}[({roff[i]} + ii * {rd} + ri) * {NC} + {coff[j]} + jj * {cd} + cj]"
                    # flags[b]: True = BSR block Mat; False = scalar CSR;
                    # ("offset", r0, c0) = block of a monolithic Mat, rows and
                    # columns shifted by its offsets in the parent's CSR
                    shift = flags[b] if isinstance(flags[b], tuple) else ("offset", 0, 0)
                    if flags[b] is True:
                        tgt = f"arg{k}_{b}[(int64_t)s_ * {rd * cd} + ri * {cd} + cj]"
                        post.append(
                            f"    for (int ii = 0; ii < {rar}; ++ii) {{ const int32_t r_ = {_mv(rm, rar // nf, 'ii')}; if (r_ < 0) continue;\n"
                            f"      for (int jj = 0; jj < {car}; ++jj) {{ const int32_t c_ = {_mv(cm, car // nf, 'jj')}; if (c_ < 0) continue;\n"
                            f"        const int32_t s_ = op2_find_slot(rowptr{k}_{b}, colidx{k}_{b}, r_, c_); if (s_ < 0) {{ atomicAdd(status, 1); continue; }}\n"
                            f"        for (int ri = 0; ri < {rd}; ++ri) for (int cj = 0; cj < {cd}; ++cj) {{ {store.format(tgt=tgt, val=val)} }} }} }}")
                    else:
                        tgt = f"arg{k}_{b}[s_]"
                        post.append(
                            f"    for (int ii = 0; ii < {rar}; ++ii) {{ const int32_t r_ = {_mv(rm, rar // nf, 'ii')}; if (r_ < 0) continue;\n"
                            f"      for (int ri = 0; ri < {rd}; ++ri) for (int jj = 0; jj < {car}; ++jj) {{ const int32_t c_ = {_mv(cm, car // nf, 'jj')}; if (c_ < 0) continue;\n"
                            f"        for (int cj = 0; cj < {cd}; ++cj) {{ const int32_t s_ = op2_find_slot(rowptr{k}_{b}, colidx{k}_{b}, {shift[1]} + r_ * {rd} + ri, {shift[2]} + c_ * {cd} + cj);"
                            f" if (s_ < 0) {{ atomicAdd(status, 1); continue; }} {store.format(tgt=tgt, val=val)} }} }} }}")
        elif isinstance(garg, MatKernelArg):
            rd, cd = garg.dims[0][0]
            params += [f"const int32_t *__restrict__ rowptr{k}", f"const int32_t *__restrict__ colidx{k}",
                       f"{t} *__restrict__ arg{k}"]
            rm, cm = map_index[(k, 0)], map_index[(k, 1)]
            rar, car = nf * garg.maps[0].arity, nf * garg.maps[1].arity
            ncol = car * cd
            pre.append(f"    {t} buf{k}[{rar * rd * ncol}];")
            pre.append(f"    for (int q = 0; q < {rar * rd * ncol}; ++q) buf{k}[q] = ({t})0;")
            call.append(f"op2_ptr_{{(void *)buf{k}}}")
            block = bool(block_flags.get(k, False))
            if acc == INC:
                store = "if (atomic) op2_atomic_add(&{tgt}, {val}); else {tgt} += {val};"
            else:
                store = "{tgt} = {val};"
            if block:
                tgt = f"arg{k}[(int64_t)s_ * {rd * cd} + ri * {cd} + cj]"
                val = f"buf{k}[(i * {rd} + ri) * {ncol} + j * {cd} + cj]"
                post.append(
                    f"    for (int i = 0; i < {rar}; ++i) {{ const int32_t r_ = {_mv(rm, rar // nf, 'i')}; if (r_ < 0) continue;\n"
                    f"      for (int j = 0; j < {car}; ++j) {{ const int32_t c_ = {_mv(cm, car // nf, 'j')}; if (c_ < 0) continue;\n"
                    f"        const int32_t s_ = op2_find_slot(rowptr{k}, colidx{k}, r_, c_); if (s_ < 0) {{ atomicAdd(status, 1); continue; }}\n"
                    f"        for (int ri = 0; ri < {rd}; ++ri) for (int cj = 0; cj < {cd}; ++cj) {{ {store.format(tgt=tgt, val=val)} }} }} }}")
            else:
                tgt = f"arg{k}[s_]"
                val = f"buf{k}[(i * {rd} + ri) * {ncol} + j * {cd} + cj]"
                post.append(
                    f"    for (int i = 0; i < {rar}; ++i) {{ const int32_t r_ = {_mv(rm, rar // nf, 'i')}; if (r_ < 0) continue;\n"
                    f"      for (int ri = 0; ri < {rd}; ++ri) for (int j = 0; j < {car}; ++j) {{ const int32_t c_ = {_mv(cm, car // nf, 'j')}; if (c_ < 0) continue;\n"
                    f"        for (int cj = 0; cj < {cd}; ++cj) {{ const int32_t s_ = op2_find_slot(rowptr{k}, colidx{k}, r_ * {rd} + ri, c_ * {cd} + cj);"
                    f" if (s_ < 0) {{ atomicAdd(status, 1); continue; }} {store.format(tgt=tgt, val=val)} }} }} }}")
        else:
            raise NotImplementedError("Passthrough arguments are not supported on the device")
    for m in range(nmaps):
        params.append(f"const int32_t *__restrict__ map{m}")
    params.append("int32_t *__restrict__ status")
    if extruded:
        params.append("const int32_t *__restrict__ layers")
    offdecl = [f"    const int32_t off{m}[{len(o)}] = {{{', '.join(str(int(v)) for v in o)}}};"
               for m, o in sorted(offsets.items())]
    offdecl += [f"    const int32_t quo{m}[{len(o)}] = {{{', '.join(str(int(v)) for v in o)}}};"
                for m, o in sorted(quots.items())]
    decls = "\n".join(s for kind, s in red if kind == "decl")
    reds = "\n".join(s for kind, s in red if kind == "post")
    src = [_PRELUDE,
           "#pragma clang force_cuda_host_device begin",
           headers, code,
           "#pragma clang force_cuda_host_device end",
           'extern "C" __global__ void __launch_bounds__(256)',
           "op2_wrap(" + ", ".join(params) + ")",
           "{",
           "  const int32_t t_ = lo + (int32_t)(blockIdx.x * blockDim.x + threadIdx.x);",
           "  const bool active = t_ < hi;",
           "  const int lane = threadIdx.x & 63;",
           "  (void)lane; (void)status;",
           decls,
           "  if (active) {",
           "    const int32_t n = list ? list[t_] : t_;",
           *([] if not extruded else [
               *offdecl,
               "    const int32_t lay_ = %s;" % ("0" if gk._constant_layers else "n"),
               "    const int32_t bottom_ = layers[2 * lay_], cend_ = layers[2 * lay_ + 1] - 1;",
               "    const int32_t lo_ = %s, hi_ = %s;" % _region_bounds(region, periodic),
               "    const int32_t nl_ = cend_ - bottom_;",
               "    (void)nl_;",
               "    for (int32_t layer_ = lo_; layer_ < hi_; ++layer_) {",
               "    const int32_t lyr_ = layer_ - bottom_;",
               "    (void)lyr_;"]),
           *pre,
           f"    {lk.name}({', '.join(call + (['layer_'] if getattr(gk, '_pass_layer_arg', False) else []))});",
           *post,
           *(["    }"] if extruded else []),
           "  }",
           reds,
           "}"]
    return "\n".join(s for s in src if s is not None)


_gen_lock = threading.Lock()
_modules = {}


def _code_object(src, opts):
    """Code object bytes for ``src`` (memory, then disk cache, then hiprtc)."""
    key = hashlib.sha256(("\0".join(opts) + "\0" + src).encode()).hexdigest()[:32]
    cdir = os.path.join(configuration["cache_dir"], "hip")
    path = os.path.join(cdir, key + ".co")
    try:
        with open(path, "rb") as f:
            return key, f.read()
    except OSError:
        pass
    L = dev.lib()
    copts = (ctypes.c_char_p * len(opts))(*[o.encode() for o in opts])
    size = ctypes.c_size_t(0)
    log = ctypes.create_string_buffer(65536)
    rc = L.op2hip_jit_code_object(src.encode(), copts, len(opts), None, ctypes.byref(size), log, len(log))
    if rc != 0:
        raise CompilationError("device compilation of the generated wrapper failed:\n"
                               + log.value.decode(errors="replace") + "\n--- source ---\n" + src)
    buf = ctypes.create_string_buffer(size.value)
    rc = L.op2hip_jit_code_object(src.encode(), copts, len(opts), buf, ctypes.byref(size), log, len(log))
    dev.check(rc, "jit_code_object")
    data = buf.raw[:size.value]
    try:
        os.makedirs(cdir, exist_ok=True)
        tmp = path + ".tmp%d" % os.getpid()
        with open(tmp, "wb") as f:
            f.write(data)
        os.replace(tmp, path)
    except OSError:
        pass
    return key, data


def _module(src):
    opts = compile_options()
    key, data = _code_object(src, opts)
    with _gen_lock:
        mod = _modules.get(key)
        if mod is None:
            dev.device()
            L = dev.lib()
            h = ctypes.c_void_p()
            buf = ctypes.create_string_buffer(data, len(data))
            dev.check(L.op2hip_jit_load(buf, ctypes.byref(h)), "jit_load")
            mod = _modules[key] = h
    return mod


class GeneratedKernel:
    """Executor of a generated wrapper (``(start, end, ptrs, parloop)``)."""

    BLOCK = 256

    def __init__(self, gk):
        self.gk = gk
        self.strategy_used = None
        self._src = {}
        self._mods = {}

    def _unique_maps(self, parloop):
        seen, index = [], {}
        for k, a in enumerate(parloop.arguments):
            for j, m in enumerate(a.maps):
                for q, s in enumerate(seen):
                    if s is m:
                        index[(k, j)] = q
                        break
                else:
                    index[(k, j)] = len(seen)
                    seen.append(m)
        return seen, index

    def _flags(self, parloop):
        out = []
        for k, (garg, a) in enumerate(zip(self.gk.arguments, parloop.arguments)):
            if isinstance(garg, MatKernelArg):
                out.append((k, a.data._block))
            elif isinstance(garg, MixedMatKernelArg):
                out.append((k, tuple(getattr(b, "_gen_flag", b._block) for b in a.data)))
        return tuple(out)

    def source(self, parloop):
        flags = self._flags(parloop)
        src = self._src.get(flags)
        if src is None:
            _, index = self._unique_maps(parloop)
            src = self._src[flags] = generate_source(self.gk, index, dict(flags))
        return src

    def _conflict_maps(self, parloop):
        """Maps whose targets must be exclusive within a launch."""
        out = []
        need = False
        for larg, garg, a in zip(self.gk.local_kernel.arguments, self.gk.arguments, parloop.arguments):
            if isinstance(garg, DatKernelArg) and garg.map_ is not None:
                if larg.access in (RW, WRITE):
                    need = True
                if larg.access in (RW, WRITE, INC, MIN, MAX):
                    out.append(a.maps[0])
            elif isinstance(garg, MixedDatKernelArg):
                if larg.access in (RW, WRITE):
                    need = True
                if larg.access in (RW, WRITE, INC, MIN, MAX):
                    out.extend(a.maps)
            elif isinstance(garg, MatKernelArg):
                if larg.access == WRITE:
                    need = True
                out.append(a.maps[0])
            elif isinstance(garg, MixedMatKernelArg):
                if larg.access == WRITE:
                    need = True
                out.extend(a.maps[:garg.shape[0]])
        return (out if need else []), need

    def __call__(self, start, end, ptrs, parloop):
        if parloop is None:
            raise RuntimeError("generated loops are launched by a Parloop")
        flags = self._flags(parloop)
        mod = self._mods.get(flags)
        if mod is None:
            mod = self._mods[flags] = _module(self.source(parloop))
        maps, _ = self._unique_maps(parloop)
        conflict, need_colour = self._conflict_maps(parloop)
        if configuration["strategy"] == "coloured" and not need_colour:
            conflict, _ = [], False
            for larg, garg, a in zip(self.gk.local_kernel.arguments, self.gk.arguments, parloop.arguments):
                if (isinstance(garg, DatKernelArg) and garg.map_ is not None and larg.access != READ) \
                        or isinstance(garg, MatKernelArg):
                    conflict.append(a.maps[0])
                elif isinstance(garg, MixedDatKernelArg) and larg.access != READ:
                    conflict.extend(a.maps)
                elif isinstance(garg, MixedMatKernelArg):
                    conflict.extend(a.maps[:garg.shape[0]])
            need_colour = bool(conflict)
        vals = []
        mat_status = []
        for garg, a in zip(self.gk.arguments, parloop.arguments):
            if isinstance(garg, MatKernelArg):
                mat = a.data
                vals += [ctypes.c_void_p(dev.ptr(mat.rowptr)), ctypes.c_void_p(dev.ptr(mat.colidx)),
                         ctypes.c_void_p(dev.ptr(mat.device_values))]
                mat_status.append(mat._status)
            elif isinstance(garg, MixedMatKernelArg):
                for blk in a.data:
                    vals += [ctypes.c_void_p(dev.ptr(blk.rowptr)), ctypes.c_void_p(dev.ptr(blk.colidx)),
                             ctypes.c_void_p(dev.ptr(blk.device_values))]
                mat_status.append(a.data[0, 0]._status)
            elif isinstance(garg, MixedDatKernelArg):
                vals += [ctypes.c_void_p(dev.ptr(d._device_array())) for d in a.data]
            else:
                vals.append(ctypes.c_void_p(dev.ptr(a.data._device_array())))
        vals += [ctypes.c_void_p(dev.ptr(m.device_values)) for m in maps]
        status = mat_status[0] if mat_status else _dummy_status()
        vals.append(ctypes.c_void_p(dev.ptr(status)))
        iterset = parloop.iterset
        extruded = getattr(self.gk, "_extruded", False)
        if extruded:
            region = getattr(self.gk, "_iteration_region", None)
            for m in maps:
                if getattr(m, "offset", None) is not None:
                    m.check_extruded(iterset, region)
            vals.append(ctypes.c_void_p(dev.ptr(iterset.device_layers)))
        L = dev.lib()
        s = dev.stream()

        def launch(lo, hi, lst, atomic):
            allv = [ctypes.c_int32(lo), ctypes.c_int32(hi), ctypes.c_void_p(lst), ctypes.c_int32(atomic)] + vals
            argv = (ctypes.c_void_p * len(allv))(*[ctypes.addressof(v) for v in allv])
            grid = (hi - lo + self.BLOCK - 1) // self.BLOCK
            with dev.timed_launch("op2_wrap:" + self.gk.name):
                rc = L.op2hip_jit_launch(mod, b"op2_wrap", grid, self.BLOCK, 0, argv, s)
            dev.check(rc, f"launch of generated {self.gk.name}")

        if need_colour and conflict:
            from pyop2_amd import plans
            if isinstance(iterset, Subset):
                elems = iterset.indices[start:end]
            else:
                elems = np.arange(start, end, dtype=np.int32)
            key = (iterset._plan_key(), start, end, tuple(id(m) for m in conflict),
                   int(getattr(self.gk, "_iteration_region", None) or 0))
            if extruded:
                # a column's targets over all its layers (a thread runs the
                # layers of its column one after the other)
                region = getattr(self.gk, "_iteration_region", None)
                cols = np.ascontiguousarray(elems, dtype=np.int32)   # column ids, in launch order
                targets = [(_ColumnTargets(m, iterset, cols, region), id(m.toset)) for m in conflict]
                plan = plans.colour_plan(conflict[0], np.arange(len(cols), dtype=np.int32), targets,
                                         ("extruded",) + key)
                plan = _column_plan(plan, cols, conflict[0], ("extruded-cols",) + key)
            else:
                plan = plans.colour_plan(conflict[0], elems, [(m, id(m.toset)) for m in conflict], key)
            ep = dev.ptr(plan.elems)
            for c in range(plan.ncolours):
                lo, hi = int(plan.offsets[c]), int(plan.offsets[c + 1])
                if hi > lo:
                    launch(lo, hi, ep, 0)
            self.strategy_used = "coloured"
        else:
            lst = dev.ptr(iterset.device_indices) if isinstance(iterset, Subset) else 0
            launch(start, end, lst, 1)
            self.strategy_used = "atomic"


class _ColumnTargets:
    """Duck-typed map for colouring the columns ``cols`` of an extruded loop
    (a range of the set, or a Subset's columns): row c lists every target
    column cols[c] touches over the layers it runs."""

    def __init__(self, m, iterset, cols, region=None):
        self._table = m.column_table(iterset, cols, region)
        self.toset = m.toset


def _column_plan(plan, cols, owner, key):
    """A colour plan over positions in ``cols`` turned into column ids."""
    from pyop2_amd import plans
    got = owner._cache.get(key)
    if got is None:
        elems = cols[dev.to_host(plan.elems).astype(np.int64)].astype(np.int32)
        got = plans.ColourPlan(plan.colours, dev.to_device(elems), plan.offsets, plan.ncolours)
        owner._cache[key] = got
    return got


_STATUS = None


def _dummy_status():
    global _STATUS
    if _STATUS is None:
        _STATUS = dev.zeros((1,), np.int32)
    return _STATUS
