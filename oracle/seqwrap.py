"""Sequential par_loop wrapper, restated in C and compiled with gcc.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

The reference JIT-generates, per par_loop signature, a C function
``wrap_<kernel>(int start, int end, void *args...)`` whose body is, for each
iteration element ``n`` in ``[start, end)`` (``pyop2/codegen/builder.py``):

1. Dat packs: READ/RW/MIN/MAX gather ``buf[i][j] = dat[map[n][i]][j]``;
   INC/WRITE buffers are zeroed (``builder.py:349-362``).  Direct Dats pass
   ``&dat[n]`` with no pack (``builder.py:366-378``).
2. Global packs: READ passes the Global; INC/WRITE use a zeroed buffer,
   RW/MIN/MAX a copy (``builder.py:254-281``).
3. Mat packs: a zeroed ``(ar_r*rdim) x (ar_c*cdim)`` row-major buffer
   (``builder.py:519-545``).
4. ``kernel(buf...)``.
5. Unpack in argument order: INC ``+=``, MIN/MAX ``min/max``, WRITE/RW
   assignment (``builder.py:384-410``, ``283-300``); a Mat arg calls
   ``MatSetValuesBlockedLocal(mat, ar_r, rmap[n], ar_c, cmap[n], buf,
   ADD_VALUES or INSERT_VALUES)`` (``builder.py:554-606``), restated here as
   a sorted-row binary search of each block's column followed by the add or
   the store, block by block in row-major order.

Mixed arguments (``builder.py:413-498``, ``609-680``): a MixedDat is one
buffer holding its components' packs one after the other, each gathered and
scattered through its own map; a MixedMat is one element matrix whose rows
are the row components' packs and whose columns are the column components'
packs, block (i, j) inserted into block Mat (i, j) as above.

A Subset iteration set indirects ``n`` through ``subset_indices``
(``builder.py:725-733``).  Map entries below zero are skipped, as PETSc skips
negative indices in ``MatSetValues``.

Extruded iteration sets (``builder.py:683-800``): column ``n`` runs the
layers of its iteration region (all cells ``[bottom, cend)``, the bottom
cell, the top cell, or the interior facets ``[bottom, cend - 1)``, with
``cend = layers[n][1] - 1``; a column with no cells runs no top cell), and
a map entry of the cell ``l`` layers above the bottom is ``base + offset *
l``; over interior facets every map names the cell below, then the cell
above (``builder.py:81-129``).  On a periodic set the layer offset wraps
with the generated loop's remainder ``a < n ? a : a - n`` over the column's
``n`` cells, shifted by the entry's offset quotient, and a column has as
many interior facets as cells (``builder.py:27-30, 103-121``).

Two builds are produced: ``strict`` (``-O2 -ffp-contract=off``; the parity
oracle) and ``fast`` (the reference's own flags ``-march=native -O3
-ffast-math``, ``pyop2/compilation.py:328``; the CPU baseline).
"""
import ctypes
import hashlib
import os
import subprocess
import tempfile
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

__all__ = ["DatArg", "GlobalArg", "MatArg", "MixedDatArg", "MixedMatArg", "CSRMat", "run", "compile_wrapper",
           "READ", "WRITE", "RW", "INC", "MIN", "MAX"]

READ, WRITE, RW, INC, MIN, MAX = "READ", "WRITE", "RW", "INC", "MIN", "MAX"

_BUILD_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_build")

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

_FLAGS = {
    "strict": ["-O2", "-fPIC", "-shared", "-std=gnu11", "-ffp-contract=off",
               "-fno-fast-math", "-w"],
    "fast": ["-march=native", "-O3", "-ffast-math", "-fPIC", "-shared",
             "-std=gnu11", "-w"],
}


@dataclass
class DatArg:
    """A Dat argument: ``data`` is (total, *dim); ``map_`` (iter_total, arity)
    or None; ``comp`` the flat component a DatView passes (the kernel sees
    that one value per entry, ``builder.py:303-410`` view_index), or None."""
    data: np.ndarray
    access: str
    map_: Optional[np.ndarray] = None
    comp: Optional[int] = None

    @property
    def stride(self):
        return int(np.prod(self.data.shape[1:], dtype=np.int64)) if self.data.ndim > 1 else 1

    @property
    def cdim(self):
        return self.stride if self.comp is None else 1


@dataclass
class GlobalArg:
    data: np.ndarray
    access: str


@dataclass
class CSRMat:
    """Host CSR (or BSR with ``rdim x cdim`` blocks) matrix on a fixed pattern."""
    rowptr: np.ndarray
    colidx: np.ndarray
    values: np.ndarray          # (nnz, rdim, cdim) for block, (nnz,) for scalar
    rdim: int = 1
    cdim: int = 1
    block: bool = False

    def to_dense(self, nrows, ncols):
        """Dense (nrows*rdim, ncols*cdim) copy (block) or (nrows, ncols)."""
        if self.block:
            out = np.zeros((nrows * self.rdim, ncols * self.cdim), dtype=self.values.dtype)
            for r in range(nrows):
                for s in range(self.rowptr[r], self.rowptr[r + 1]):
                    c = self.colidx[s]
                    out[r * self.rdim:(r + 1) * self.rdim,
                        c * self.cdim:(c + 1) * self.cdim] += self.values[s]
            return out
        out = np.zeros((nrows, ncols), dtype=self.values.dtype)
        for r in range(nrows):
            for s in range(self.rowptr[r], self.rowptr[r + 1]):
                out[r, self.colidx[s]] += self.values[s]
        return out


@dataclass
class MatArg:
    mat: CSRMat
    access: str
    rmap: np.ndarray
    cmap: np.ndarray
    rdim: int = 1
    cdim: int = 1


@dataclass
class MixedDatArg:
    """A MixedDat argument: its components (DatArgs with maps, one access)."""
    parts: Sequence[DatArg]
    access: str


@dataclass
class MixedMatArg:
    """A MixedMat argument: the MatArgs of its blocks, row-major over
    ``shape`` (row components, column components)."""
    blocks: Sequence[MatArg]
    shape: tuple
    access: str


def _ctype(dtype):
    try:
        return _CTYPES[np.dtype(dtype)]
    except KeyError:
        raise TypeError(f"unsupported dtype {dtype}") from None


def _generate(code, name, args, nmaps_ids, subset, headers=(), extruded=None):
    """C source of the sequential wrapper (see module docstring).

    ``extruded``: None, or ``(constant_layers, pass_layer, offsets, region,
    periodic, quotients)`` where ``offsets[m]`` is unique map m's per-layer
    offset (or None), ``region`` one of "all", "bottom", "top", "interior",
    and ``quotients[m]`` a periodic map's offset quotients (or None)."""
    offsets = extruded[2] if extruded else {}
    region = extruded[3] if extruded else "all"
    periodic = bool(extruded[4]) if extruded else False
    quots = extruded[5] if extruded else {}

    def step(m, i, f):
        # cells above the column's bottom, wrapped on a periodic column by
        # the generated loop's remainder a < n ? a : a - n (builder.py:27-30,
        # 103-121)
        st = "(layer - bottom)" if f is None else f"(layer - bottom + {f})"
        if not periodic:
            return st
        if quots.get(m) is None:
            return f"op2o_rem({st}, ncell)"
        return f"(op2o_rem({st} + quo{m}[{i}], ncell) - op2o_rem(quo{m}[{i}], ncell))"
    # interior facets (builder.py:81-129, 825, 904): every map entry is taken
    # for the cell below (f = 0) and the cell above (f = 1), f outermost
    nf = 2 if region == "interior" else 1

    def mv(m, ar, i):
        if nf == 1:
            base = f"map{m}[(int64_t)n * {ar} + {i}]"
            if offsets.get(m) is not None:
                return f"(({base}) >= 0 ? ({base}) + off{m}[{i}] * {step(m, i, None)} : -1)"
            return base
        assert offsets.get(m) is not None, "interior facets need maps with offsets"
        base = f"map{m}[(int64_t)n * {ar} + ({i}) % {ar}]"
        return f"(({base}) >= 0 ? ({base}) + off{m}[({i}) % {ar}] * {step(m, f'({i}) % {ar}', f'({i}) / {ar}')} : -1)"
    lines = ["#include <math.h>", "#include <stdint.h>", "#include <string.h>",
             "#include <stdlib.h>"]
    lines += [h for h in headers]
    lines.append(code)
    lines.append(r"""
static int32_t op2o_rem(int32_t a, int32_t n) { return a < n ? a : a - n; }

static int32_t op2o_slot(const int32_t *rowptr, const int32_t *colidx,
                         int32_t row, int32_t col)
{
  int32_t lo = rowptr[row], hi = rowptr[row + 1] - 1;
  while (lo <= hi) {
    const int32_t mid = lo + ((hi - lo) >> 1);
    const int32_t c = colidx[mid];
    if (c == col) return mid;
    if (c < col) lo = mid + 1; else hi = mid - 1;
  }
  return -1;
}
""")
    params = ["int32_t start", "int32_t end"]
    if subset:
        params.append("const int32_t *subset_indices")
    for k, a in enumerate(args):
        if isinstance(a, MatArg):
            vt = _ctype(a.mat.values.dtype)
            params += [f"const int32_t *rowptr{k}", f"const int32_t *colidx{k}", f"{vt} *arg{k}"]
        elif isinstance(a, MixedMatArg):
            for b, blk in enumerate(a.blocks):
                vt = _ctype(blk.mat.values.dtype)
                params += [f"const int32_t *rowptr{k}_{b}", f"const int32_t *colidx{k}_{b}", f"{vt} *arg{k}_{b}"]
        elif isinstance(a, MixedDatArg):
            params += [f"{_ctype(p.data.dtype)} *arg{k}_{c}" for c, p in enumerate(a.parts)]
        else:
            params.append(f"{_ctype(a.data.dtype)} *arg{k}")
    nmaps = len(set(nmaps_ids.values()))
    for m in range(nmaps):
        params.append(f"const int32_t *map{m}")
    if extruded:
        params.append("const int32_t *layers")
    body = []
    body.append("  int32_t err = 0;")
    body.append("  for (int32_t i_ = start; i_ < end; ++i_) {")
    body.append("    const int32_t n = %s;" % ("subset_indices[i_]" if subset else "i_"))
    if extruded:
        # builder.py:683-800: cells bottom..top-1 of column n, maps offset by
        # the layer above the column's bottom (builder.py:81-129)
        for m, o in sorted(offsets.items()):
            if o is not None:
                body.append(f"    const int32_t off{m}[{len(o)}] = {{{', '.join(str(int(v)) for v in o)}}};")
        for m, o in sorted(quots.items()):
            if o is not None:
                body.append(f"    const int32_t quo{m}[{len(o)}] = {{{', '.join(str(int(v)) for v in o)}}};")
        # builder.py:771-795: the region's layers; a column with no cells
        # runs no top cell
        row = "0" if extruded[0] else "n"
        body.append(f"    const int32_t bottom = layers[2 * {row}], cend = layers[2 * {row} + 1] - 1;")
        body.append("    const int32_t ncell = cend - bottom; (void)ncell;")
        # a periodic column has as many interior facets as cells
        lo, hi = {"all": ("bottom", "cend"), "bottom": ("bottom", "bottom + 1"),
                  "top": ("(cend - 1 > bottom ? cend - 1 : bottom)", "cend"),
                  "interior": ("bottom", "cend" if periodic else "cend - 1")}[region]
        body.append(f"    for (int32_t layer = {lo}; layer < {hi}; ++layer) {{")
    call = []
    unpack = []
    for k, a in enumerate(args):
        if isinstance(a, DatArg):
            t = _ctype(a.data.dtype)
            dim, stride, comp = a.cdim, a.stride, (a.comp or 0)
            if a.map_ is None:
                call.append(f"(void *)(arg{k} + (int64_t)n * {stride} + {comp})")
                continue
            m = nmaps_ids[k]
            bar = a.map_.shape[1]
            ar = nf * bar
            body.append(f"    {t} buf{k}[{ar * dim}];")
            if a.access in (INC, WRITE):
                body.append(f"    for (int t_ = 0; t_ < {ar * dim}; ++t_) buf{k}[t_] = 0;")
            else:
                body.append(f"    for (int i = 0; i < {ar}; ++i) {{ const int32_t m_ = {mv(m, bar, 'i')};"
                            f" for (int j = 0; j < {dim}; ++j) buf{k}[i * {dim} + j] = m_ >= 0 ? arg{k}[(int64_t)m_ * {stride} + {comp} + j] : 0; }}")
            call.append(f"(void *)buf{k}")
            if a.access == READ:
                continue
            op = {INC: "arg{k}[o_] += buf{k}[i * {dim} + j];",
                  WRITE: "arg{k}[o_] = buf{k}[i * {dim} + j];",
                  RW: "arg{k}[o_] = buf{k}[i * {dim} + j];",
                  MIN: "{{ const {t} v_ = buf{k}[i * {dim} + j]; if (v_ < arg{k}[o_]) arg{k}[o_] = v_; }}",
                  MAX: "{{ const {t} v_ = buf{k}[i * {dim} + j]; if (v_ > arg{k}[o_]) arg{k}[o_] = v_; }}"}[a.access]
            op = op.format(k=k, dim=dim, t=t)
            unpack.append(f"    for (int i = 0; i < {ar}; ++i) {{ const int32_t m_ = {mv(m, bar, 'i')};"
                          f" if (m_ < 0) continue; for (int j = 0; j < {dim}; ++j) {{ const int64_t o_ = (int64_t)m_ * {stride} + {comp} + j; {op} }} }}")
        elif isinstance(a, GlobalArg):
            t = _ctype(a.data.dtype)
            dim = int(a.data.size)
            if a.access == READ:
                call.append(f"(void *)arg{k}")
                continue
            body.append(f"    {t} buf{k}[{dim}];")
            if a.access in (INC, WRITE):
                body.append(f"    for (int t_ = 0; t_ < {dim}; ++t_) buf{k}[t_] = 0;")
            else:
                body.append(f"    for (int t_ = 0; t_ < {dim}; ++t_) buf{k}[t_] = arg{k}[t_];")
            call.append(f"(void *)buf{k}")
            op = {INC: "arg{k}[t_] += buf{k}[t_];",
                  WRITE: "arg{k}[t_] = buf{k}[t_];",
                  RW: "arg{k}[t_] = buf{k}[t_];",
                  MIN: "if (buf{k}[t_] < arg{k}[t_]) arg{k}[t_] = buf{k}[t_];",
                  MAX: "if (buf{k}[t_] > arg{k}[t_]) arg{k}[t_] = buf{k}[t_];"}[a.access]
            unpack.append(f"    for (int t_ = 0; t_ < {dim}; ++t_) {{ {op.format(k=k)} }}")
        elif isinstance(a, MixedDatArg):
            t = _ctype(a.parts[0].data.dtype)
            total = sum(nf * p.map_.shape[1] * p.cdim for p in a.parts)
            body.append(f"    {t} buf{k}[{total}];")
            if a.access in (INC, WRITE):
                body.append(f"    for (int t_ = 0; t_ < {total}; ++t_) buf{k}[t_] = 0;")
            off = 0
            for c, p in enumerate(a.parts):
                m, bar, dim = nmaps_ids[(k, c)], p.map_.shape[1], p.cdim
                ar = nf * bar
                if a.access not in (INC, WRITE):
                    body.append(f"    for (int i = 0; i < {ar}; ++i) {{ const int32_t m_ = {mv(m, bar, 'i')};"
                                f" for (int j = 0; j < {dim}; ++j) buf{k}[{off} + i * {dim} + j] ="
                                f" m_ >= 0 ? arg{k}_{c}[(int64_t)m_ * {dim} + j] : 0; }}")
                if a.access != READ:
                    src_ = f"buf{k}[{off} + i * {dim} + j]"
                    op = {INC: f"arg{k}_{c}[o_] += {src_};", WRITE: f"arg{k}_{c}[o_] = {src_};",
                          RW: f"arg{k}_{c}[o_] = {src_};",
                          MIN: f"{{ const {t} v_ = {src_}; if (v_ < arg{k}_{c}[o_]) arg{k}_{c}[o_] = v_; }}",
                          MAX: f"{{ const {t} v_ = {src_}; if (v_ > arg{k}_{c}[o_]) arg{k}_{c}[o_] = v_; }}"}[a.access]
                    unpack.append(f"    for (int i = 0; i < {ar}; ++i) {{ const int32_t m_ = {mv(m, bar, 'i')};"
                                  f" if (m_ < 0) continue; for (int j = 0; j < {dim}; ++j) {{"
                                  f" const int64_t o_ = (int64_t)m_ * {dim} + j; {op} }} }}")
                off += ar * dim
            call.append(f"(void *)buf{k}")
        elif isinstance(a, MixedMatArg):
            nr, nc = a.shape
            t = _ctype(a.blocks[0].mat.values.dtype)
            rds = [a.blocks[i * nc].rdim for i in range(nr)]
            cds = [a.blocks[j].cdim for j in range(nc)]
            rars = [nf * a.blocks[i * nc].rmap.shape[1] for i in range(nr)]
            cars = [nf * a.blocks[j].cmap.shape[1] for j in range(nc)]
            roff = [sum(rars[q] * rds[q] for q in range(i)) for i in range(nr)]
            coff = [sum(cars[q] * cds[q] for q in range(j)) for j in range(nc)]
            NR, NC = roff[-1] + rars[-1] * rds[-1], coff[-1] + cars[-1] * cds[-1]
            body.append(f"    {t} buf{k}[{NR * NC}];")
            body.append(f"    for (int t_ = 0; t_ < {NR * NC}; ++t_) buf{k}[t_] = 0;")
            call.append(f"(void *)buf{k}")
            assign = "+=" if a.access == INC else "="
            for i in range(nr):
                for j in range(nc):
                    b = i * nc + j
                    blk = a.blocks[b]
                    rm, cm = nmaps_ids[(k, i)], nmaps_ids[(k, nr + j)]
                    rar, car, rd, cd = rars[i], cars[j], rds[i], cds[j]
                    val = f"buf{k}[({roff[i]} + ii * {rd} + ri) * {NC} + {coff[j]} + jj * {cd} + cj]"
                    if blk.mat.block:
                        unpack.append(
                            f"    for (int ii = 0; ii < {rar}; ++ii) {{ const int32_t r_ = {mv(rm, rar // nf, 'ii')}; if (r_ < 0) continue;\n"
                            f"      for (int jj = 0; jj < {car}; ++jj) {{ const int32_t c_ = {mv(cm, car // nf, 'jj')}; if (c_ < 0) continue;\n"
                            f"        const int32_t s_ = op2o_slot(rowptr{k}_{b}, colidx{k}_{b}, r_, c_); if (s_ < 0) {{ ++err; continue; }}\n"
                            f"        for (int ri = 0; ri < {rd}; ++ri) for (int cj = 0; cj < {cd}; ++cj)\n"
                            f"          arg{k}_{b}[(int64_t)s_ * {rd * cd} + ri * {cd} + cj] {assign} {val}; }} }}")
                    else:
                        unpack.append(
                            f"    for (int ii = 0; ii < {rar}; ++ii) {{ const int32_t r_ = {mv(rm, rar // nf, 'ii')}; if (r_ < 0) continue;\n"
                            f"      for (int ri = 0; ri < {rd}; ++ri)\n"
                            f"      for (int jj = 0; jj < {car}; ++jj) {{ const int32_t c_ = {mv(cm, car // nf, 'jj')}; if (c_ < 0) continue;\n"
                            f"        for (int cj = 0; cj < {cd}; ++cj) {{\n"
                            f"          const int32_t s_ = op2o_slot(rowptr{k}_{b}, colidx{k}_{b}, r_ * {rd} + ri, c_ * {cd} + cj); if (s_ < 0) {{ ++err; continue; }}\n"
                            f"          arg{k}_{b}[s_] {assign} {val}; }} }} }}")
        else:  # MatArg
            t = _ctype(a.mat.values.dtype)
            rm, cm = nmaps_ids[(k, 0)], nmaps_ids[(k, 1)]
            rar, car = nf * a.rmap.shape[1], nf * a.cmap.shape[1]
            rd, cd = a.rdim, a.cdim
            ncol = car * cd
            body.append(f"    {t} buf{k}[{rar * rd * ncol}];")
            body.append(f"    for (int t_ = 0; t_ < {rar * rd * ncol}; ++t_) buf{k}[t_] = 0;")
            call.append(f"(void *)buf{k}")
            assign = "+=" if a.access == INC else "="
            if a.mat.block:
                unpack.append(
                    f"    for (int i = 0; i < {rar}; ++i) {{ const int32_t r_ = {mv(rm, rar // nf, 'i')}; if (r_ < 0) continue;\n"
                    f"      for (int j = 0; j < {car}; ++j) {{ const int32_t c_ = {mv(cm, car // nf, 'j')}; if (c_ < 0) continue;\n"
                    f"        const int32_t s_ = op2o_slot(rowptr{k}, colidx{k}, r_, c_); if (s_ < 0) {{ ++err; continue; }}\n"
                    f"        for (int ri = 0; ri < {rd}; ++ri) for (int cj = 0; cj < {cd}; ++cj)\n"
                    f"          arg{k}[(int64_t)s_ * {rd * cd} + ri * {cd} + cj] {assign} buf{k}[(i * {rd} + ri) * {ncol} + j * {cd} + cj]; }} }}")
            else:
                unpack.append(
                    f"    for (int i = 0; i < {rar}; ++i) {{ const int32_t r_ = {mv(rm, rar // nf, 'i')}; if (r_ < 0) continue;\n"
                    f"      for (int ri = 0; ri < {rd}; ++ri)\n"
                    f"      for (int j = 0; j < {car}; ++j) {{ const int32_t c_ = {mv(cm, car // nf, 'j')}; if (c_ < 0) continue;\n"
                    f"        for (int cj = 0; cj < {cd}; ++cj) {{\n"
                    f"          const int32_t s_ = op2o_slot(rowptr{k}, colidx{k}, r_ * {rd} + ri, c_ * {cd} + cj); if (s_ < 0) {{ ++err; continue; }}\n"
                    f"          arg{k}[s_] {assign} buf{k}[(i * {rd} + ri) * {ncol} + j * {cd} + cj]; }} }} }}")
    if extruded and extruded[1]:
        call.append("layer")
    body.append(f"    {name}({', '.join(call)});")
    body += unpack
    if extruded:
        body.append("    }")
    body.append("  }")
    body.append("  return err;")
    src = "\n".join(lines) + "\nint op2o_wrap(" + ", ".join(params) + ")\n{\n" + "\n".join(body) + "\n}\n"
    return src


_LOADED = {}


def compile_wrapper(src, flavour="strict"):
    """Compile ``src`` with gcc into ``oracle/_build`` (cached by hash)."""
    key = hashlib.sha256((flavour + "\0" + src).encode()).hexdigest()[:24]
    if key in _LOADED:
        return _LOADED[key]
    os.makedirs(_BUILD_DIR, exist_ok=True)
    so = os.path.join(_BUILD_DIR, f"seq_{key}.so")
    if not os.path.exists(so):
        with tempfile.TemporaryDirectory() as td:
            c = os.path.join(td, "w.c")
            with open(c, "w") as f:
                f.write(src)
            tmp = os.path.join(td, "w.so")
            r = subprocess.run(["gcc", *_FLAGS[flavour], c, "-o", tmp, "-lm"],
                               capture_output=True, text=True)
            if r.returncode != 0:
                raise RuntimeError("oracle wrapper failed to compile:\n" + r.stderr + "\n" + src)
            os.replace(tmp, so)
    lib = ctypes.CDLL(so)
    fn = lib.op2o_wrap
    fn.restype = ctypes.c_int32
    _LOADED[key] = fn
    return fn


def _map_ids(args):
    """Unique maps by identity in first-use order (``global_kernel.py:21-23``)."""
    ids = {}
    seen = []

    def idx(m):
        for q, s in enumerate(seen):
            if s is m:
                return q
        seen.append(m)
        return len(seen) - 1
    for k, a in enumerate(args):
        if isinstance(a, DatArg) and a.map_ is not None:
            ids[k] = idx(a.map_)
        elif isinstance(a, MatArg):
            ids[(k, 0)] = idx(a.rmap)
            ids[(k, 1)] = idx(a.cmap)
        elif isinstance(a, MixedDatArg):
            for c, p in enumerate(a.parts):
                ids[(k, c)] = idx(p.map_)
        elif isinstance(a, MixedMatArg):
            nr, nc = a.shape
            for i in range(nr):
                ids[(k, i)] = idx(a.blocks[i * nc].rmap)
            for j in range(nc):
                ids[(k, nr + j)] = idx(a.blocks[j].cmap)
    return ids, seen


def run(code, name, args: Sequence, parts, subset=None, flavour="strict", headers=(), layers=None,
        offsets=None, pass_layer=False, region="all", periodic=False, quotients=None):
    """Execute the sequential wrapper over each ``(start, end)`` in ``parts``.

    Arrays are modified in place (C-contiguous numpy arrays required).
    Returns the number of Mat insertions that missed the pattern.
    Extruded loops: ``layers`` the (columns, 2) or (1, 2) int32 [bottom, top)
    layers, ``offsets`` (map array, per-layer offsets) pairs (maps matched by
    identity), ``pass_layer`` to pass the layer as the kernel's last argument,
    ``region`` the iteration region ("all", "bottom", "top", "interior"),
    ``periodic`` for a periodic extrusion with ``quotients`` (map array,
    offset quotients) pairs.
    """
    ids, maps = _map_ids(args)
    ext = None
    if layers is not None:
        layers = np.ascontiguousarray(layers, dtype=np.int32)
        offs = {}
        for q, m in enumerate(maps):
            for key, o in (offsets or ()):
                if key is m:
                    offs[q] = o
        quo = {}
        for q, m in enumerate(maps):
            for key, o in (quotients or ()):
                if key is m:
                    quo[q] = o
        ext = (layers.shape[0] == 1, pass_layer, offs, region, periodic, quo)
    src = _generate(code, name, args, ids, subset is not None, headers, ext)
    fn = compile_wrapper(src, flavour)
    cargs = []
    if subset is not None:
        subset = np.ascontiguousarray(subset, dtype=np.int32)
        cargs.append(subset.ctypes.data_as(ctypes.c_void_p))
    for a in args:
        if isinstance(a, MatArg):
            for arr in (a.mat.rowptr, a.mat.colidx, a.mat.values):
                assert arr.flags.c_contiguous
                cargs.append(arr.ctypes.data_as(ctypes.c_void_p))
        elif isinstance(a, MixedMatArg):
            for blk in a.blocks:
                for arr in (blk.mat.rowptr, blk.mat.colidx, blk.mat.values):
                    assert arr.flags.c_contiguous
                    cargs.append(arr.ctypes.data_as(ctypes.c_void_p))
        elif isinstance(a, MixedDatArg):
            for p in a.parts:
                assert p.data.flags.c_contiguous
                cargs.append(p.data.ctypes.data_as(ctypes.c_void_p))
        else:
            assert a.data.flags.c_contiguous
            cargs.append(a.data.ctypes.data_as(ctypes.c_void_p))
    for m in maps:
        assert m.dtype == np.int32 and m.flags.c_contiguous
        cargs.append(m.ctypes.data_as(ctypes.c_void_p))
    if layers is not None:
        cargs.append(layers.ctypes.data_as(ctypes.c_void_p))
    err = 0
    for start, end in parts:
        err += fn(ctypes.c_int32(start), ctypes.c_int32(end), *cargs)
    return err
