"""Sparsity pattern closed form and greedy colouring -- TEST INFRASTRUCTURE ONLY.

``pattern`` restates ``pyop2/sparsity.pyx`` for non-extruded, non-mixed maps:
``build_sparsity`` (``:121-175``) fills a PETSc preallocator at block level
through ``fill_with_zeros`` (``:178-284``): the diagonal block (i, i) for
``i < min(nrows, ncols)`` when the row and column DataSets are the same and
the block is diagonal (``:214-219``; ``types/mat.py:72``), then the dense
block ``rmap[e, :] x cmap[e, :]`` for every owned iteration element
``e < iterset.size`` (``:263-284``).  The resulting pattern is the sorted
unique union of those pairs; negative map entries are ignored (PETSc skips
negative indices).  ``nnz`` is per block row, unwound to dof rows when the
matrix is not block-sparse (``:168-173``).
"""
import ctypes
import os

import numpy as np

__all__ = ["pattern", "dof_nnz", "expand_block_pattern", "monolithic_pattern", "greedy_colour"]


def pattern(nrows, ncols, pairs, set_diag=True):
    """Block CSR pattern ``(rowptr, colidx)`` (int32) of the sparsity.

    ``pairs`` is a list of ``(rmap, cmap, iter_size)``.
    """
    keys = []
    for rmap, cmap, size in pairs:
        r = np.asarray(rmap[:size], dtype=np.int64)
        c = np.asarray(cmap[:size], dtype=np.int64)
        if r.size == 0:
            continue
        rr = np.repeat(r[:, :, None], c.shape[1], axis=2)
        cc = np.repeat(c[:, None, :], r.shape[1], axis=1)
        ok = (rr >= 0) & (cc >= 0)
        keys.append(rr[ok] * ncols + cc[ok])
    if set_diag:
        i = np.arange(min(nrows, ncols), dtype=np.int64)
        keys.append(i * ncols + i)
    if keys:
        k = np.unique(np.concatenate(keys))
    else:
        k = np.zeros(0, dtype=np.int64)
    rows = k // max(ncols, 1)
    cols = k - rows * ncols
    counts = np.bincount(rows, minlength=nrows)
    rowptr = np.zeros(nrows + 1, dtype=np.int64)
    np.cumsum(counts, out=rowptr[1:])
    return rowptr.astype(np.int32), cols.astype(np.int32)


def dof_nnz(rowptr, rdim, cdim, block_sparse=True):
    """Per-row nnz as ``Sparsity.nnz`` reports it (``sparsity.pyx:168-173``)."""
    nnz = np.diff(rowptr).astype(np.int32)
    if not (block_sparse and rdim == cdim):
        nnz = np.repeat(nnz * cdim, rdim)
    return nnz


def expand_block_pattern(rowptr, colidx, rdim, cdim):
    """Scalar (dof) CSR pattern of a block pattern with rdim x cdim blocks."""
    nb = len(rowptr) - 1
    rows = []
    cols = []
    for r in range(nb):
        bc = colidx[rowptr[r]:rowptr[r + 1]].astype(np.int64)
        sc = (bc[:, None] * cdim + np.arange(cdim)[None, :]).ravel()
        for _ in range(rdim):
            rows.append(len(sc))
            cols.append(sc)
    rp = np.zeros(nb * rdim + 1, dtype=np.int64)
    np.cumsum(np.array(rows, dtype=np.int64), out=rp[1:])
    ci = np.concatenate(cols) if cols else np.zeros(0, dtype=np.int64)
    return rp.astype(np.int32), ci.astype(np.int32)


def monolithic_pattern(blocks, rdims, cdims, rsizes, csizes):
    """Scalar pattern of a monolithic mixed matrix (``sparsity.pyx:141-160``:
    one dof-level preallocator, each block (i, j) filled through the local
    submatrix of rows of component i and columns of component j with
    rdims[i] x cdims[j] blocks).  ``blocks[i][j]`` is the block-level
    ``(rowptr, colidx)`` of block (i, j) (diagonal included where the
    reference sets it), or None for an empty block; component i's dofs start
    at the sum of rsizes[q] * rdims[q] over q < i (``dataset.py:166-180``),
    likewise for the columns.  Row R of the result is, column component by
    column component, the unwound row of each block shifted by that
    component's offset."""
    nr, nc = len(rdims), len(cdims)
    coff = np.concatenate([[0], np.cumsum([csizes[j] * cdims[j] for j in range(nc)])]).astype(np.int64)
    rows = []
    for i in range(nr):
        expanded = []
        for j in range(nc):
            if blocks[i][j] is None:
                rp = np.zeros(rsizes[i] + 1, dtype=np.int32)
                ci = np.zeros(0, dtype=np.int32)
            else:
                rp, ci = blocks[i][j]
            erp, eci = expand_block_pattern(rp, ci, rdims[i], cdims[j])
            expanded.append((erp.astype(np.int64), eci.astype(np.int64) + coff[j]))
        for r in range(rsizes[i] * rdims[i]):
            rows.append(np.concatenate([eci[erp[r]:erp[r + 1]] for erp, eci in expanded]))
    rowptr = np.zeros(len(rows) + 1, dtype=np.int64)
    np.cumsum([len(r) for r in rows], out=rowptr[1:])
    colidx = np.concatenate(rows) if rows else np.zeros(0, dtype=np.int64)
    return rowptr.astype(np.int32), colidx.astype(np.int32)


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        here = os.path.dirname(os.path.abspath(__file__))
        so = os.path.join(here, "_build", "liboracle.so")
        if not os.path.exists(so):
            import subprocess
            subprocess.run(["make", "-s", "-C", here], check=True)
        _LIB = ctypes.CDLL(so)
        _LIB.op2o_greedy_colour.restype = ctypes.c_int32
    return _LIB


def greedy_colour(nelem, maps, target_ids, target_sizes):
    """Greedy colouring of ``nelem`` elements (``doc/.../plan.rst:46-80``).

    ``maps`` are (>= nelem, arity) int32 arrays; ``target_ids[m]`` names the
    target set of map m (maps into one set share colour masks).
    Returns ``(colours int32[nelem], ncolours)``.
    """
    maps = [np.ascontiguousarray(m, dtype=np.int32) for m in maps]
    nm = len(maps)
    ptrs = (ctypes.c_void_p * max(nm, 1))(*[m.ctypes.data for m in maps])
    ar = np.array([m.shape[1] for m in maps] or [0], dtype=np.int32)
    tid = np.array(list(target_ids) or [0], dtype=np.int32)
    tsz = np.array(list(target_sizes) or [0], dtype=np.int32)
    ntargets = len(target_sizes)
    out = np.empty(max(nelem, 1), dtype=np.int32)
    nc = _lib().op2o_greedy_colour(
        ctypes.c_int32(nelem), ctypes.c_int32(nm), ptrs,
        ar.ctypes.data_as(ctypes.c_void_p), tid.ctypes.data_as(ctypes.c_void_p),
        tsz.ctypes.data_as(ctypes.c_void_p), ctypes.c_int32(ntargets),
        out.ctypes.data_as(ctypes.c_void_p))
    return out[:nelem].copy(), int(nc)
