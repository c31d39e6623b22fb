"""Many independent meshes' loops as one persistent launch in a HIP graph.

The reference runs one ``par_loop`` per mesh; for many small meshes the GPU
then spends its time in launch overhead and half-empty grids.  A
:class:`BatchedLaplace` takes, per mesh, the pair of loops

    op2.ParLoop(p1_laplace_matrix, cells, A(op2.INC, (c2v, c2v)), X(op2.READ, c2v))
    op2.ParLoop(p1_laplace_residual, cells, R(op2.INC, c2v), X(op2.READ, c2v), u(op2.READ, c2v))

built exactly as for separate execution, plans every mesh once (row batches
and slot table of its CSR assembly, compact tiles of its residual), and runs
all work items of all meshes -- (mesh, row batch) and (mesh, tile) -- in one
launch of a persistent kernel whose workgroups stride over the item list.
The launch is captured in a HIP graph, so a step is one graph replay.  The
grid size (resident workgroups per CU) is chosen by timing, once per batch
shape, and kept with the other autotuning results.
"""
import ctypes

import numpy as np

from pyop2_amd import device as dev
from pyop2_amd.configuration import configuration
from pyop2_amd.exceptions import DeviceError
from pyop2_amd.types.mat import Mat

__all__ = ["BatchedLaplace", "BatchMesh"]

BATCH_TILE = 0x40000000
_ROW_BUDGET = (1024, 2048)   # (max row incidences, max nnz) of a row batch
# elements per residual tile: the batched tiles add the nodes they share with
# device atomics, so larger tiles (fewer shared nodes, more work per item)
# pay until their LDS limits the workgroups a CU holds.  512 meshes of
# 16^2..128^2 squares: 346 / 291 / 273 / 257 / 262 / 278 us with tiles of
# 256 / 512 / 768 / 1024 / 1280 / 1536 elements.
_TILE = 1024


class BatchMesh(ctypes.Structure):
    """``op2hip_batch_mesh``."""
    _fields_ = [
        ("A", dev.MatDesc),
        ("batch_desc", ctypes.c_void_p),
        ("slot_table", ctypes.c_void_p),
        ("nbatches", ctypes.c_int32),
        ("ncells", ctypes.c_int32),
        ("X", ctypes.c_void_p),
        ("c2v", ctypes.c_void_p),
        ("u", ctypes.c_void_p),
        ("R", ctypes.c_void_p),
        ("ntiles", ctypes.c_int32),
        ("max_tile_nodes", ctypes.c_int32),
        ("tile_ptr", ctypes.c_void_p),
        ("tile_elems", ctypes.c_void_p),
        ("tile_node_ptr", ctypes.c_void_p),
        ("tile_nodes", ctypes.c_void_p),
        ("tile_nexcl", ctypes.c_void_p),
        ("elem_local", ctypes.c_void_p),
    ]


def _family(pl):
    ex = pl.global_kernel.compile()
    return getattr(ex, "family", None)


class BatchedLaplace:
    """Laplace CSR assembly + residual INC of many P1 meshes in one launch.

    ``pairs``: one ``(matrix_loop, residual_loop)`` per mesh (see the module
    docstring).  ``graph``: capture the launch in a HIP graph (replayed by
    each call).  ``grid``: workgroups of the persistent launch (None: tuned).
    """

    def __init__(self, pairs, graph=True, grid=None):
        from pyop2_amd import plans
        self._keep = []
        self.meshes = []
        descs = []
        items = []
        max_nnz = max_nodes = 0
        for k, (pm, pr) in enumerate(pairs):
            if _family(pm) != "p1_laplace_matrix" or _family(pr) != "p1_laplace_residual":
                raise TypeError(f"mesh {k}: expected (p1_laplace_matrix, p1_laplace_residual) loops")
            A = pm.arguments[0].data
            c2v = pm.arguments[0].maps[0]
            X = pm.arguments[1].data
            R, u = pr.arguments[0].data, pr.arguments[2].data
            if not (pm.arguments[0].maps[1] is c2v and pm.arguments[1].map_ is c2v
                    and all(a.map_ is c2v for a in pr.arguments) and pr.arguments[1].data is X):
                raise ValueError(f"mesh {k}: both loops must use one cell->vertex Map and one X")
            if pm.iterset is not pr.iterset or pm.iterset.total_size != pm.iterset.size:
                raise ValueError(f"mesh {k}: both loops iterate one Set without a halo")
            if A.block_shape != (1, 1) or A.dtype != np.float64:
                raise TypeError(f"mesh {k}: the Mat must be float64 CSR")
            ncells = pm.iterset.size
            tr = plans.transpose(c2v, A.sparsity.nrows, ncells)
            rb = plans.row_batches(c2v, tr, A.rowptr, _ROW_BUDGET[0], _ROW_BUDGET[1],
                                   (id(A.sparsity), A.block_shape, "f64"))
            hit = rb.slot_table(tr, A.rowptr, A.colidx, c2v, 3)
            if hit is None:
                raise ValueError(f"mesh {k}: a row has more than 255 entries")
            tp = plans.tile_plan([c2v], 0, ncells, _TILE)
            d = BatchMesh()
            d.A = A.descriptor(tr.inc_ptr, tr.inc_ent)
            d.batch_desc, d.slot_table = dev.ptr(hit[1]), dev.ptr(hit[0])
            d.nbatches, d.ncells = rb.nbatches, ncells
            d.X, d.c2v = dev.ptr(X._device_array()), dev.ptr(c2v.device_values)
            d.u, d.R = dev.ptr(u._device_array()), dev.ptr(R._device_array())
            d.ntiles, d.max_tile_nodes = tp.ntiles, tp.max_nodes[0]
            d.tile_ptr, d.tile_elems = dev.ptr(tp.tile_ptr), dev.ptr(tp.tile_elems)
            d.tile_node_ptr, d.tile_nodes = dev.ptr(tp.node_ptr[0]), dev.ptr(tp.nodes[0])
            d.tile_nexcl, d.elem_local = dev.ptr(tp.nexcl[0]), dev.ptr(tp.local[0])
            descs.append(d)
            items += [(k, b) for b in range(rb.nbatches)]
            items += [(k, BATCH_TILE | t) for t in range(tp.ntiles)]
            max_nnz = max(max_nnz, rb.max_nnz)
            max_nodes = max(max_nodes, tp.max_nodes[0])
            self._keep += [tr, rb, hit, tp]
            self.meshes.append((A, R, X, u))
        raw = (BatchMesh * max(len(descs), 1))(*descs)
        host = np.frombuffer(bytes(raw), dtype=np.uint8)
        self._descs = dev.to_device(host)
        self._items = dev.to_device(np.asarray(items, dtype=np.int32).reshape(-1, 2))
        self.nitems = len(items)
        self.max_nnz, self.max_nodes = max_nnz, max_nodes
        self.ncells = sum(p[0].iterset.size for p in pairs)
        self.grid = grid if grid is not None else self._tune()
        self._graph = None
        if graph:
            self._capture()

    def _launch(self, grid=None):
        L = dev.lib()
        g = self.grid if grid is None else grid
        with dev.timed_launch("batched_p1_laplace"):
            dev.check(L.op2hip_batched_p1_laplace(dev.ptr(self._descs), dev.ptr(self._items),
                                                  self.nitems, g, self.max_nnz, self.max_nodes,
                                                  dev.stream()), "batched_p1_laplace")

    def _tune(self):
        """Resident workgroups per CU of the persistent launch, by timing (the
        result is kept in the autotuning cache under the batch's shape)."""
        from pyop2_amd import autotune
        torch = dev.torch()
        cus = torch.cuda.get_device_properties(dev.device()).multi_processor_count
        key = ["batched_p1_laplace", self.nitems, self.max_nnz, self.max_nodes]
        cands = [cus * w for w in (2, 4, 8)] + [0]
        best = autotune.cached(key)
        if best is not None or not configuration["autotune"]:
            return int(best) if best is not None else cus * 4
        saved = [(A.device_values.clone(), R._device_array().clone()) for A, R, _, _ in self.meshes]
        times = {}
        for g in cands:
            self._launch(g)
            torch.cuda.synchronize()
            ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
            ev[0].record()
            for _ in range(3):
                self._launch(g)
            ev[1].record()
            torch.cuda.synchronize()
            times[g] = ev[0].elapsed_time(ev[1])
        for (A, R, _, _), (a, r) in zip(self.meshes, saved):
            A.device_values.copy_(a)
            R._device_array().copy_(r)
        best = min(times, key=times.get)
        autotune.store(key, best)
        return best

    def _capture(self):
        torch = dev.torch()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                L = dev.lib()
                dev.check(L.op2hip_batched_p1_laplace(dev.ptr(self._descs), dev.ptr(self._items),
                                                      self.nitems, self.grid, self.max_nnz,
                                                      self.max_nodes, side.cuda_stream),
                          "batched_p1_laplace (capture)")
        torch.cuda.current_stream().wait_stream(side)
        self._graph = g

    def __call__(self):
        """One step: every mesh's assembly and residual increment.  The batch
        holds the device storage of its Dats and Mats (host-side changes are
        copied there first, as a par_loop does)."""
        for A, R, X, u in self.meshes:
            X._device_array()
            u._device_array()
            R._device_array()
            A.change_assembly_state(Mat.ADD_VALUES)
        if self._graph is not None:
            with dev.timed_launch("batched_p1_laplace"):
                self._graph.replay()
        else:
            self._launch()
        for A, R, _, _ in self.meshes:
            R.halo_valid = False
            R._device_modified()
            R.increment_dat_version()
