"""Registry of hand-written HIP element loops.

``pyop2_amd.elements`` registers the digest of each element kernel's C
source with a family name.  ``lookup(global_kernel)`` returns an executor
for a GlobalKernel whose local kernel is registered and whose argument
signature (access modes, dims, arities, dtypes, map sharing) is the one the
HIP entry point implements; anything else goes to the generated path.

Conflict resolution (``configuration["strategy"]``):
* ``atomic``   -- device-scope atomics (FP64 atomics are native on gfx950);
* ``coloured`` -- greedy colouring, one launch per colour, plain adds;
* ``rowgather``-- Mat arguments: owner-computes rows, no atomics;
* ``auto``     -- the autotuner's choice (``pyop2_amd.autotune``).
"""
import ctypes

import numpy as np

from pyop2_amd import device as dev
from pyop2_amd.configuration import configuration
from pyop2_amd.elements import code_digest
from pyop2_amd.global_kernel import DatKernelArg, MatKernelArg
from pyop2_amd.types.access import INC, READ
from pyop2_amd.types.set import Subset

__all__ = ["register_element_kernel", "lookup", "FAMILIES"]

_REGISTERED = {}


def register_element_kernel(code, name, family, params):
    _REGISTERED[code_digest(code, name)] = (family, dict(params))


# family -> (argument specs, map sharing, entry points)
# spec: ("dat", access, dim, arity) | ("mat", access, (rdim, cdim), (rarity, carity))
FAMILIES = {
    "p1_lumped_mass": dict(
        args=[("dat", INC, (1,), 3), ("dat", READ, (2,), 3)],
        maps=[[(0, 0), (1, 0)]], entry={"f64": "op2hip_wrap_p1_lumped_mass"},
        staged=[0], staged_vals=[(2, 1)]),
    "p1_laplace_residual": dict(
        args=[("dat", INC, (1,), 3), ("dat", READ, (2,), 3), ("dat", READ, (1,), 3)],
        maps=[[(0, 0), (1, 0), (2, 0)]], entry={"f64": "op2hip_wrap_p1_laplace_residual"},
        staged=[0], staged_vals=[(3, 1)]),
    "p1_laplace_matrix": dict(
        args=[("mat", INC, (1, 1), (3, 3)), ("dat", READ, (2,), 3)],
        maps=[[(0, 0), (0, 1), (1, 0)]], entry={"f64": "op2hip_wrap_p1_laplace_matrix"}),
    "p2_elasticity_residual": dict(
        args=[("dat", INC, (3,), 10), ("dat", READ, (3,), 4), ("dat", READ, (3,), 10)],
        maps=[[(0, 0), (2, 0)]],
        entry={"f64": "op2hip_wrap_p2_elasticity_residual",
               "f32": "op2hip_wrap_p2_elasticity_residual_f32"},
        staged=[0, 1], staged_vals=[(3, 3), (3, 0)]),
    "p2_elasticity_matrix": dict(
        args=[("mat", INC, (3, 3), (10, 10)), ("dat", READ, (3,), 4)],
        maps=[[(0, 0), (0, 1)]],
        entry={"f64": "op2hip_wrap_p2_elasticity_matrix",
               "f32": "op2hip_wrap_p2_elasticity_matrix_f32"}),
    "p2_elasticity_fused": dict(
        args=[("mat", INC, (3, 3), (10, 10)), ("dat", INC, (3,), 10), ("dat", READ, (3,), 4),
              ("dat", READ, (3,), 10)],
        maps=[[(0, 0), (0, 1), (1, 0), (3, 0)]],
        entry={"f64": "op2hip_wrap_p2_elasticity_fused",
               "f32": "op2hip_wrap_p2_elasticity_fused_f32"},
        # "split": the same results as the assembly loop (args 0, 2) and the
        # residual loop (args 1, 2, 3) launched one after the other
        split=[("p2_elasticity_matrix", [0, 2]), ("p2_elasticity_residual", [1, 2, 3])]),
    "p2_elasticity_adjoint": dict(
        args=[("dat", INC, (3,), 10), ("dat", INC, (3,), 4), ("dat", READ, (3,), 10),
              ("dat", READ, (3,), 4), ("dat", READ, (3,), 10)],
        maps=[[(0, 0), (2, 0), (4, 0)], [(1, 0), (3, 0)]],
        entry={"f64": "op2hip_wrap_p2_elasticity_adjoint"},
        staged=[0, 1], staged_vals=[(6, 3), (3, 3)]),
}


def _register_known_sources():
    """Register the element kernels' sources at their default parameters
    (the Lame parameters 1 and 0.5 of the P2 kernels), so a loop whose
    kernel text is one of them -- up to formatting, comments and its name --
    takes the hand-written entry point however the kernel was built."""
    from pyop2_amd.elements import kernel_source
    for family in FAMILIES:
        params = {} if family.startswith("p1_") else {"lam": 1.0, "mu": 0.5}
        register_element_kernel(kernel_source(family, **params), family, family, params)


_register_known_sources()

# "staged": arguments whose maps are the staged maps 0 and 1 of the entry
# point's OP2HIP_STAGED launch (tiles of _TILE elements); "staged_vals":
# per staged map, the values staged (Mat dtype) and accumulated (double) per
# node, which size the launch's LDS.
_TILE = 256
_LDS = 64 * 1024

_side_streams = {}


def _side_stream(device):
    """The side stream (default priority) split loops overlap parts on."""
    s = _side_streams.get(device)
    if s is None:
        s = _side_streams[device] = dev.torch().cuda.Stream(device=device)
    return s


def _tile_lds(spec, itemsize, max_nodes):
    b = 0
    for (nin, nout), mx in zip(spec["staged_vals"], max_nodes):
        b = (b + mx * nin * itemsize + 7) // 8 * 8 + mx * nout * 8
    return b

# Row-batch budgets (max row incidences, max blocks) of the row-gather
# assembly (LDS <= 64 KiB per workgroup; float64 P2 batches of at most 560
# blocks take 40,320 bytes, so four workgroups share a CU).
_BATCH = {("p1_laplace_matrix", "f64"): (1024, 2048),
          ("p2_elasticity_matrix", "f64"): (448, 560),
          ("p2_elasticity_matrix", "f32"): (512, 640),
          ("p2_elasticity_fused", "f64"): (448, 560),
          ("p2_elasticity_fused", "f32"): (512, 640)}
# The staged variant holds one batch's element block rows in LDS: at most
# 256 / (column groups of the element) incidences per batch.
# The cell-gather variant stages the geometry of at most 256 elements (one
# per row incidence at most), then the batch's new block values, in the same
# LDS (256 P2 elements or 512 3x3 blocks both take 36 KiB).
_BATCH_CELLS = {("p1_laplace_matrix", "f64"): (256, 512),
                ("p2_elasticity_matrix", "f64"): (256, 512),
                ("p2_elasticity_matrix", "f32"): (256, 512),
                ("p2_elasticity_fused", "f64"): (256, 512),
                ("p2_elasticity_fused", "f32"): (256, 512)}
_BATCH_STAGED = {("p1_laplace_matrix", "f64"): (256, 512),
                 ("p2_elasticity_matrix", "f64"): (64, 256),
                 ("p2_elasticity_matrix", "f32"): (64, 256),
                 ("p2_elasticity_fused", "f64"): (64, 256),
                 ("p2_elasticity_fused", "f32"): (64, 256)}


def _variant(dtypes):
    dts = {np.dtype(d) for d in dtypes}
    if dts == {np.dtype(np.float64)}:
        return "f64"
    if dts == {np.dtype(np.float32)}:
        return "f32"
    return None


def _refuse_generated(family, variant, why):
    """The generated loop hands the kernel buffers of the data's type, and the
    registered kernel sources take double arguments: their float32 variants
    exist only as hand-written loops."""
    from pyop2_amd.exceptions import DataTypeError
    raise DataTypeError(f"{family}: {variant or 'mixed-precision'} data runs only in the hand-written "
                        f"loops, which this loop does not fit ({why}); the kernel's C source takes "
                        "double arguments")


def lookup(gk):
    """Executor for ``gk`` if a hand-written entry point implements it."""
    lk = gk.local_kernel
    hit = _REGISTERED.get(code_digest(lk.code, lk.name))
    if hit is None:
        return None
    family, params = hit
    ex = _match(gk, lk, family, params)
    if ex is None:
        variant = _variant(lk.dtypes)
        if variant != "f64":
            _refuse_generated(family, variant, "extruded set or other arguments than the entry point's")
    return ex


def _match(gk, lk, family, params):
    if getattr(gk, "_extruded", False):
        return None
    spec = FAMILIES[family]
    if len(gk.arguments) != len(spec["args"]):
        return None
    for (kind, access, dims, arity), larg, garg in zip(spec["args"], lk.arguments, gk.arguments):
        if larg.access != access:
            return None
        if kind == "dat":
            if not isinstance(garg, DatKernelArg) or garg.map_ is None:
                return None
            if tuple(garg.dim) != dims or garg.map_.arity != arity or garg.index is not None:
                return None
        else:
            if not isinstance(garg, MatKernelArg) or garg.unroll:
                return None
            if tuple(garg.dims[0][0]) != dims or tuple(m.arity for m in garg.maps) != arity:
                return None
    variant = _variant(lk.dtypes)
    if variant not in spec["entry"]:
        return None
    return HandWrittenKernel(gk, family, params, variant)


class _SubLoop:
    """The part of a Parloop a split strategy hands to one of its loops."""

    def __init__(self, iterset, arguments):
        self.iterset = iterset
        self.arguments = arguments


def fits_32bit_indexing(parloop):
    """True if every element and node index product of the loop (element x
    map arity, node x 3 components) stays below 2**31, the range of the
    hand-written loops' int32 arithmetic."""
    limit = 2 ** 31 - 1
    for a in parloop.arguments:
        for m in getattr(a, "maps", ()):
            if m.iterset.total_size * m.arity > limit or m.toset.total_size * 3 > limit:
                return False
    return True


class HandWrittenKernel:
    """Launches a registered element loop through its C entry point."""

    def __init__(self, gk, family, params, variant):
        self.gk = gk
        self.family = family
        self.params = params
        self.variant = variant
        self.spec = FAMILIES[family]
        self.fn = getattr(dev.lib(), self.spec["entry"][variant])
        self.strategy_used = None

    def _check_maps(self, parloop):
        """Arguments that the entry point reads through one map must share it,
        and the loop must fit the hand-written loops' 32-bit indexing."""
        if not fits_32bit_indexing(parloop):
            return False   # the generated loops index in 64 bits
        maps = [a.maps if hasattr(a, "maps") else () for a in parloop.arguments]
        if any(m.has_negative for ms in maps for m in ms):
            return False   # the hand-written loops assume every entry has a target
        for group in self.spec["maps"]:
            objs = {id(maps[i][j]) for i, j in group}
            if len(objs) != 1:
                return False
        return True

    def applicable(self, parloop, strategy):
        """``strategy`` if this loop can run it, else the fallback it takes."""
        has_mat = any(s[0] == "mat" for s in self.spec["args"])
        subset = isinstance(parloop.iterset, Subset)
        if strategy.startswith("rowgather") and (not has_mat or subset):
            return "atomic"
        if strategy in ("staged", "staged_coloured") and (subset or "staged" not in self.spec):
            return "atomic"
        if strategy == "split" and "split" not in self.spec:
            return "atomic"
        if has_mat and strategy.startswith("rowgather"):
            mat = parloop.arguments[0].data
            if mat.block_shape != self.spec["args"][0][2]:
                return "atomic"
        return strategy

    def _choose(self, parloop):
        """The launch configuration: ``{"strategy": ..., parameters}`` (the
        autotuner's choice under ``auto``)."""
        want = configuration["strategy"]
        if want == "auto":
            from pyop2_amd import autotune
            cfg = dict(autotune.choice(self, parloop))
        else:
            cfg = {"strategy": want}
        cfg["strategy"] = self.applicable(parloop, cfg["strategy"])
        return cfg

    def _tile_plan(self, args, start, end, tile=_TILE, lds_budget=_LDS):
        """Tile plan of ``tile`` elements whose LDS fits ``lds_budget``: the
        node cap of map 0 shrinks until it does (None if even small tiles do
        not fit)."""
        from pyop2_amd import plans
        maps = [args[k].maps[0] for k in self.spec["staged"]]
        isz = 8 if self.variant == "f64" else 4
        nin, nout = self.spec["staged_vals"][0]
        cap = 0
        for _ in range(8):
            tp = plans.tile_plan(maps, start, end, tile, cap)
            if _tile_lds(self.spec, isz, tp.max_nodes) <= lds_budget:
                return tp
            per0 = nin * isz + nout * 8
            cap = max(maps[0].arity, int(0.9 * min(tp.max_nodes[0], lds_budget // per0)))
        return None

    def _inc_maps(self, parloop):
        out = []
        for (kind, access, _, _), arg in zip(self.spec["args"], parloop.arguments):
            if access == INC:
                out.append(arg.maps[0])
        return out

    def __call__(self, start, end, ptrs, parloop):
        if parloop is None:
            raise RuntimeError("hand-written loops are launched by a Parloop")
        if not self._check_maps(parloop):
            # a loop the entry point does not take: the generated path
            if self.variant != "f64":
                _refuse_generated(self.family, self.variant,
                                  "a map with negative entries, or indices beyond 32 bits")
            from pyop2_amd.codegen import GeneratedKernel
            return GeneratedKernel(self.gk)(start, end, ptrs, parloop)
        self.launch(start, end, parloop, self._choose(parloop))

    @property
    def last_config(self):
        """The configuration of the last launch (strategy and parameters)."""
        return getattr(self, "_last_config", None)

    def _split_parts(self):
        parts = getattr(self, "_parts", None)
        if parts is None:
            parts = self._parts = [(HandWrittenKernel(self.gk, fam, self.params, self.variant), idx)
                                   for fam, idx in self.spec["split"]]
        return parts

    def _launch_split(self, start, end, parloop, outputs, cfg=None):
        """The loop as its split parts: the assembly with its best row gather,
        the residual with LDS-staged tiles (one element pass each, no
        redundant gathers of the solution per row incidence); ``cfg["parts"]``
        holds each part's launch configuration.

        The parts write disjoint outputs and only read what they share, so
        ``cfg["overlap"]`` may run the parts that write no matrix on a side
        stream (default priority), forked from and joined back into the
        calling stream: ``"before"`` issues them ahead of the assembly,
        ``"after"`` behind it, and the residual's kernels run beside the
        assembly's instead of after them."""
        cfg = cfg or {}
        parts_cfg = cfg.get("parts") or [None] * len(self.spec["split"])
        jobs = []
        for (sub, idx), pcfg in zip(self._split_parts(), parts_cfg):
            part = _SubLoop(parloop.iterset, [parloop.arguments[k] for k in idx])
            sub_out = None
            if outputs:
                sub_out = {}
                for n, k in enumerate(idx):
                    if k in outputs:
                        sub_out[n] = outputs[k]
                if "mat" in outputs:
                    sub_out["mat"] = outputs["mat"]
            if pcfg is None:
                pcfg = {"strategy": "rowgather_slots" if sub.spec["args"][0][0] == "mat" else "staged"}
            pcfg = dict(pcfg, strategy=sub.applicable(part, pcfg["strategy"]))
            jobs.append((sub, part, pcfg, sub_out))
        self.strategy_used = "split"
        overlap = cfg.get("overlap")
        if overlap not in ("before", "after"):
            for sub, part, pcfg, sub_out in jobs:
                sub.launch(start, end, part, pcfg, sub_out)
            return
        torch = dev.torch()
        main = torch.cuda.current_stream(dev.device())
        side = _side_stream(dev.device())
        side.wait_stream(main)
        on_side = [j for j in jobs if j[0].spec["args"][0][0] != "mat"]
        on_main = [j for j in jobs if j[0].spec["args"][0][0] == "mat"]

        def issue_side():
            with torch.cuda.stream(side):
                for sub, part, pcfg, sub_out in on_side:
                    sub.launch(start, end, part, pcfg, sub_out)

        try:
            if overlap == "before":
                issue_side()
            for sub, part, pcfg, sub_out in on_main:
                sub.launch(start, end, part, pcfg, sub_out)
            if overlap == "after":
                issue_side()
        finally:
            # the calling stream always rejoins what the side stream has
            main.wait_stream(side)

    def launch(self, start, end, parloop, cfg, outputs=None):
        """Launch with the configuration ``cfg`` (a strategy name, or a dict
        of the strategy and its parameters: ``tile`` / ``lds`` / ``block``
        for staged tiles, ``batch`` (max incidences, max blocks) / ``block``
        for row gathers, ``parts`` for a split loop); ``outputs`` optionally
        replaces the written arguments' device pointers (autotuning on scratch
        copies)."""
        if isinstance(cfg, str):
            cfg = {"strategy": cfg}
        strategy = cfg["strategy"]
        self._last_config = cfg
        if strategy == "split":
            with dev.timed_launch(self.family):
                return self._launch_split(start, end, parloop, outputs, cfg)
        ex = dev.ExecPlan()
        ex.strategy = dev.STRATEGIES[strategy]
        ex.block_size = int(cfg.get("block") or 0)
        keep = []
        iterset = parloop.iterset
        if isinstance(iterset, Subset):
            ex.subset = dev.ptr(iterset.device_indices)
        args = parloop.arguments
        desc = None
        if strategy == "coloured":
            from pyop2_amd import plans
            if isinstance(iterset, Subset):
                elems = iterset.indices[start:end]
            else:
                elems = np.arange(start, end, dtype=np.int32)
            owner = self._inc_maps(parloop)[0]
            key = (iterset._plan_key(), start, end, tuple(id(m) for m in self._inc_maps(parloop)))
            plan = plans.colour_plan(owner, elems,
                                     [(m, id(m.toset)) for m in self._inc_maps(parloop)], key)
            plan.exec_fields(ex)
            keep.append(plan)
        if strategy in ("staged", "staged_coloured"):
            tp = self._tile_plan(args, start, end, int(cfg.get("tile") or _TILE), int(cfg.get("lds") or _LDS))
            if tp is None:
                strategy = "atomic"
                ex.strategy = dev.STRATEGIES[strategy]
            elif strategy == "staged_coloured":
                # tiles coloured over the nodes of the maps they increment:
                # one launch per colour, every node written without atomics
                inc = [m for m, (_, nout) in enumerate(self.spec["staged_vals"]) if nout > 0]
                sizes = [args[self.spec["staged"][m]].maps[0].toset.total_size for m in inc]
                col = tp.colouring(inc, sizes)
                tp.exec_fields(ex, None)
                col.exec_fields(ex)
                keep += [tp, col]
            else:
                nouts = [nout for _, nout in self.spec["staged_vals"]]
                partials = cfg.get("partials", configuration["staged_partials"])
                tp.exec_fields(ex, nouts if partials else None)
                keep.append(tp)
        if self.spec["args"][0][0] == "mat":
            mat = args[0].data
            rmap = args[0].maps[0]
            if strategy == "rowgather_rows":
                from pyop2_amd import plans
                tr = plans.transpose(rmap, mat.sparsity.nrows, rmap.iterset.size)
                budget = tuple(cfg.get("batch") or _BATCH[(self.family, self.variant)])
                rl = plans.row_list(rmap, args[0].maps[1], tr, mat, budget[0], budget[1],
                                    (id(mat.sparsity), mat.block_shape, self.variant))
                rl.exec_fields(ex)
                keep.append(rl)
                desc = mat.descriptor(tr.inc_ptr, tr.inc_ent)
            elif strategy.startswith("rowgather"):
                from pyop2_amd import plans
                tr = plans.transpose(rmap, mat.sparsity.nrows, rmap.iterset.size)
                key = (id(mat.sparsity), mat.block_shape, self.variant)
                rb = None
                if strategy in ("rowgather_staged", "rowgather_cells"):
                    table = _BATCH_STAGED if strategy == "rowgather_staged" else _BATCH_CELLS
                    budget = table[(self.family, self.variant)]
                    rb = plans.row_batches(rmap, tr, mat.rowptr, budget[0], budget[1], key)
                    if rb.max_inc > budget[0]:
                        # a row with more incidences than one batch can stage
                        rb, strategy = None, "rowgather_slots"
                if rb is None:
                    budget = tuple(cfg.get("batch") or _BATCH[(self.family, self.variant)])
                    rb = plans.row_batches(rmap, tr, mat.rowptr, budget[0], budget[1], key)
                rb.exec_fields(ex)
                keep.append(rb)
                if strategy in ("rowgather_slots", "rowgather_staged", "rowgather_cells"):
                    hit = rb.slot_table(tr, mat.rowptr, mat.colidx, args[0].maps[1], rmap.arity)
                    if hit is None:
                        strategy = "rowgather_thread"
                    elif strategy == "rowgather_cells":
                        cg = rb.cell_gather(hit, tr, rmap.arity, mat.sparsity.block_nnz)
                        if cg is None:
                            strategy = "rowgather_slots"
                        else:
                            rb.cell_fields(ex, cg)
                            keep.append(cg)
                        rb.slot_fields(ex, hit)
                    elif strategy == "rowgather_staged" and hit[3]:
                        # elements repeating a column need one add per occurrence
                        strategy = "rowgather_thread"
                    else:
                        rb.slot_fields(ex, hit)
                        if strategy == "rowgather_staged":
                            gt = rb.gather_table(hit, mat.sparsity.block_nnz)
                            if gt is None:
                                strategy = "rowgather_thread"
                            else:
                                rb.gather_fields(ex, gt)
                                keep.append(gt)
                ex.strategy = dev.STRATEGIES[strategy]
                desc = mat.descriptor(tr.inc_ptr, tr.inc_ent)
            else:
                desc = mat.descriptor()
            if outputs and "mat" in outputs:
                desc.values = dev.ptr(outputs["mat"])
        cargs = []
        for k, ((kind, access, _, _), arg) in enumerate(zip(self.spec["args"], args)):
            if kind == "mat":
                cargs.append(ctypes.addressof(desc))
            else:
                t = outputs.get(k) if outputs else None
                cargs.append(dev.ptr(t if t is not None else arg.data._device_array()))
        seen = []
        for a in args:
            for m in a.maps:
                if not any(m is s for s in seen):
                    seen.append(m)
        cargs += [dev.ptr(m.device_values) for m in seen]
        if self.family.startswith("p2_"):
            cargs += [ctypes.c_double(self.params["lam"]), ctypes.c_double(self.params["mu"])]
        with dev.timed_launch(self.family):
            rc = self.fn(start, end, *cargs, ctypes.byref(ex), dev.stream())
        dev.check(rc, self.spec["entry"][self.variant])
        self.strategy_used = strategy
        del keep
