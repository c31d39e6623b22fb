"""Execution plans: map transposes, greedy colourings and row batches.

Plans depend on the data (map values), so they are built on first use and
cached on the Map (``Map._cache``), the way the reference caches objects
built on top of Sets and Maps.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from pyop2_amd import device as dev

__all__ = ["transpose", "colour_plan", "row_batches", "tile_plan", "Transpose", "ColourPlan",
           "RowBatches", "TilePlan"]


@dataclass
class Transpose:
    """Row incidences of a map: ``inc_ent[inc_ptr[r]:inc_ptr[r+1]]`` lists
    ``e * arity + i`` for every ``map[e][i] == r`` (e < iter_size), ascending."""
    inc_ptr: object
    inc_ent: object
    total: int


def transpose(m, nrows, iter_size):
    key = ("transpose", int(nrows), int(iter_size))
    tr = m._cache.get(key)
    if tr is not None:
        return tr
    L = dev.lib()
    s = dev.stream()
    need = ctypes.c_size_t()
    dev.check(L.op2hip_transpose_workspace(nrows, ctypes.byref(need)), "transpose_workspace")
    work = dev.workspace(need.value)
    inc_ptr = dev.empty((nrows + 1,), np.int32)
    total = ctypes.c_int64()
    mv = m.device_values
    dev.check(L.op2hip_transpose_count(nrows, dev.ptr(mv), m.arity, iter_size, dev.ptr(inc_ptr),
                                       ctypes.byref(total), dev.ptr(work), work.numel(), s),
              "transpose_count")
    inc_ent = dev.empty((max(total.value, 1),), np.int32)
    dev.check(L.op2hip_transpose_fill(nrows, dev.ptr(mv), m.arity, iter_size, dev.ptr(inc_ptr),
                                      dev.ptr(inc_ent), dev.ptr(work), work.numel(), s),
              "transpose_fill")
    tr = Transpose(inc_ptr, inc_ent, int(total.value))
    m._cache[key] = tr
    return tr


@dataclass
class ColourPlan:
    colours: np.ndarray        # host colour of each element of the range
    elems: object              # device: element ids sorted by colour
    offsets: np.ndarray        # host int32, ncolours+1
    ncolours: int
    _offsets_c: object = None

    def exec_fields(self, ex):
        self._offsets_c = np.ascontiguousarray(self.offsets, dtype=np.int32)
        ex.colour_elems = dev.ptr(self.elems)
        ex.colour_offsets = self._offsets_c.ctypes.data
        ex.ncolours = self.ncolours


def greedy_colours(elements, maps_and_targets):
    """Greedy colouring (``doc/sphinx/source/plan.rst:46-80``) of ``elements``
    (host int32 ids) against the rows of each ``(map, target_key)``."""
    L = dev.lib()
    targets = {}
    rows = []
    for m, tkey in maps_and_targets:
        tid = targets.setdefault(tkey, (len(targets), m.toset.total_size))[0]
        rows.append((np.ascontiguousarray(m._table[elements], dtype=np.int32), tid))
    nmaps = len(rows)
    ptrs = (ctypes.c_void_p * max(nmaps, 1))(*[r.ctypes.data for r, _ in rows])
    ar = np.array([r.shape[1] for r, _ in rows] or [0], dtype=np.int32)
    tid = np.array([t for _, t in rows] or [0], dtype=np.int32)
    tsz = np.zeros(max(len(targets), 1), dtype=np.int32)
    for _, (i, size) in targets.items():
        tsz[i] = size
    out = np.empty(max(len(elements), 1), dtype=np.int32)
    nc = ctypes.c_int32()
    dev.check(L.op2hip_colour_greedy(len(elements), nmaps, ptrs, ar.ctypes.data, tid.ctypes.data,
                                     tsz.ctypes.data, len(targets), out.ctypes.data,
                                     ctypes.byref(nc)), "colour_greedy")
    return out[:len(elements)], int(nc.value)


def colour_plan(owner, elements, maps_and_targets, key):
    """Colour plan of host element ids ``elements``, cached on ``owner``."""
    ck = ("colour", key)
    plan = owner._cache.get(ck)
    if plan is not None:
        return plan
    elements = np.ascontiguousarray(elements, dtype=np.int32)
    colours, nc = greedy_colours(elements, maps_and_targets)
    order = np.argsort(colours, kind="stable")
    counts = np.bincount(colours, minlength=nc)
    offsets = np.zeros(nc + 1, dtype=np.int32)
    np.cumsum(counts, out=offsets[1:])
    plan = ColourPlan(colours, dev.to_device(elements[order]), offsets, nc)
    owner._cache[ck] = plan
    return plan


@dataclass
class RowBatches:
    batch_rows: object   # device int32, nbatches+1
    nbatches: int
    max_inc: int
    max_nnz: int
    max_rows: int
    slots: dict = None   # (cmap id) -> (table, descriptors)
    max_row_nnz: int = None
    cells_plans: dict = None

    def exec_fields(self, ex):
        ex.batch_rows = dev.ptr(self.batch_rows)
        ex.nbatches = self.nbatches
        ex.batch_max_inc = self.max_inc
        ex.batch_max_nnz = self.max_nnz
        ex.batch_max_rows = self.max_rows

    def slot_table(self, tr, rowptr, colidx, cmap, rarity):
        """Slot table of these batches for column map ``cmap`` (built once),
        or None when a row is too long for 8-bit slots or the arity has no
        table layout (the caller then uses another strategy)."""
        if self.slots is None:
            self.slots = {}
        if self.max_row_nnz is None:
            self.max_row_nnz = int(np.diff(dev.to_host(rowptr)).max(initial=0))
        if self.max_row_nnz > 255 or cmap.arity not in (3, 4, 10):
            return None
        hit = self.slots.get(id(cmap))
        if hit is None:
            L = dev.lib()
            eb = ctypes.c_size_t()
            dev.check(L.op2hip_slot_table_bytes(cmap.arity, ctypes.byref(eb)), "slot_table_bytes")
            table = dev.workspace(max(tr.total, 1) * eb.value)
            desc = dev.empty((max(self.nbatches, 1), 8), np.int32)
            bad = dev.zeros((2,), np.int32)
            dev.check(L.op2hip_build_slot_table(self.nbatches, dev.ptr(self.batch_rows), dev.ptr(rowptr),
                                                dev.ptr(colidx), dev.ptr(tr.inc_ptr), dev.ptr(tr.inc_ent),
                                                dev.ptr(cmap.device_values), rarity, cmap.arity,
                                                dev.ptr(table), dev.ptr(desc), dev.ptr(bad), dev.stream()),
                      "build_slot_table")
            nbad, nrep = (int(v) for v in dev.to_host(bad))
            if nbad:
                from pyop2_amd.exceptions import PatternError
                raise PatternError(f"{nbad} map entries are not in the sparsity pattern")
            hit = self.slots[id(cmap)] = (table, desc, cmap, nrep)
        return hit

    def slot_fields(self, ex, hit):
        ex.slot_table = dev.ptr(hit[0])
        ex.batch_desc = dev.ptr(hit[1])

    def gather_table(self, hit, nblocks):
        """Gather table (per-block counts, per-batch lists) of slot table
        ``hit`` for a Mat of ``nblocks`` blocks, or None when a block takes
        more than 255 element blocks (built once)."""
        if len(hit) > 4:
            return hit[4]
        L = dev.lib()
        table, desc, cmap = hit[0], hit[1], hit[2]
        gcount = dev.empty((max(nblocks, 1),), np.uint8)
        totals = dev.empty((max(self.nbatches, 1),), np.int32)
        over = dev.zeros((1,), np.int32)
        dev.check(L.op2hip_gather_table_count(self.nbatches, self.max_nnz, dev.ptr(desc), dev.ptr(table),
                                              cmap.arity, dev.ptr(gcount), dev.ptr(totals), dev.ptr(over),
                                              dev.stream()), "gather_table_count")
        gt = None
        if not int(over.item()):
            tot = dev.to_host(totals)[:self.nbatches].astype(np.int64)
            off = np.zeros(max(self.nbatches, 1), dtype=np.int64)
            np.cumsum(tot[:-1], out=off[1:self.nbatches])
            total = int(tot.sum())
            if total < 2 ** 31:
                glist = dev.empty((max(total, 1),), np.uint16)
                doff = dev.to_device(off.astype(np.int32))
                dev.check(L.op2hip_gather_table_fill(self.nbatches, self.max_nnz, dev.ptr(desc), dev.ptr(table),
                                                     cmap.arity, dev.ptr(doff), dev.ptr(totals),
                                                     dev.ptr(gcount), dev.ptr(glist), dev.stream()),
                          "gather_table_fill")
                gt = (gcount, glist)
        key = id(cmap)
        self.slots[key] = hit = hit + (gt,)
        return gt

    def cell_gather(self, hit, tr, rarity, nblocks):
        """Cell-gather plan (op2hip_plan_cell_gather) of slot table ``hit``,
        built once on the host from copies of the device plan."""
        if self.cells_plans is None:
            self.cells_plans = {}
        cmap = hit[2]
        cg = self.cells_plans.get(id(cmap))
        if cg is not None:
            return cg
        L = dev.lib()
        nb = self.nbatches
        desc = np.ascontiguousarray(dev.to_host(hit[1]))
        table = np.ascontiguousarray(dev.to_host(hit[0]))
        inc_ent = np.ascontiguousarray(dev.to_host(tr.inc_ent), dtype=np.int32)
        ninc = int(tr.total)
        cell_ptr = np.empty(nb + 1, dtype=np.int32)
        cells = np.empty(max(ninc, 1), dtype=np.int32)
        gcount = np.zeros(max(nblocks, 1), dtype=np.uint8)
        order = np.zeros(max(nblocks, 1), dtype=np.uint16)
        contrib_ptr = np.empty(nb + 1, dtype=np.int32)
        contrib = np.empty(max(ninc * cmap.arity, 1), dtype=np.uint16)
        rc = L.op2hip_plan_cell_gather(nb, desc.ctypes.data, table.ctypes.data, rarity, cmap.arity,
                                       inc_ent.ctypes.data, cell_ptr.ctypes.data, cells.ctypes.data,
                                       gcount.ctypes.data, order.ctypes.data, contrib_ptr.ctypes.data,
                                       contrib.ctypes.data)
        if rc:
            cg = None
        else:
            nc, ng = int(cell_ptr[nb]), int(contrib_ptr[nb])
            cg = dict(cell_ptr=dev.to_device(cell_ptr), cells=dev.to_device(cells[:max(nc, 1)]),
                      gcount=dev.to_device(gcount), order=dev.to_device(order),
                      contrib_ptr=dev.to_device(contrib_ptr), contrib=dev.to_device(contrib[:max(ng, 1)]),
                      max_cells=int(np.diff(cell_ptr).max(initial=0)),
                      max_contrib=int(np.diff(contrib_ptr).max(initial=0)))
        self.cells_plans[id(cmap)] = cg
        return cg

    @staticmethod
    def cell_fields(ex, cg):
        ex.cell_ptr = dev.ptr(cg["cell_ptr"])
        ex.cells = dev.ptr(cg["cells"])
        ex.gather_count = dev.ptr(cg["gcount"])
        ex.block_order = dev.ptr(cg["order"])
        ex.contrib_ptr = dev.ptr(cg["contrib_ptr"])
        ex.contrib = dev.ptr(cg["contrib"])
        ex.max_batch_cells = cg["max_cells"]
        ex.max_batch_contrib = cg["max_contrib"]

    @staticmethod
    def gather_fields(ex, gt):
        ex.gather_count = dev.ptr(gt[0])
        ex.gather_list = dev.ptr(gt[1])


def row_batches(owner, tr, rowptr, max_inc, max_nnz, key):
    """Batches of consecutive rows for the row-gather assembly."""
    ck = ("rowbatch", key, int(max_inc), int(max_nnz))
    plan = owner._cache.get(ck)
    if plan is not None:
        return plan
    L = dev.lib()
    ip = np.ascontiguousarray(dev.to_host(tr.inc_ptr), dtype=np.int32)
    rp = np.ascontiguousarray(dev.to_host(rowptr), dtype=np.int32)
    nrows = len(ip) - 1
    br = np.empty(nrows + 1, dtype=np.int32)
    nb, mi, mn, mr = (ctypes.c_int32() for _ in range(4))
    dev.check(L.op2hip_plan_row_batches(nrows, ip.ctypes.data, rp.ctypes.data, int(max_inc),
                                        int(max_nnz), br.ctypes.data, ctypes.byref(nb),
                                        ctypes.byref(mi), ctypes.byref(mn), ctypes.byref(mr)),
              "plan_row_batches")
    plan = RowBatches(dev.to_device(br[:nb.value + 1]), nb.value, mi.value, mn.value, mr.value)
    owner._cache[ck] = plan
    return plan


@dataclass
class TilePlan:
    """Compact tiles of the elements of [start, end) and, per staged map, each
    tile's distinct nodes (exclusive ones first) and the elements' local ids;
    per map, the shared-node reduction plan (partial slot offsets per tile,
    distinct shared nodes, their slots) of ``op2hip_plan_tile_reduce``."""
    ntiles: int
    tile_ptr: object
    tile_elems: object
    node_ptr: list
    nodes: list
    nexcl: list
    local: list
    max_nodes: list
    shared_off: list = None
    nshared: list = None
    shared_nodes: list = None
    shared_ptr: list = None
    shared_slots: list = None
    nslot_total: list = None
    workspaces: dict = field(default_factory=dict)
    host_node_ptr: list = None     # per map, host copies for the tile colouring
    host_nodes: list = None
    colourings: dict = field(default_factory=dict)

    def colouring(self, maps, sizes):
        """Greedy colouring of the tiles (the plan's partition colouring,
        ``doc/sphinx/source/plan.rst:46-51, 78-80``): tiles of one colour share
        no node of the staged maps ``maps`` (indices; ``sizes`` their target
        set sizes), so they can write every node without atomics.  Built once
        per map selection."""
        key = tuple(maps)
        hit = self.colourings.get(key)
        if hit is None:
            hit = self.colourings[key] = colour_tiles(
                self.ntiles, [self.host_node_ptr[m] for m in maps], [self.host_nodes[m] for m in maps],
                list(sizes))
        return hit

    def partials(self, m, nout):
        """Device workspace of map m's partial slots (nout float64 each),
        allocated on first use and kept with the plan."""
        buf = self.workspaces.get((m, nout))
        if buf is None:
            buf = dev.zeros((max(int(self.nslot_total[m]) * nout, 1),), np.float64)
            self.workspaces[(m, nout)] = buf
        return buf

    def exec_fields(self, ex, nouts=None):
        """Fill the tile fields of ``ex``; with ``nouts`` (increments per node
        of each staged map), also the partial-slot plan of the maps that have
        increments."""
        ex.ntiles = self.ntiles
        ex.tile_ptr = dev.ptr(self.tile_ptr)
        ex.tile_elems = dev.ptr(self.tile_elems)
        for m in range(len(self.nodes)):
            ex.tile_node_ptr[m] = dev.ptr(self.node_ptr[m])
            ex.tile_nodes[m] = dev.ptr(self.nodes[m])
            ex.tile_nexcl[m] = dev.ptr(self.nexcl[m])
            ex.elem_local[m] = dev.ptr(self.local[m])
            ex.max_tile_nodes[m] = self.max_nodes[m]
            ex.tile_shared_off[m] = None
            if nouts is not None and nouts[m] > 0 and self.shared_off is not None:
                ex.tile_shared_off[m] = dev.ptr(self.shared_off[m])
                ex.nshared[m] = self.nshared[m]
                ex.shared_nodes[m] = dev.ptr(self.shared_nodes[m])
                ex.shared_ptr[m] = dev.ptr(self.shared_ptr[m])
                ex.shared_slots[m] = dev.ptr(self.shared_slots[m])
                ex.partials[m] = dev.ptr(self.partials(m, nouts[m]))


@dataclass
class TileColouring:
    """Tiles sorted by colour (device) and the colour boundaries (host)."""
    order: object
    offsets: np.ndarray
    ncolours: int

    def exec_fields(self, ex):
        ex.colour_elems = dev.ptr(self.order)
        ex.colour_offsets = self.offsets.ctypes.data
        ex.ncolours = self.ncolours


def tile_node_table(ntiles, node_ptr, nodes):
    """The tiles' node lists as a fixed-arity table (rows padded with -1)."""
    counts = np.diff(node_ptr[:ntiles + 1]).astype(np.int64)
    width = int(counts.max(initial=0))
    table = np.full((max(ntiles, 1), max(width, 1)), -1, dtype=np.int32)
    rows = np.repeat(np.arange(ntiles, dtype=np.int64), counts)
    cols = np.arange(int(counts.sum()), dtype=np.int64) - np.repeat(node_ptr[:ntiles].astype(np.int64), counts)
    table[rows, cols] = nodes[:int(counts.sum())]
    return table


def tile_colours(ntiles, node_ptrs, nodes, sizes):
    """Greedy colouring of tiles over their node lists (one list per map):
    the same first-free-colour bitmask colouring as the elements'
    (``op2hip_colour_greedy``), with each tile's nodes as its map row.
    Returns (host colours int32[ntiles], number of colours)."""
    L = dev.lib()
    tables = [tile_node_table(ntiles, p, n) for p, n in zip(node_ptrs, nodes)]
    nm = len(tables)
    ptrs = (ctypes.c_void_p * max(nm, 1))(*[t.ctypes.data for t in tables])
    ar = np.array([t.shape[1] for t in tables] or [0], dtype=np.int32)
    tid = np.arange(max(nm, 1), dtype=np.int32)
    tsz = np.array(sizes or [0], dtype=np.int32)
    out = np.empty(max(ntiles, 1), dtype=np.int32)
    nc = ctypes.c_int32()
    dev.check(L.op2hip_colour_greedy(int(ntiles), nm, ptrs, ar.ctypes.data, tid.ctypes.data,
                                     tsz.ctypes.data, nm, out.ctypes.data, ctypes.byref(nc)),
              "colour_greedy (tiles)")
    return out[:ntiles], int(nc.value)


def colour_tiles(ntiles, node_ptrs, nodes, sizes):
    """Tiles sorted by their greedy colour (``tile_colours``), on the device."""
    colours, nc = tile_colours(ntiles, node_ptrs, nodes, sizes)
    order = np.argsort(colours, kind="stable").astype(np.int32)
    counts = np.bincount(colours, minlength=nc)
    offsets = np.zeros(nc + 1, dtype=np.int32)
    np.cumsum(counts, out=offsets[1:])
    return TileColouring(dev.to_device(order), offsets, nc)


def tile_reduce_plan(L, ntiles, node_ptr, nodes, nexcl, nnodes):
    """Host arrays of ``op2hip_plan_tile_reduce`` for one staged map:
    (shared_off, nshared, shared_nodes, shared_ptr, shared_slots)."""
    cnt = np.diff(node_ptr[:ntiles + 1]).astype(np.int64) - nexcl[:ntiles]
    nslots = int(cnt.sum()) if ntiles else 0
    off = np.empty(ntiles + 1, dtype=np.int32)
    sn = np.empty(max(nslots, 1), dtype=np.int32)
    sp = np.empty(nslots + 1, dtype=np.int32)
    ss = np.empty(max(nslots, 1), dtype=np.int32)
    ns = ctypes.c_int32()
    dev.check(L.op2hip_plan_tile_reduce(int(ntiles), node_ptr.ctypes.data, nodes.ctypes.data,
                                        nexcl.ctypes.data, int(nnodes), off.ctypes.data,
                                        ctypes.byref(ns), sn.ctypes.data, sp.ctypes.data,
                                        ss.ctypes.data), "plan_tile_reduce")
    ns = ns.value
    return off, ns, sn[:max(ns, 1)], sp[:ns + 1], ss, nslots


def tile_plan(maps, start, end, tile_size, node_cap=0):
    """Tile plan of the elements [start, end) over ``maps`` (staged map order)
    with at most ``tile_size`` elements and ``node_cap`` nodes of map 0 (0: no
    cap) per tile, cached on the first map."""
    owner = maps[0]
    ck = ("tiles", tuple(id(m) for m in maps), int(start), int(end), int(tile_size), int(node_cap))
    plan = owner._cache.get(ck)
    if plan is not None:
        return plan
    L = dev.lib()
    n = int(end - start)
    nm = len(maps)
    vals = [np.ascontiguousarray(m._table, dtype=np.int32) for m in maps]
    ar = np.array([m.arity for m in maps], dtype=np.int32)
    ts = np.array([m.toset.total_size for m in maps], dtype=np.int32)
    tp = np.empty(n + 1, dtype=np.int32)
    te = np.empty(max(n, 1), dtype=np.int32)
    nps = [np.empty(n + 1, dtype=np.int32) for _ in maps]
    nds = [np.empty(max(n * m.arity, 1), dtype=np.int32) for m in maps]
    nxs = [np.empty(max(n, 1), dtype=np.int32) for _ in maps]
    lcs = [np.empty(max(n * m.arity, 1), dtype=np.uint16) for m in maps]
    mx = np.zeros(2, dtype=np.int32)
    nt = ctypes.c_int32()

    def arr(a):
        return (ctypes.c_void_p * nm)(*[x.ctypes.data for x in a])
    dev.check(L.op2hip_plan_tiles(int(start), int(end), nm, arr(vals), ar.ctypes.data, ts.ctypes.data,
                                  int(tile_size), int(node_cap), ctypes.byref(nt), tp.ctypes.data,
                                  te.ctypes.data, arr(nps), arr(nds), arr(nxs), arr(lcs), mx.ctypes.data),
              "plan_tiles")
    nt = nt.value
    red = [tile_reduce_plan(L, nt, nps[k], nds[k], nxs[k], ts[k]) for k in range(nm)]
    plan = TilePlan(nt, dev.to_device(tp[:nt + 1]), dev.to_device(te),
                    [dev.to_device(x[:nt + 1]) for x in nps],
                    [dev.to_device(nds[k][:max(int(nps[k][nt]), 1)]) for k in range(nm)],
                    [dev.to_device(x[:max(nt, 1)]) for x in nxs],
                    [dev.to_device(x) for x in lcs], [int(v) for v in mx[:nm]],
                    [dev.to_device(r[0]) for r in red], [r[1] for r in red],
                    [dev.to_device(r[2]) for r in red], [dev.to_device(r[3]) for r in red],
                    [dev.to_device(r[4]) for r in red], [r[5] for r in red],
                    host_node_ptr=[x[:nt + 1] for x in nps],
                    host_nodes=[nds[k][:int(nps[k][nt])] for k in range(nm)])
    owner._cache[ck] = plan
    return plan


@dataclass
class RowList:
    """Batches of rows in element-local order (op2hip_plan_row_list)."""
    nbatches: int
    desc: object        # device, nbatches x 8 int32
    row_order: object   # device, nrows x 2 int32
    ent: object         # device, incidences in batch order
    table: object       # device, slot entries in batch order
    max_inc: int
    max_nnz: int
    max_rows: int
    row_base: object = None   # device, rowptr[row] per row_order entry
    packed: object = None     # device, one word per incidence (arity-3 maps), or None

    def exec_fields(self, ex):
        ex.nbatches = self.nbatches
        ex.batch_desc = dev.ptr(self.desc)
        ex.row_order = dev.ptr(self.row_order)
        ex.batch_ent = dev.ptr(self.ent)
        ex.row_base = dev.ptr(self.row_base) if self.row_base is not None else None
        ex.batch_packed = dev.ptr(self.packed) if self.packed is not None else None
        ex.slot_table = dev.ptr(self.table)
        ex.batch_max_inc = self.max_inc
        ex.batch_max_nnz = self.max_nnz
        ex.batch_max_rows = self.max_rows


def pack_row_list(ent, table, rarity, carity):
    """The row list's per-incidence words for arity-3 maps (None otherwise):
    bits 0-31 the incidence, 32-47 the row's block offset in its batch,
    48-62 the three column slots, 5 bits each (31: not in the pattern).
    ``table`` holds the slot entries as bytes: row offset (u16), batch row
    (u16), one u8 slot per column (255: missing)."""
    if rarity != 3 or carity != 3 or table.shape[1] < 7 or len(ent) == 0:
        return None
    t = table.astype(np.uint64)
    slots = t[:, 4:7]
    if ((slots >= 31) & (slots != 255)).any():
        return None
    slots = np.where(slots == 255, np.uint64(31), slots)
    word = ent.astype(np.int64).astype(np.uint64) & np.uint64(0xFFFFFFFF)
    word |= (t[:, 0] | (t[:, 1] << np.uint64(8))) << np.uint64(32)
    for j in range(3):
        word |= slots[:, j] << np.uint64(48 + 5 * j)
    return np.ascontiguousarray(word)


def row_list(rmap, cmap, tr, mat, max_inc, max_nnz, key):
    """Row-list plan of ``mat``'s pattern and the maps, cached on ``rmap``."""
    ck = ("rowlist", key, id(cmap), int(max_inc), int(max_nnz))
    plan = rmap._cache.get(ck)
    if plan is not None:
        return plan
    L = dev.lib()
    nrows = int(mat.sparsity.nrows)
    rowptr = np.ascontiguousarray(dev.to_host(mat.rowptr), dtype=np.int32)
    colidx = np.ascontiguousarray(dev.to_host(mat.colidx), dtype=np.int32)
    inc_ptr = np.ascontiguousarray(dev.to_host(tr.inc_ptr), dtype=np.int32)
    inc_ent = np.ascontiguousarray(dev.to_host(tr.inc_ent), dtype=np.int32)
    cvals = np.ascontiguousarray(cmap._table, dtype=np.int32)
    eb = ctypes.c_size_t()
    dev.check(L.op2hip_slot_table_bytes(cmap.arity, ctypes.byref(eb)), "slot_table_bytes")
    ninc = int(tr.total)
    order = np.empty(max(2 * nrows, 2), dtype=np.int32)
    desc = np.zeros((max(nrows, 1), 8), dtype=np.int32)
    ent = np.empty(max(ninc, 1), dtype=np.int32)
    table = np.zeros(max(ninc, 1) * eb.value, dtype=np.uint8)
    nb, mi, mn, mr = (ctypes.c_int32() for _ in range(4))
    missing = ctypes.c_int64()
    dev.check(L.op2hip_plan_row_list(nrows, rowptr.ctypes.data, colidx.ctypes.data, inc_ptr.ctypes.data,
                                     inc_ent.ctypes.data, cvals.ctypes.data, rmap.arity, cmap.arity,
                                     int(max_inc), int(max_nnz), order.ctypes.data, ctypes.byref(nb),
                                     desc.ctypes.data, ent.ctypes.data, table.ctypes.data,
                                     ctypes.byref(missing), ctypes.byref(mi), ctypes.byref(mn),
                                     ctypes.byref(mr)), "plan_row_list")
    if missing.value:
        from pyop2_amd.exceptions import PatternError
        raise PatternError(f"{missing.value} map entries are not in the sparsity pattern")
    base = rowptr[order[0:2 * nrows:2]] if nrows else np.zeros(1, dtype=np.int32)
    packed = pack_row_list(ent[:ninc], table.reshape(max(ninc, 1), -1)[:ninc], rmap.arity, cmap.arity)
    plan = RowList(nb.value, dev.to_device(desc[:max(nb.value, 1)]), dev.to_device(order),
                   dev.to_device(ent), dev.to_device(table), mi.value, mn.value, mr.value,
                   dev.to_device(np.ascontiguousarray(base, dtype=np.int32)),
                   dev.to_device(packed.view(np.int64)) if packed is not None else None)
    rmap._cache[ck] = plan
    return plan
