"""The C-ABI library on the CPU: it loads, exports every entry point declared
in ``include/op2hip.h``, and its host-side routines (greedy colouring, row
batching) agree with the oracle; generated wrappers compile for gfx950."""
import ctypes
import os
import re

import numpy as np
import pytest

from helpers import repeated_edge_nodes
from oracle import sparsity as osp
from pyop2_amd import device as dev
from pyop2_amd.meshes import unit_square_tri, unit_cube_tet_p2, workload_rng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "op2hip.h")).read()
    return sorted(set(re.findall(r"\b(op2hip_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = dev.lib()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(L, n), n
    assert L.op2hip_abi_version() == 2


def test_binding_covers_every_declared_symbol():
    assert set(declared_symbols()) <= set(dev._SIGNATURES)


def _lib_colour(nelem, maps, tids, tsizes):
    L = dev.lib()
    maps = [np.ascontiguousarray(m, dtype=np.int32) for m in maps]
    ptrs = (ctypes.c_void_p * len(maps))(*[m.ctypes.data for m in maps])
    ar = np.array([m.shape[1] for m in maps], dtype=np.int32)
    tid = np.array(tids, dtype=np.int32)
    tsz = np.array(tsizes, dtype=np.int32)
    out = np.empty(nelem, dtype=np.int32)
    nc = ctypes.c_int32()
    rc = L.op2hip_colour_greedy(nelem, len(maps), ptrs, ar.ctypes.data, tid.ctypes.data, tsz.ctypes.data,
                                len(tsizes), out.ctypes.data, ctypes.byref(nc))
    assert rc == 0
    return out, nc.value


@pytest.mark.parametrize("what", ["tri", "tet", "clique"])
def test_library_colouring_bit_exact_with_oracle(what):
    if what == "tri":
        m = unit_square_tri(48, rng=workload_rng(2), permute=True)
        args = (m.ncells, [m.cell_vertex], [0], [m.nverts])
    elif what == "tet":
        m = unit_cube_tet_p2(4, rng=workload_rng(3))
        args = (m.ncells, [m.cell_node, m.cell_vertex], [0, 1], [m.nnodes, m.nverts])
    else:
        args = (70, [np.zeros((70, 1), dtype=np.int32)], [0], [1])
    a, na = _lib_colour(*args)
    b, nb = osp.greedy_colour(*args)
    assert na == nb and np.array_equal(a, b)


def test_row_batches_cover_rows_within_budget():
    L = dev.lib()
    rng = np.random.default_rng(0)
    inc = rng.integers(0, 30, 5000).astype(np.int32)
    nz = rng.integers(1, 70, 5000).astype(np.int32)
    ip = np.concatenate([[0], np.cumsum(inc)]).astype(np.int32)
    rp = np.concatenate([[0], np.cumsum(nz)]).astype(np.int32)
    br = np.empty(5001, dtype=np.int32)
    nb, mi, mn, mr = (ctypes.c_int32() for _ in range(4))
    rc = L.op2hip_plan_row_batches(5000, ip.ctypes.data, rp.ctypes.data, 256, 300, br.ctypes.data,
                                   ctypes.byref(nb), ctypes.byref(mi), ctypes.byref(mn), ctypes.byref(mr))
    assert rc == 0
    b = br[:nb.value + 1]
    assert b[0] == 0 and b[-1] == 5000 and np.all(np.diff(b) > 0)
    for lo, hi in zip(b[:-1], b[1:]):
        if hi - lo > 1:
            assert ip[hi] - ip[lo] <= 256 and rp[hi] - rp[lo] <= 300


def _tiles(maps, sizes, start, end, tile, cap=0):
    L = dev.lib()
    n = end - start
    nm = len(maps)
    tp, te = np.empty(n + 1, np.int32), np.empty(max(n, 1), np.int32)
    nps = [np.empty(n + 1, np.int32) for _ in maps]
    nds = [np.empty(max(n * m.shape[1], 1), np.int32) for m in maps]
    nxs = [np.empty(max(n, 1), np.int32) for _ in maps]
    lcs = [np.empty(max(n * m.shape[1], 1), np.uint16) for m in maps]
    mx = np.zeros(2, np.int32)
    ar = np.array([m.shape[1] for m in maps], np.int32)
    ts = np.array(sizes, np.int32)
    nt = ctypes.c_int32()

    def arr(a):
        return (ctypes.c_void_p * nm)(*[x.ctypes.data for x in a])
    rc = L.op2hip_plan_tiles(start, end, nm, arr(maps), ar.ctypes.data, ts.ctypes.data, tile, cap,
                             ctypes.byref(nt), tp.ctypes.data, te.ctypes.data, arr(nps), arr(nds),
                             arr(nxs), arr(lcs), mx.ctypes.data)
    assert rc == 0, dev.lib().op2hip_last_error()
    nt = nt.value
    return nt, tp[:nt + 1], te[:n], nps, nds, nxs, lcs, mx


@pytest.mark.parametrize("start,end,tile,cap,repeated", [(0, 3072, 256, 0, False), (100, 3000, 128, 0, False),
                                                        (0, 37, 64, 0, False), (0, 3072, 256, 300, False),
                                                        (5, 2000, 256, 40, False), (0, 3072, 256, 0, True),
                                                        (100, 3000, 128, 0, True)])
def test_tiles_partition_elements_and_renumber_nodes(start, end, tile, cap, repeated):
    """Every element of [start, end) in exactly one tile of at most ``tile``;
    local ids give back the map entries; a tile's first nexcl nodes occur in
    no other tile, the others in at least two; node lists strictly increase
    within each part.  Also with map rows that name a node twice (both local
    ids then give the one tile node)."""
    m = unit_cube_tet_p2(8, rng=workload_rng(3))
    if repeated:
        m = repeated_edge_nodes(m)
        assert (m.cell_node[:, 8] == m.cell_node[:, 9]).any()
    maps = [np.ascontiguousarray(m.cell_node, np.int32), np.ascontiguousarray(m.cell_vertex, np.int32)]
    nt, tp, te, nps, nds, nxs, lcs, mx = _tiles(maps, [m.nnodes, m.nverts], start, end, tile, cap)
    if cap:
        assert mx[0] <= cap
    assert tp[0] == 0 and tp[-1] == end - start and np.all(np.diff(tp) <= tile) and np.all(np.diff(tp) > 0)
    assert np.array_equal(np.sort(te), np.arange(start, end))
    tile_of = np.repeat(np.arange(nt), np.diff(tp))
    for k, mp in enumerate(maps):
        loc = lcs[k][:(end - start) * mp.shape[1]].reshape(end - start, -1).astype(np.int64)
        assert np.array_equal(nds[k][nps[k][tile_of][:, None] + loc], mp[te])
        owners = {}
        for t in range(nt):
            ids = nds[k][nps[k][t]:nps[k][t + 1]]
            assert len(ids) <= mx[k]
            ex, sh = ids[:nxs[k][t]], ids[nxs[k][t]:]
            assert np.all(np.diff(ex) > 0) and np.all(np.diff(sh) > 0)
            for v in ids:
                owners.setdefault(int(v), []).append((t, bool(len(ex) and v in set(ex.tolist()))))
        for v, occ in owners.items():
            assert all(flag == (len(occ) == 1) for _, flag in occ)


@pytest.mark.parametrize("start,end,tile", [(0, 3072, 256), (7, 1500, 96)])
def test_tile_reduce_plan_lists_shared_slots_in_tile_order(start, end, tile):
    """Shared-node reduction plan: tile b's shared nodes (positions nexcl..)
    get consecutive slots from shared_off[b]; each shared node lists exactly
    the slots of its occurrences, in tile order; nodes in increasing order.
    Checked against a dictionary built from the tile lists."""
    from pyop2_amd import plans
    m = unit_cube_tet_p2(8, rng=workload_rng(3))
    c2n = np.ascontiguousarray(m.cell_node, np.int32)
    nt, tp, te, nps, nds, nxs, lcs, mx = _tiles([c2n], [m.nnodes], start, end, tile)
    off, ns, sn, sp, ss, nslots = plans.tile_reduce_plan(dev.lib(), nt, nps[0], nds[0], nxs[0], m.nnodes)
    want = {}
    slot = 0
    for b in range(nt):
        assert off[b] == slot
        ids = nds[0][nps[0][b]:nps[0][b + 1]]
        for v in ids[nxs[0][b]:]:
            want.setdefault(int(v), []).append(slot)
            slot += 1
    assert off[nt] == slot == nslots and ns == len(want)
    assert list(sn[:ns]) == sorted(want)
    for s_, v in enumerate(sn[:ns]):
        assert list(ss[sp[s_]:sp[s_ + 1]]) == want[int(v)]
        assert len(want[int(v)]) >= 2            # a shared node is in two tiles at least


def test_tiles_are_compact_regions():
    """On a structured mesh the grown regions touch far fewer nodes than
    runs of consecutive elements (a strip of the numbering)."""
    m = unit_cube_tet_p2(24, rng=workload_rng(3))
    c2n = np.ascontiguousarray(m.cell_node, np.int32)
    nt, tp, te, nps, nds, nxs, lcs, mx = _tiles([c2n], [m.nnodes], 0, m.ncells, 256)
    grown = np.diff(nps[0])[:-1].mean()
    strips = np.mean([len(np.unique(c2n[a:a + 256])) for a in range(0, m.ncells - 256, 256)])
    assert grown < 0.8 * strips


def test_row_list_plan_covers_rows_and_finds_slots():
    """Row-list plan on a permuted P1 mesh: every row once; batches within
    budget; rows ordered by first element; each incidence's slots point at
    the entries of its element's columns in its row."""
    m = unit_square_tri(24, rng=workload_rng(2), permute=True)
    cv = np.ascontiguousarray(m.cell_vertex, np.int32)
    rp, ci = osp.pattern(m.nverts, m.nverts, [(cv, cv, m.ncells)])
    rp, ci = rp.astype(np.int32), ci.astype(np.int32)
    ent = np.arange(cv.size, dtype=np.int32)
    rows = cv.ravel()
    order = np.lexsort((ent, rows))
    inc_ent = ent[order].astype(np.int32)
    inc_ptr = np.zeros(m.nverts + 1, np.int32)
    np.cumsum(np.bincount(rows, minlength=m.nverts), out=inc_ptr[1:])
    L = dev.lib()
    n = m.nverts
    ro = np.empty(2 * n, np.int32)
    desc = np.zeros((n, 8), np.int32)
    bent = np.empty(cv.size, np.int32)
    tab = np.zeros((cv.size, 8), np.uint8)
    nb, mi, mn, mr = (ctypes.c_int32() for _ in range(4))
    miss = ctypes.c_int64()
    rc = L.op2hip_plan_row_list(n, rp.ctypes.data, ci.ctypes.data, inc_ptr.ctypes.data, inc_ent.ctypes.data,
                                cv.ctypes.data, 3, 3, 40, 64, ro.ctypes.data, ctypes.byref(nb), desc.ctypes.data,
                                bent.ctypes.data, tab.ctypes.data, ctypes.byref(miss), ctypes.byref(mi),
                                ctypes.byref(mn), ctypes.byref(mr))
    assert rc == 0 and miss.value == 0
    order_rows = ro[0::2]
    assert np.array_equal(np.sort(order_rows), np.arange(n))
    first = np.array([inc_ent[inc_ptr[r]] // 3 for r in order_rows])
    assert np.all(np.diff(first) >= 0)
    for b in range(nb.value):
        r0, nr, _, nz, i0, ni = desc[b, :6]
        rows_b = order_rows[r0:r0 + nr]
        lens = rp[rows_b + 1] - rp[rows_b]
        assert nz == lens.sum() and np.array_equal(ro[1::2][r0:r0 + nr], np.concatenate([[0], np.cumsum(lens)[:-1]]))
        if nr > 1:
            assert nz <= 64 and ni <= 40
        for t in range(i0, i0 + ni):
            row_off, lr = tab[t, 0] | (tab[t, 1].astype(int) << 8), tab[t, 2] | (tab[t, 3].astype(int) << 8)
            r = rows_b[lr]
            e = bent[t] // 3
            assert r == cv[e, bent[t] % 3]
            for j in range(3):
                assert ci[rp[r] + tab[t, 4 + j]] == cv[e, j]


def test_ctypes_structs_match_the_c_layout(tmp_path):
    """Sizes and field offsets of op2hip_exec, op2hip_mat and
    op2hip_batch_mesh as the C compiler lays them out equal the ctypes
    mirrors (a mismatch would hand the kernels wrong pointers)."""
    import shutil
    import subprocess
    from pyop2_amd.batch import BatchMesh
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = {"op2hip_exec": dev.ExecPlan, "op2hip_mat": dev.MatDesc, "op2hip_batch_mesh": BatchMesh}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "op2hip.h"', "int main(void) {"]
    for cname, ct in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in ct._fields_:
            lines.append(f'  printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, ct in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(ct), cname
        for f in ct._fields_:
            assert got[(cname, f[0])] == getattr(ct, f[0]).offset, (cname, f[0])


def test_generated_wrappers_compile_for_gfx950():
    """Every known-answer kernel's generated wrapper compiles with hiprtc."""
    import kat_cases  # noqa: F401  (the kernels are built through pyop2_amd.op2 below)
    from pyop2_amd import op2
    from pyop2_amd.codegen import GeneratedKernel, compile_options
    from pyop2_amd.parloop import LegacyParloop
    it = op2.Set(8)
    ind = op2.Set(8)
    m = op2.Map(it, ind, 2, np.zeros((8, 2), dtype=np.int32))
    x = op2.Dat(ind ** 2, dtype=np.float64)
    y = op2.Dat(it, dtype=np.int32)
    g = op2.Global(1, 0.0, np.float64)
    h = op2.Global(2, [0, 0], np.int64)
    sp = op2.Sparsity((ind ** 1, ind ** 1), [(m, m, None)])
    A = op2.Mat(sp)
    cases = [
        ("static void k1(double x[2][2], int *y, double *g) { x[0][0] += y[0]; g[0] += 1.0; }", "k1",
         (x(op2.INC, m), y(op2.READ), g(op2.INC))),
        ("static void k2(double *x, int *y, long long *h) { y[0] = (int)x[0]; h[0] = h[0] > 3 ? h[0] : 3; }", "k2",
         (x(op2.READ, m), y(op2.WRITE), h(op2.MAX))),
        ("static void k3(double A[2][2], const double *x) { A[0][1] = x[0]; }", "k3",
         (A(op2.INC, (m, m)), x(op2.READ, m))),
        ("static void k4(double *x, double *g) { x[0] = fmin(x[0], g[0]); }", "k4",
         (x(op2.MIN, m), g(op2.READ))),
    ]
    L = dev.lib()
    opts = compile_options()
    copts = (ctypes.c_char_p * len(opts))(*[o.encode() for o in opts])
    for code, name, args in cases:
        pl = LegacyParloop(op2.Kernel(code, name), it, *args)
        src = GeneratedKernel(pl.global_kernel).source(pl)
        size = ctypes.c_size_t(0)
        log = ctypes.create_string_buffer(16384)
        rc = L.op2hip_jit_code_object(src.encode(), copts, len(opts), None, ctypes.byref(size), log, len(log))
        assert rc == 0, log.value.decode()
        assert size.value > 1000


def test_generated_extruded_region_wrappers_compile_for_gfx950():
    """Generated wrappers of extruded loops over each iteration region
    (doubled packs for interior facets, Dat, Mat and Global arguments, the
    layer argument), plain and periodic, compile with hiprtc."""
    from pyop2_amd import op2
    from pyop2_amd.codegen import GeneratedKernel, compile_options
    from pyop2_amd.parloop import LegacyParloop
    base = op2.Set(4)
    ext = op2.ExtrudedSet(base, np.array([[0, 3], [1, 5], [2, 3], [0, 6]], dtype=np.int32))
    cells = op2.Set(40)
    m = op2.Map(ext, cells, 2, np.array([[0, 1], [10, 11], [20, 21], [30, 31]], dtype=np.int32),
                offset=np.full(2, 2, dtype=np.int32))
    x = op2.Dat(cells ** 1, dtype=np.float64)
    g = op2.Global(1, 0.0, np.float64)
    sp = op2.Sparsity((cells ** 1, cells ** 1), [(m, m, (op2.ON_INTERIOR_FACETS, op2.ON_TOP))])
    A = op2.Mat(sp)
    cases = [
        ("static void f1(double x[4], double *g) { x[0] += x[3]; g[0] += 1.0; }", "f1",
         (x(op2.INC, m), g(op2.INC)), op2.ON_INTERIOR_FACETS, False),
        ("static void f2(double A[4][4], const double *x) { A[0][3] = x[2]; }", "f2",
         (A(op2.INC, (m, m)), x(op2.READ, m)), op2.ON_INTERIOR_FACETS, False),
        ("static void f3(double x[2], int layer) { x[1] = layer; }", "f3",
         (x(op2.WRITE, m),), op2.ON_TOP, True),
        ("static void f4(double x[2]) { x[0] += 1.0; }", "f4",
         (x(op2.INC, m),), op2.ON_BOTTOM, False),
    ]
    pext = op2.ExtrudedSet(base, 5, extruded_periodic=True)
    pm = op2.Map(pext, cells, 2, np.array([[0, 1], [4, 5], [8, 9], [12, 13]], dtype=np.int32),
                 offset=np.ones(2, dtype=np.int32), offset_quotient=np.array([0, 1], dtype=np.int32))
    psp = op2.Sparsity((cells ** 1, cells ** 1), [(pm, pm, (op2.ALL, op2.ON_INTERIOR_FACETS))])
    pA = op2.Mat(psp)
    cases_p = [
        ("static void p1(double A[4][4], double x[4]) { A[0][3] = x[2]; x[1] += 1.0; }", "p1",
         (pA(op2.INC, (pm, pm)), x(op2.INC, pm)), op2.ON_INTERIOR_FACETS, False),
        ("static void p2(double x[2], int layer) { x[1] = layer; }", "p2",
         (x(op2.WRITE, pm),), None, True),
    ]
    L = dev.lib()
    opts = compile_options()
    copts = (ctypes.c_char_p * len(opts))(*[o.encode() for o in opts])
    for it, (code, name, args, region, layer) in [(ext, c) for c in cases] + [(pext, c) for c in cases_p]:
        pl = LegacyParloop(op2.Kernel(code, name), it, *args, iteration_region=region, pass_layer_arg=layer)
        src = GeneratedKernel(pl.global_kernel).source(pl)
        size = ctypes.c_size_t(0)
        log = ctypes.create_string_buffer(16384)
        rc = L.op2hip_jit_code_object(src.encode(), copts, len(opts), None, ctypes.byref(size), log, len(log))
        assert rc == 0, log.value.decode()
        assert size.value > 1000


def test_extruded_map_range_check():
    """An extruded map whose offsets step past its target set in the layers
    a region runs is refused before any launch; the same map passes over
    the regions that stay inside."""
    from pyop2_amd import op2
    base = op2.Set(2)
    ext = op2.ExtrudedSet(base, 4)          # 3 cells a column
    tgt = op2.Set(7)
    m = op2.Map(ext, tgt, 1, np.array([0, 5], dtype=np.int32), offset=np.ones(1, dtype=np.int32))
    m.check_extruded(ext, op2.ON_BOTTOM)    # entries 0 and 5
    for region in (op2.ALL, op2.ON_TOP, op2.ON_INTERIOR_FACETS):
        with pytest.raises(op2.MapValueError):
            m.check_extruded(ext, region)   # column 1 reaches 5 + 2 = 7
    d = op2.Dat(base ** 1, np.zeros(2), np.float64)
    with pytest.raises(ValueError):
        op2.par_loop(op2.Kernel("static void k(double *x) { *x += 1.0; }", "k"), base, d(op2.INC),
                     iteration_region=op2.ON_INTERIOR_FACETS)


def test_integration_stub_struct_matches_the_abi():
    """The ctypes op2hip_exec in the integration guide's binding stub lists
    the fields of the C struct in order (the same list the package binds)."""
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "INTEGRATION.md")
    text = open(path).read()
    stub = text[text.index("class op2hip_exec"):text.index("_f = _lib")]
    assert re.findall(r'\("(\w+)"', stub) == [f[0] for f in dev.ExecPlan._fields_]


def test_row_list_packing_round_trips():
    """The packed row-list words hold each incidence, its row offset and
    its three slots (255 -> 31 -> missing); a slot the 5 bits cannot hold,
    or maps of another arity, keep the unpacked plan."""
    from pyop2_amd.plans import pack_row_list
    rng = np.random.default_rng(5)
    n = 1000
    ent = rng.integers(0, 2 ** 31 - 1, n).astype(np.int32)
    table = np.zeros((n, 8), dtype=np.uint8)
    row_off = rng.integers(0, 2 ** 16, n)
    table[:, 0], table[:, 1] = row_off & 0xFF, row_off >> 8
    table[:, 2:4] = rng.integers(0, 256, (n, 2))       # batch row: not packed
    slots = rng.integers(0, 31, (n, 3))
    slots[rng.random((n, 3)) < 0.1] = 255
    table[:, 4:7] = slots
    w = pack_row_list(ent, table, 3, 3)
    assert w.dtype == np.uint64
    assert np.array_equal((w & 0xFFFFFFFF).astype(np.int64), ent.astype(np.int64))
    assert np.array_equal((w >> 32) & 0xFFFF, row_off.astype(np.uint64))
    for j in range(3):
        s = (w >> np.uint64(48 + 5 * j)) & np.uint64(31)
        assert np.array_equal(np.where(s == 31, 255, s), slots[:, j])
    table[0, 5] = 31
    assert pack_row_list(ent, table, 3, 3) is None
    assert pack_row_list(ent, table, 4, 4) is None


def test_integration_p2_stub_argtypes_match_the_abi():
    """The guide's P2 fused stub declares the entry point's C signature (the
    one the package binds): start, end, the Mat descriptor and five data /
    map pointers, lam and mu, then the plan and the stream."""
    import ctypes
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "INTEGRATION.md")
    text = open(path).read()
    stub = text[text.index("_g.argtypes"):text.index("def compile_p2_fused")]
    names = re.findall(r"ctypes\.(c_\w+|POINTER)", stub)
    kinds = {"c_int32": ctypes.c_int32, "c_void_p": ctypes.c_void_p, "c_double": ctypes.c_double,
             "POINTER": ctypes.c_void_p}
    want = [getattr(t, "_type_", None) for t in dev._SIGNATURES["op2hip_wrap_p2_elasticity_fused"][1]]
    got = [kinds[n]._type_ for n in names]
    assert got == want


@pytest.mark.parametrize("maps_used,tile", [((0,), 256), ((0, 1), 128), ((0,), 37)])
def test_tile_colouring_is_conflict_free_and_matches_the_oracle(maps_used, tile):
    """Partition colouring of the staged tiles (``plan.rst:46-51, 78-80``):
    tiles of one colour share no node of the coloured maps, and the colours
    equal the oracle's greedy colouring of the tiles' padded node lists (the
    whole tile's nodes as one row).  A planted mistake -- colouring over the
    tiles' exclusive nodes only, dropping the shared ones -- must show a
    conflict."""
    from pyop2_amd import plans
    m = unit_cube_tet_p2(8, rng=workload_rng(3))
    maps = [np.ascontiguousarray(m.cell_node, np.int32), np.ascontiguousarray(m.cell_vertex, np.int32)]
    sizes = [m.nnodes, m.nverts]
    nt, tp, te, nps, nds, nxs, lcs, mx = _tiles(maps, sizes, 0, m.ncells, tile)
    sel = list(maps_used)
    colours, nc = plans.tile_colours(nt, [nps[k] for k in sel], [nds[k] for k in sel], [sizes[k] for k in sel])
    tables = [plans.tile_node_table(nt, nps[k], nds[k]) for k in sel]
    want, nwant = osp.greedy_colour(nt, tables, list(range(len(sel))), [sizes[k] for k in sel])
    assert nc == nwant and np.array_equal(colours, want)
    assert colours.min() == 0 and colours.max() == nc - 1

    def conflicts(col, k):
        seen = {}
        for t in range(nt):
            for v in nds[k][nps[k][t]:nps[k][t + 1]]:
                key = (int(col[t]), int(v))
                if key in seen and seen[key] != t:
                    return True
                seen[key] = t
        return False
    for k in sel:
        assert not conflicts(colours, k)
    k = sel[0]
    excl = [nds[k][nps[k][t]:nps[k][t] + nxs[k][t]] for t in range(nt)]
    eptr = np.concatenate([[0], np.cumsum([len(x) for x in excl])]).astype(np.int32)
    wrong, _ = plans.tile_colours(nt, [eptr], [np.concatenate(excl).astype(np.int32)], [sizes[k]])
    assert conflicts(wrong, k)
