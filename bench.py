"""Benchmark of the indirect INC par_loop and CSR assembly on one MI355X.

Headline: one step = one ``op2.par_loop`` of the fused P2 tetrahedral
elasticity kernel over the 64^3 unit cube (1,572,864 cells, 2,146,689 P2
nodes): INC of the 30x30 element matrices into the float64 BSR(3) Mat
(60,859,905 blocks) and INC of the element residuals into a vector Dat,
inputs resident in HBM.  Before the W warmup steps the loop runs untimed
for half a second, so the device has left its idle clocks.  ``value`` is
cells/s over the timed steps; the
kernel duration behind ``roofline`` is the time between one HIP event pair
recorded on the launch stream around those steps, divided by their number
(the parts' durations come from further steps with an event pair around
each launch, reported beside it).  The timed
loop's result from zero is compared with the oracle on the whole Mat and
the whole residual, and ``cpu_baseline`` times the oracle's restatement of
the reference's sequential wrapper (reference compiler flags, one core) on
the whole mesh.

The same run measures, as ``secondary``: the P1 Laplacian CSR assembly on
the permuted 1024^2 square (plus a cache-flushed variant), the P1 lumped
mass / Laplace residual on the 256^2 square, the P2 residual and assembly
alone and the float32 variants, the adjoint/forward time ratio, the batched
512-mesh launch against plain launches, and the streaming bandwidth.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--quick]
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "cells/s + eff. HBM GB/s: indirect INC par_loop & CSR assembly; adjoint/fwd ratio"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_ms(timer, name):
    t = timer.times(name)
    return float(np.mean(t)) if t else float("nan")


def time_loop(pl, reps, name):
    """Average duration (HIP events around each launch) of ``reps`` runs."""
    from pyop2_amd import device as dev
    pl()
    dev.torch().cuda.synchronize()
    with dev.launch_timer() as t:
        for _ in range(reps):
            pl()
    dev.torch().cuda.synchronize()
    return kernel_ms(t, name)


def time_many(pl, n):
    """Average duration of ``n`` back-to-back runs between one event pair:
    the rate at which the host issues the loop (for a loop this short the
    Python side of a par_loop, not the GPU, sets it)."""
    from pyop2_amd import device as dev
    torch = dev.torch()
    pl()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream(dev.device())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(n):
        pl()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / n


def graph_kernel_ms(pl, n):
    """GPU time per run of ``n`` runs captured in one HIP graph and replayed
    between one event pair (no host issue cost, no event per launch), or
    None if the loop cannot be captured."""
    from pyop2_amd import device as dev
    torch = dev.torch()
    pl()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    try:
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                for _ in range(n):
                    pl()
    except RuntimeError as e:
        log(f"graph capture of {n} runs failed ({e}); per-launch events instead")
        torch.cuda.synchronize()
        return None
    torch.cuda.current_stream().wait_stream(side)
    g.replay()
    torch.cuda.synchronize()
    s = torch.cuda.current_stream(dev.device())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    g.replay()
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / n


def cpu_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count()


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def measured_traffic(key):
    """HBM bytes per launch recorded for ``key`` by a FETCH_SIZE / WRITE_SIZE
    profile under profiles/ (``traffic_bytes_per_launch KEY BYTES``), or None."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*.txt"))):
        with open(path) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == key:
                    return float(parts[2]), os.path.relpath(path, ROOT)
    return None


def copy_bandwidth(nbytes=1 << 29, reps=10, calibration=True):
    """Streaming kernel over ``nbytes`` each way (HIP events): the measured
    copy bandwidth.  With ``calibration`` it is y = -x (torch's neg kernel),
    the calibration kernel of known bytes of the FETCH_SIZE / WRITE_SIZE
    profiles, so only those launches may use it; otherwise y = 2 x (torch's
    multiply kernel, the same streaming)."""
    from pyop2_amd import device as dev
    torch = dev.torch()
    x = torch.ones(nbytes // 8, dtype=torch.float64, device=dev.device())
    y = torch.empty_like(x)

    def step():
        if calibration:
            torch.neg(x, out=y)
        else:
            torch.mul(x, 2.0, out=y)
    step()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, b in ev:
        a.record()
        step()
        b.record()
    torch.cuda.synchronize()
    ms = float(np.median([a.elapsed_time(b) for a, b in ev]))
    del x, y
    return 2 * nbytes / (ms * 1e-3) / 1e9, ms


def cold_kernel_ms(pl, reps, name):
    """Like time_loop, with 1 GiB written between launches so that the
    256 MB last-level cache holds none of the loop's data."""
    from pyop2_amd import device as dev
    torch = dev.torch()
    scratch = torch.empty(1 << 27, dtype=torch.float64, device=dev.device())
    pl()
    torch.cuda.synchronize()
    ts = []
    for r in range(reps):
        scratch.fill_(float(r))
        torch.cuda.synchronize()
        with dev.launch_timer() as t:
            pl()
        torch.cuda.synchronize()
        ts.append(kernel_ms(t, name))
    del scratch
    return float(np.median(ts))


# ---------------------------------------------------------------------------
# P2 workload: loops and algorithmic bytes

def p2_loops(m3, dt):
    """op2 objects of the P2 problem in dtype ``dt`` and its three loops
    (residual, assembly, fused) with their algorithmic bytes per launch."""
    from pyop2_amd import elements, op2
    isz = np.dtype(dt).itemsize
    cells = op2.Set(m3.ncells, "cells")
    vset = op2.Set(m3.nverts, "verts")
    nodes = op2.Set(m3.nnodes, "nodes")
    c2v = op2.Map(cells, vset, 4, m3.cell_vertex, "c2v")
    c2n = op2.Map(cells, nodes, 10, m3.cell_node, "c2n")
    X = op2.Dat(vset ** 3, m3.coords.astype(dt), dt, name="X")
    u = op2.Dat(nodes ** 3, m3.u.astype(dt), dt, name="u")
    R = op2.Dat(nodes ** 3, dtype=dt, name="R")
    sp = op2.Sparsity((nodes ** 3, nodes ** 3), [(c2n, c2n, None)])
    nnzb = sp.block_nnz
    A = op2.Mat(sp, dt)
    maps_b = m3.ncells * (40 + 16)
    x_b = m3.nverts * 3 * isz
    u_b = m3.nnodes * 3 * isz
    r_b = 2 * u_b
    asm_b = nnzb * 9 * isz * 2 + nnzb * 4 + (m3.nnodes + 1) * 4 + maps_b + x_b
    loops = {
        "residual": (op2.ParLoop(elements.p2_elasticity_residual(), cells, R(op2.INC, c2n), X(op2.READ, c2v),
                                 u(op2.READ, c2n)), "p2_elasticity_residual", maps_b + x_b + u_b + r_b),
        "assembly": (op2.ParLoop(elements.p2_elasticity_matrix(), cells, A(op2.INC, (c2n, c2n)),
                                 X(op2.READ, c2v)), "p2_elasticity_matrix", asm_b),
        "fused": (op2.ParLoop(elements.p2_elasticity_fused(), cells, A(op2.INC, (c2n, c2n)), R(op2.INC, c2n),
                              X(op2.READ, c2v), u(op2.READ, c2n)), "p2_elasticity_fused", asm_b + u_b + r_b),
    }
    objs = dict(cells=cells, c2v=c2v, c2n=c2n, X=X, u=u, R=R, A=A, sp=sp, nnzb=nnzb)
    return objs, loops


def c3_oracle(m3, cells_sel=None, pattern=None, flavour="strict"):
    """The oracle's residual and BSR(3) Mat of the P2 loop over the cells
    ``cells_sel`` (all by default) on ``pattern`` (host rowptr, colidx;
    default: the oracle's own closed-form pattern of those cells).  Returns
    (R, Mat, seconds of the residual pass, seconds of the assembly pass,
    Mat insertions that missed the pattern)."""
    from oracle import seqwrap, sparsity as osp
    from pyop2_amd import elements
    if cells_sel is None:
        cn, cv, n = m3.cell_node, m3.cell_vertex, m3.ncells
    else:
        cn = np.ascontiguousarray(m3.cell_node[cells_sel], dtype=np.int32)
        cv = np.ascontiguousarray(m3.cell_vertex[cells_sel], dtype=np.int32)
        n = len(cells_sel)
    if pattern is None:
        rp, ci = osp.pattern(m3.nnodes, m3.nnodes, [(cn, cn, n)])
    else:
        rp, ci = pattern
    Ao = seqwrap.CSRMat(rp, ci, np.zeros((len(ci), 3, 3)), 3, 3, block=True)
    Ro = np.zeros((m3.nnodes, 3))
    rcode = elements.kernel_source("p2_elasticity_residual", lam=1.0, mu=0.5)
    mcode = elements.kernel_source("p2_elasticity_matrix", lam=1.0, mu=0.5)
    t0 = time.perf_counter()
    seqwrap.run(rcode, "p2_elasticity_residual", [seqwrap.DatArg(Ro, "INC", cn), seqwrap.DatArg(m3.coords, "READ", cv),
                                                  seqwrap.DatArg(m3.u, "READ", cn)], [(0, n)], flavour=flavour)
    t1 = time.perf_counter()
    miss = seqwrap.run(mcode, "p2_elasticity_matrix", [seqwrap.MatArg(Ao, "INC", cn, cn, 3, 3),
                                                       seqwrap.DatArg(m3.coords, "READ", cv)], [(0, n)],
                       flavour=flavour)
    t2 = time.perf_counter()
    return Ro, Ao, t1 - t0, t2 - t1, miss


def normwise(a, b):
    d = np.linalg.norm((np.asarray(a, dtype=np.float64) - b).ravel())
    return float(d / np.linalg.norm(np.asarray(b).ravel()))


def c3_parity(m3, pl, A, R, sample_rows=8192):
    """One pass of the timed fused loop from zero against the oracle:

    * pattern: the device pattern must hold every element block the oracle
      inserts (no missed insertion), have sorted unique rows and exactly the
      closed-form count (60,859,905 at 64^3) -- together the same pattern --
      and ``sample_rows`` rows spread evenly over the mesh (so over every
      part of the row-batch sequence) must equal the oracle's own pattern;
    * values: the whole Mat and the whole residual, normwise."""
    from oracle import sparsity as osp
    from pyop2_amd import device as dev
    A.zero()
    R.zero()
    pl()
    A.assemble()
    rp = dev.to_host(A.rowptr).astype(np.int32)
    ci = dev.to_host(A.colidx)
    steps = np.diff(ci.astype(np.int64))
    inside = np.ones(len(steps), dtype=bool)
    starts = rp[1:-1].astype(np.int64)
    inside[starts[(starts > 0) & (starts < len(ci))] - 1] = False
    rows_sorted = bool(np.all(steps[inside] > 0))
    Ro, Ao, tr, ta, miss = c3_oracle(m3, pattern=(rp, ci))
    sel = np.linspace(0, m3.nnodes - 1, min(sample_rows, m3.nnodes)).astype(np.int64)
    cells = np.nonzero(np.isin(m3.cell_node, sel).any(axis=1))[0]
    cn = np.ascontiguousarray(m3.cell_node[cells], dtype=np.int32)
    srp, sci = osp.pattern(m3.nnodes, m3.nnodes, [(cn, cn, len(cells))])
    sample_ok = all(np.array_equal(ci[rp[r]:rp[r + 1]], sci[srp[r]:srp[r + 1]]) for r in sel)
    gv = dev.to_host(A.device_values).astype(np.float64).reshape(-1, 3, 3)[:len(ci)]
    mat_err = normwise(gv, Ao.values)
    del gv
    res_err = normwise(dev.to_host(R.device_data_ro).reshape(-1, 3), Ro)
    return dict(pattern_missed_insertions=int(miss), pattern_rows_sorted_unique=rows_sorted,
                block_nnz=int(len(ci)), pattern_sample_rows=int(len(sel)), pattern_sample_bit_exact=bool(sample_ok),
                relerr_mat_vs_cpu=mat_err, relerr_residual_vs_cpu=res_err,
                oracle_strict_s=tr + ta), (rp, ci)


def c3_headline(args):
    """The fused P2 residual + BSR(3) assembly, float64, 64^3."""
    from pyop2_amd import device as dev
    torch = dev.torch()
    from pyop2_amd.meshes import unit_cube_tet_p2, workload_rng
    n3 = 16 if args.quick else 64
    m3 = unit_cube_tet_p2(n3, rng=workload_rng(3))
    objs, loops = p2_loops(m3, np.float64)
    pl, name, alg = loops["fused"]
    log(f"C3 mesh {n3}^3: {m3.ncells} cells, {m3.nnodes} P2 nodes, {objs['nnzb']} blocks")
    # first launch (tunes on a first encounter), then half a second of the
    # loop so the device clocks have left their idle state, then the W warmup
    # steps, all untimed
    pl()
    torch.cuda.synchronize()
    settle = time.perf_counter() + 0.5
    while time.perf_counter() < settle:
        pl()
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        pl()
    torch.cuda.synchronize()
    # the timed region: one event pair on the launch stream around the K
    # back-to-back launches (the host issues a launch in about 0.2 ms, so the
    # stream never drains and the pair sees device time only)
    s = torch.cuda.current_stream(dev.device())
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(s)
    for _ in range(args.steps):
        pl()
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kms = e0.elapsed_time(e1) / args.steps
    # the parts, each launch bracketed by events of its own (a report: the
    # extra event records between kernels lengthen each launch a little)
    with dev.launch_timer() as timer:
        for _ in range(args.steps):
            pl()
        torch.cuda.synchronize()
    kms_bracketed = kernel_ms(timer, name)
    parts = {k: kernel_ms(timer, k) * 1e3 for k in ("p2_elasticity_matrix", "p2_elasticity_residual")
             if timer.times(k)}
    executor = pl.global_kernel.compile()
    strategy = executor.strategy_used
    launch_config = getattr(executor, "last_config", None)
    ms_per_step = wall * 1e3 / args.steps
    achieved = alg / (kms * 1e-3) / 1e9
    log(f"C3 fused f64: {ms_per_step:.3f} ms/step, kernel {kms * 1e3:.1f} us per launch over {args.steps} "
        f"launches ({kms_bracketed * 1e3:.1f} us with an event pair per launch; {strategy}; parts {parts}), "
        f"{achieved:.0f} GB/s algorithmic = {achieved / HBM_PEAK_GBS:.1%} of 8 TB/s")
    parity, pattern = c3_parity(m3, pl, objs["A"], objs["R"])
    log(f"C3 parity (one pass from zero, whole Mat and residual): {parity}")
    cpu = None
    if not args.no_cpu:
        ts = []
        budget = time.time() + 30.0
        while len(ts) < 3 and (time.time() < budget or not ts):
            _, _, tr, ta, _ = c3_oracle(m3, pattern=pattern, flavour="fast")
            ts.append((tr, ta))
        tr, ta = (float(np.median([t[i] for t in ts])) for i in (0, 1))
        cpu_s = tr + ta
        cpu = {"value": m3.ncells / cpu_s, "unit": "cells/s", "cores": 1, "kind": "port",
               "sample": f"the whole {n3}^3 workload ({m3.ncells} cells, residual {tr:.2f} s + BSR(3) "
                         f"assembly {ta:.2f} s per pass), median of {len(ts)} passes of the oracle's "
                         f"sequential wrapper (gcc -march=native -O3 -ffast-math) on one core of {cpu_model()} "
                         f"({cpu_cores()} cores visible)"}
        log(f"C3 CPU baseline: {cpu_s:.2f} s per pass = {cpu['value']:.3e} cells/s; GPU kernel "
            f"{cpu_s / (kms * 1e-3):.0f}x faster")
    tkey = f"p2_elasticity_fused:{strategy}:{n3}:f64"
    tr_meas = measured_traffic(tkey)
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": tr_meas[0] if tr_meas else None,
            "traffic_source": tr_meas[1] if tr_meas else None, "traffic_key": tkey,
            "kernel": (f"p2_elasticity_fused ({strategy}"
                       + (f", residual overlapped {launch_config['overlap']} the assembly"
                          if isinstance(launch_config, dict) and launch_config.get("overlap") else "")
                       + "): " + ", ".join(f"{k} {v:.1f} us" for k, v in parts.items())
                       + (" between each part's own event pair" if parts else "")),
            "kernel_us": kms * 1e3, "alg_bytes_per_launch": alg}
    detail = dict(cells=m3.ncells, kernel_us=kms * 1e3, kernel_us_bracketed=kms_bracketed * 1e3,
                  kernel_parts_us=parts, launch_config=launch_config,
                  ms_per_step=ms_per_step, strategy=strategy, alg_bytes=alg, gbs=achieved,
                  frac=achieved / HBM_PEAK_GBS, **parity)
    if cpu:
        detail["speedup_vs_cpu"] = (m3.ncells / cpu["value"]) / (kms * 1e-3)
    del objs, loops, pl
    torch.cuda.empty_cache()
    return m3, dict(value=m3.ncells / (wall / args.steps), ms_per_step=ms_per_step, roofline=roof,
                    cpu_baseline=cpu, detail=detail, workload=n3)


# ---------------------------------------------------------------------------
# C2: P1 Laplacian CSR assembly on the permuted 1024^2 square

def c2_bytes(m, nnz):
    return (m.ncells * 3 * 4 + m.nverts * 2 * 8 + nnz * 8 * 2 + nnz * 4 + (m.nverts + 1) * 4)


def c2(args):
    from pyop2_amd import device as dev, elements, op2
    from pyop2_amd.meshes import unit_square_tri, workload_rng
    from oracle import seqwrap, sparsity as osp
    torch = dev.torch()
    n2 = 256 if args.quick else 1024
    m2 = unit_square_tri(n2, rng=workload_rng(2), permute=True)
    cells = op2.Set(m2.ncells, "cells")
    verts = op2.Set(m2.nverts, "verts")
    c2v = op2.Map(cells, verts, 3, m2.cell_vertex, "c2v")
    X = op2.Dat(verts ** 2, m2.coords, name="X")
    t0 = time.time()
    sp = op2.Sparsity((verts ** 1, verts ** 1), [(c2v, c2v, None)])
    nnz = sp.block_nnz
    torch.cuda.synchronize()
    t_sp = time.time() - t0
    A = op2.Mat(sp, np.float64)
    step = op2.ParLoop(elements.p1_laplace_matrix(), cells, A(op2.INC, (c2v, c2v)), X(op2.READ, c2v))
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    with dev.launch_timer() as timer:
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
    kms = kernel_ms(timer, "p1_laplace_matrix")
    strategy = step.global_kernel.compile().strategy_used
    alg = c2_bytes(m2, nnz)
    achieved = alg / (kms * 1e-3) / 1e9
    kms_cold = float("nan") if args.no_cold else cold_kernel_ms(step, max(3, min(args.steps, 10)),
                                                                "p1_laplace_matrix")
    # one pass from zero against the oracle
    A.zero()
    step()
    A.assemble()
    rp, ci = osp.pattern(m2.nverts, m2.nverts, [(m2.cell_vertex, m2.cell_vertex, m2.ncells)])
    ref = seqwrap.CSRMat(rp, ci, np.zeros(len(ci)))
    code = elements.kernel_source("p1_laplace_matrix")
    cargs = [seqwrap.MatArg(ref, "INC", m2.cell_vertex, m2.cell_vertex),
             seqwrap.DatArg(m2.coords, "READ", m2.cell_vertex)]
    seqwrap.run(code, "p1_laplace_matrix", cargs, [(0, m2.ncells)])
    grp, gci, gv = A.csr()
    pattern_exact = bool(np.array_equal(grp, rp) and np.array_equal(gci, ci))
    err = normwise(gv, ref.values)
    tr = measured_traffic(f"p1_laplace_matrix:{strategy}:{n2}")
    out = dict(workload=f"P1 Laplacian CSR assembly, {n2}^2 permuted", cells=m2.ncells, nnz=nnz,
               kernel_us=kms * 1e3, kernel_us_cold=kms_cold * 1e3, ms_per_step=wall * 1e3 / args.steps,
               cells_per_s=m2.ncells / (wall / args.steps), strategy=strategy, alg_bytes=alg, gbs=achieved,
               frac=achieved / HBM_PEAK_GBS, traffic=tr[0] if tr else None,
               pattern_bit_exact=pattern_exact, relerr_vs_cpu=err, sparsity_build_ms_host=t_sp * 1e3)
    if not args.no_cpu:
        ref.values[:] = 0
        seqwrap.run(code, "p1_laplace_matrix", cargs, [(0, m2.ncells)], flavour="fast")
        ts = []
        budget = time.time() + 10.0
        while len(ts) < 5 and (time.time() < budget or len(ts) < 2):
            t1 = time.perf_counter()
            seqwrap.run(code, "p1_laplace_matrix", cargs, [(0, m2.ncells)], flavour="fast")
            ts.append(time.perf_counter() - t1)
        out["cpu_cells_per_s_1core"] = m2.ncells / float(np.median(ts))
    log(f"C2: kernel {kms * 1e3:.1f} us ({strategy}), cold {kms_cold * 1e3:.1f} us, "
        f"{achieved:.0f} GB/s = {achieved / HBM_PEAK_GBS:.1%}; pattern bit-exact {pattern_exact}, "
        f"rel err {err:.1e}")
    return out


# ---------------------------------------------------------------------------
# secondary measurements

def secondary(args, m3=None):
    """C1, the P2 loops alone and in float32, and C4."""
    from pyop2_amd import device as dev, elements, op2
    from pyop2_amd.meshes import unit_cube_tet_p2, unit_square_tri, workload_rng
    torch = dev.torch()
    reps = args.reps
    res = {}

    # C1: P1 lumped mass and Laplace residual, 256^2 (launch-bound: timed
    # over 200 back-to-back launches per event pair)
    m1 = unit_square_tri(256, rng=workload_rng(1))
    cells = op2.Set(m1.ncells, "cells")
    verts = op2.Set(m1.nverts, "verts")
    c2v = op2.Map(cells, verts, 3, m1.cell_vertex, "c2v")
    X = op2.Dat(verts ** 2, m1.coords, name="X")
    u = op2.Dat(verts, m1.u, name="u")
    M = op2.Dat(verts, name="M")
    R = op2.Dat(verts, name="R")
    pm = op2.ParLoop(elements.p1_lumped_mass(), cells, M(op2.INC, c2v), X(op2.READ, c2v))
    pr = op2.ParLoop(elements.p1_laplace_residual(), cells, R(op2.INC, c2v), X(op2.READ, c2v), u(op2.READ, c2v))
    base = m1.ncells * 12 + m1.nverts * 16
    for key, pl, name, alg in (("c1a_mass", pm, "p1_lumped_mass", base + m1.nverts * 16),
                               ("c1b_residual", pr, "p1_laplace_residual", base + m1.nverts * 24)):
        host_ms = time_many(pl, 200)
        ms = graph_kernel_ms(pl, 100)
        timing = "100 runs in one HIP graph, one event pair"
        if ms is None:
            ms, timing = time_loop(pl, reps, name), "HIP events around each launch"
        res[key] = dict(cells=m1.ncells, kernel_us=ms * 1e3, cells_per_s=m1.ncells / (ms * 1e-3),
                        gbs=alg / (ms * 1e-3) / 1e9, alg_bytes=alg, timing=timing,
                        host_issue_us_per_loop=host_ms * 1e3,
                        strategy=pl.global_kernel.compile().strategy_used,
                        relerr_vs_cpu=c1_relerr(m1, pl, name))
        log(f"C1 {key}: {ms * 1e3:.2f} us per run on the GPU ({timing}), {res[key]['gbs']:.0f} GB/s "
            f"({res[key]['strategy']}); issued from Python every {host_ms * 1e3:.1f} us; "
            f"rel err vs CPU {res[key]['relerr_vs_cpu']:.1e}")

    # C3: P2 residual and assembly alone (f64), and all three in f32
    n3 = 16 if args.quick else 64
    if m3 is None:
        m3 = unit_cube_tet_p2(n3, rng=workload_rng(3))
    for dt, tag in ((np.float64, "f64"), (np.float32, "f32")):
        objs, loops = p2_loops(m3, dt)
        for key in (("residual", "assembly") if tag == "f64" else ("residual", "assembly", "fused")):
            pl, name, alg = loops[key]
            ms = time_loop(pl, reps, name)
            k = f"c3_{key}_{tag}"
            res[k] = dict(cells=m3.ncells, block_nnz=objs["nnzb"], kernel_us=ms * 1e3,
                          launch_config=getattr(pl.global_kernel.compile(), "last_config", None),
                          cells_per_s=m3.ncells / (ms * 1e-3), gbs=alg / (ms * 1e-3) / 1e9,
                          frac=alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, alg_bytes=alg,
                          strategy=pl.global_kernel.compile().strategy_used)
            log(f"C3 {k}: {ms * 1e3:.1f} us, {res[k]['gbs']:.0f} GB/s = {res[k]['frac']:.1%} ({res[k]['strategy']})")
        if tag == "f32":
            pl, _, _ = loops["fused"]
            res["c3_fused_f32"].update(c3_f32_relerr(m3, pl, objs["A"], objs["R"]))
            log(f"C3 fused f32 vs the float64 oracle: Mat {res['c3_fused_f32']['relerr_mat_vs_cpu']:.1e}, "
                f"residual {res['c3_fused_f32']['relerr_residual_vs_cpu']:.1e}")
        if tag == "f64":
            res["c4_adjoint"] = adjoint_ratio(objs, reps)
        del objs, loops
        torch.cuda.empty_cache()
    return res


def c3_f32_relerr(m3, pl, A, R):
    """One pass of the float32 fused loop from zero against the float64
    oracle on the float32-rounded inputs (whole Mat and residual)."""
    from pyop2_amd import device as dev
    import dataclasses
    A.zero()
    R.zero()
    pl()
    A.assemble()
    rp = dev.to_host(A.rowptr).astype(np.int32)
    ci = dev.to_host(A.colidx)
    m32 = dataclasses.replace(m3, coords=m3.coords.astype(np.float32).astype(np.float64),
                              u=m3.u.astype(np.float32).astype(np.float64))
    Ro, Ao, _, _, miss = c3_oracle(m32, pattern=(rp, ci))
    gv = dev.to_host(A.device_values).astype(np.float64).reshape(-1, 3, 3)[:len(ci)]
    return dict(pattern_missed_insertions=int(miss), relerr_mat_vs_cpu=normwise(gv, Ao.values),
                relerr_residual_vs_cpu=normwise(dev.to_host(R.device_data_ro).reshape(-1, 3), Ro))


def c1_relerr(m1, pl, name):
    """One pass of the timed C1 loop from zero against the oracle's
    sequential wrapper (float64, strict IEEE) on the whole mesh."""
    from oracle import seqwrap
    from pyop2_amd import elements
    out = pl.arguments[0].data
    out.zero()
    pl()
    got = np.asarray(out.data_ro, dtype=np.float64).reshape(-1)
    ref = np.zeros(m1.nverts)
    args = [seqwrap.DatArg(ref, "INC", m1.cell_vertex), seqwrap.DatArg(m1.coords, "READ", m1.cell_vertex)]
    if name == "p1_laplace_residual":
        args.append(seqwrap.DatArg(m1.u, "READ", m1.cell_vertex))
    seqwrap.run(elements.kernel_source(name), name, args, [(0, m1.ncells)])
    return normwise(got, ref)


def adjoint_ratio(objs, reps):
    """t(adjoint loop) / t(forward residual loop), P2, float64."""
    from pyop2_amd import adjoint, device as dev
    torch = dev.torch()
    prob = adjoint.P2Elasticity(objs["cells"], objs["c2v"], objs["c2n"], objs["X"], objs["u"])
    uu = objs["u"].device_data_ro.clone().requires_grad_(True)
    xx = objs["X"].device_data_ro.clone().requires_grad_(True)
    for _ in range(2):
        J = prob.objective(uu, xx)
        J.backward()
    torch.cuda.synchronize()
    with dev.launch_timer() as t:
        for _ in range(reps):
            J = prob.objective(uu, xx)
            J.backward()
    torch.cuda.synchronize()
    fwd = kernel_ms(t, "p2_elasticity_residual")
    adj = kernel_ms(t, "p2_elasticity_adjoint")
    log(f"C4 adjoint: forward {fwd * 1e3:.1f} us, adjoint {adj * 1e3:.1f} us, ratio {adj / fwd:.2f}")
    return dict(forward_us=fwd * 1e3, adjoint_us=adj * 1e3, ratio=adj / fwd)


def batched(args):
    """Config 5: Laplace CSR assembly + residual INC of 512 unit squares (n ~
    U{16..128}) as one persistent launch in a HIP graph, against the same
    loops as 512 x 2 plain par_loop launches (wall clock per full pass); the
    batched results of sampled meshes against the oracle."""
    from pyop2_amd import device as dev, elements, op2
    from pyop2_amd.batch import BatchedLaplace
    from pyop2_amd.configuration import configuration
    from pyop2_amd.meshes import batched_tri_meshes
    from oracle import seqwrap, sparsity as osp
    torch = dev.torch()
    count = 32 if args.quick else 512
    t0 = time.time()
    meshes = batched_tri_meshes(count=count)
    plain, pairs = [], []
    for m in meshes:
        cells = op2.Set(m.ncells)
        verts = op2.Set(m.nverts)
        c2v = op2.Map(cells, verts, 3, m.cell_vertex)
        X = op2.Dat(verts ** 2, m.coords)
        u = op2.Dat(verts, m.u)
        sp = op2.Sparsity((verts ** 1, verts ** 1), [(c2v, c2v, None)])
        ls = []
        for _ in range(2):
            A = op2.Mat(sp, np.float64)
            R = op2.Dat(verts)
            ls.append((op2.ParLoop(elements.p1_laplace_matrix(), cells, A(op2.INC, (c2v, c2v)), X(op2.READ, c2v)),
                       op2.ParLoop(elements.p1_laplace_residual(), cells, R(op2.INC, c2v), X(op2.READ, c2v),
                                   u(op2.READ, c2v)), A, R))
        plain.append(ls[0])
        pairs.append(ls[1])
    ncells = sum(m.ncells for m in meshes)
    nnz = sum(p[2].sparsity.block_nnz for p in plain)
    alg = sum(c2_bytes(m, p[2].sparsity.block_nnz) + m.ncells * 12 + m.nverts * 40
              for m, p in zip(meshes, plain))
    saved = configuration["strategy"]
    saved_partials = configuration["staged_partials"]

    def plain_pass():
        configuration["strategy"] = "rowgather_slots"
        for pm, _, _, _ in plain:
            pm()
        configuration["strategy"] = "staged"
        for _, pr, _, _ in plain:
            pr()
    reps = max(2, min(args.reps, 5))
    t_plain, plain_partials = None, None
    for partials in (True, False):
        configuration["staged_partials"] = partials
        plain_pass()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(reps):
            plain_pass()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t1) / reps
        if t_plain is None or t < t_plain:
            t_plain, plain_partials = t, partials
    configuration["staged_partials"] = saved_partials
    configuration["strategy"] = saved
    tb = time.time()
    batch = BatchedLaplace([(pm, pr) for pm, pr, _, _ in pairs])
    t_setup = time.time() - tb
    batch()
    torch.cuda.synchronize()
    with dev.launch_timer() as tm:
        t1 = time.perf_counter()
        for _ in range(reps * 4):
            batch()
        torch.cuda.synchronize()
        t_batch = (time.perf_counter() - t1) / (reps * 4)
    kms = kernel_ms(tm, "batched_p1_laplace")
    nb = 1 + 4 * reps   # batched passes from zero: the warm-up and the timed ones (tuning restores)
    # sampled meshes against the oracle (first, last and six in between)
    worst = 0.0
    for k in sorted(set(np.linspace(0, count - 1, 8).astype(int))):
        m, (_, _, A1, R1) = meshes[k], pairs[k]
        rp, ci = osp.pattern(m.nverts, m.nverts, [(m.cell_vertex, m.cell_vertex, m.ncells)])
        ref = seqwrap.CSRMat(rp, ci, np.zeros(len(ci)))
        seqwrap.run(elements.kernel_source("p1_laplace_matrix"), "p1_laplace_matrix",
                    [seqwrap.MatArg(ref, "INC", m.cell_vertex, m.cell_vertex),
                     seqwrap.DatArg(m.coords, "READ", m.cell_vertex)], [(0, m.ncells)])
        rr = np.zeros(m.nverts)
        seqwrap.run(elements.kernel_source("p1_laplace_residual"), "p1_laplace_residual",
                    [seqwrap.DatArg(rr, "INC", m.cell_vertex), seqwrap.DatArg(m.coords, "READ", m.cell_vertex),
                     seqwrap.DatArg(m.u, "READ", m.cell_vertex)], [(0, m.ncells)])
        worst = max(worst, normwise(dev.to_host(A1.device_values)[:len(ci)] / nb, ref.values),
                    normwise(dev.to_host(R1._device_array()) / nb, rr))
    achieved = alg / (kms * 1e-3) / 1e9
    trm = measured_traffic(f"batched_p1_laplace:{count}")
    res = dict(meshes=count, cells=ncells, nnz=nnz, plain_ms_per_pass=t_plain * 1e3,
               plain_staged_partials=plain_partials,
               batched_ms_per_pass=t_batch * 1e3, batched_kernel_us=kms * 1e3,
               speedup=t_plain / t_batch, cells_per_s=ncells / t_batch, grid=batch.grid,
               items=batch.nitems, relerr_sampled_meshes_vs_cpu=worst, setup_s=time.time() - t0,
               batch_plan_s=t_setup,
               roofline={"bound": "hbm", "alg_bytes_per_launch": alg, "achieved": achieved,
                         "frac": achieved / HBM_PEAK_GBS, "traffic": trm[0] if trm else None})
    log(f"C5 batched {count} meshes ({ncells} cells, {nnz} nnz): plain {t_plain * 1e3:.2f} ms/pass, "
        f"batched {t_batch * 1e3:.3f} ms/pass (kernel {kms * 1e3:.1f} us = {achieved:.0f} GB/s, "
        f"{achieved / HBM_PEAK_GBS:.1%}; grid {batch.grid}, {batch.nitems} items): {t_plain / t_batch:.1f}x; "
        f"8 sampled meshes vs oracle, worst rel err {worst:.1e}")
    return res


def run(args):
    from pyop2_amd import device as dev
    from pyop2_amd.configuration import configuration
    torch = dev.torch()
    gpu_name = torch.cuda.get_device_name(dev.device())
    log(f"device {dev.device()} {gpu_name}")
    configuration["strategy"] = args.strategy
    m3, head = c3_headline(args)
    out = {"c3_fused_f64": head["detail"], "c2": c2(args)}
    if not args.primary_only:
        out.update(secondary(args, m3))
        out["c5_batched"] = batched(args)
    # 512 MiB each way: the calibration kernel of the FETCH_SIZE / WRITE_SIZE
    # profiles; 4 GiB each way: the streaming rate with the 256 MB last-level
    # cache out of play, which the roofline's copy fraction uses
    cal_gbs, cal_ms = copy_bandwidth()
    copy_gbs, copy_ms = copy_bandwidth(1 << 32, reps=5, calibration=False)
    out["stream_copy"] = dict(bytes_each_way=1 << 32, ms=copy_ms, gbs=copy_gbs,
                              calibration=dict(bytes_each_way=1 << 29, ms=cal_ms, gbs=cal_gbs))
    # every launch configuration the autotuner holds (key: family, precision,
    # iteration set sizes, map content digests)
    from pyop2_amd import autotune
    out["autotune"] = autotune.decisions()
    log(f"streaming y = -x over 512 MiB: {cal_ms * 1e3:.1f} us = {cal_gbs:.0f} GB/s; y = 2 x over 4 GiB: "
        f"{copy_ms * 1e3:.1f} us = {copy_gbs:.0f} GB/s (read + write)")
    head["roofline"]["frac_of_measured_copy"] = head["roofline"]["achieved"] / copy_gbs
    n3 = head["workload"]
    d = head["detail"]
    return {
        "metric": METRIC,
        "value": head["value"],
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"P2 tetrahedral linear elasticity, residual + BSR(3) assembly in one par_loop "
                               f"(INC of 30x30 element matrices into a float64 Mat and of element residuals "
                               f"into a vector Dat), {n3}^3 unit cube x 6 tets ({d['cells']} cells, "
                               f"{d['block_nnz']} 3x3 blocks), lambda=1, mu=0.5",
                   "strategy": d["strategy"], "launch_config": d.get("launch_config"), "device": gpu_name},
        "roofline": head["roofline"],
        "cpu_baseline": head["cpu_baseline"],
        "secondary": out,
    }


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--reps", type=int, default=10, help="repetitions of each secondary loop")
    ap.add_argument("--strategy", default="auto")
    ap.add_argument("--quick", action="store_true", help="small meshes (smoke of the harness)")
    ap.add_argument("--primary-only", action="store_true", help="the headline and C2 only")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cold", action="store_true",
                    help="skip the cache-flushed C2 variant (profiles of the warm kernels)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: pyop2_amd has no multi-GPU path "
              "(its sharding is designed, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    result = run(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
