"""Summaries of rocprofv3 output (the SQLite databases it writes).

Usage::

    python3 -m pyop2_amd.profiling stats DIR [DIR ...]
        per-kernel dispatch count and average / min / max duration
    python3 -m pyop2_amd.profiling counters DIR [DIR ...]
        per-kernel mean value per dispatch of every collected counter
    python3 -m pyop2_amd.profiling traffic --fetch DIR --write DIR --kernel SUB
            [--copy SUB --copy-bytes N]
        HBM bytes per dispatch of the kernels whose name contains SUB, from a
        FETCH_SIZE run and a WRITE_SIZE run (both in KiB per dispatch), each
        scaled by bytes / counter of a plain copy kernel of known size;
        with --per LOOPSUB, bytes per dispatch of the kernel LOOPSUB instead
        (the run's total over SUB's dispatches / LOOPSUB's dispatch count:
        per loop launch, for a kernel dispatched once per colour)

The ``traffic`` report ends with one line per kernel,
``traffic_bytes_per_launch <kernel-substring> <bytes>``, that the benchmark
reads to fill its roofline ``traffic`` field.
"""
import argparse
import collections
import glob
import os
import sqlite3
import sys

__all__ = ["kernel_stats", "counter_means", "traffic"]


def _dbs(dirs):
    out = []
    for d in dirs:
        if os.path.isfile(d):
            out.append(d)
        else:
            out.extend(sorted(glob.glob(os.path.join(d, "**", "*.db"), recursive=True)))
    if not out:
        raise SystemExit(f"no rocprofv3 database under {dirs}")
    return out


def kernel_stats(dirs):
    """{kernel name: dict(calls, avg_us, min_us, max_us, grid, wg, vgpr, lds)}."""
    acc = collections.defaultdict(list)
    meta = {}
    for db in _dbs(dirs):
        c = sqlite3.connect(db).cursor()
        q = ("select name, duration, grid_x, workgroup_x, vgpr_count, accum_vgpr_count, lds_size "
             "from kernels")
        for name, dur, gx, wx, vg, ag, lds in c.execute(q):
            acc[name].append(dur)
            meta[name] = (gx, wx, vg, ag, lds)
    out = {}
    for name, ds in acc.items():
        gx, wx, vg, ag, lds = meta[name]
        srt = sorted(ds)
        out[name] = dict(calls=len(ds), avg_us=sum(ds) / len(ds) / 1e3, med_us=srt[len(srt) // 2] / 1e3,
                         min_us=min(ds) / 1e3,
                         max_us=max(ds) / 1e3, total_us=sum(ds) / 1e3, grid=gx, wg=wx, vgpr=vg,
                         agpr=ag, lds=lds)
    return out


def counter_means(dirs):
    """{kernel name: {counter: (mean per dispatch, dispatches)}}."""
    acc = collections.defaultdict(lambda: collections.defaultdict(dict))
    for db in _dbs(dirs):
        c = sqlite3.connect(db).cursor()
        q = "select dispatch_id, kernel_name, counter_name, value from counters_collection"
        for did, name, ctr, val in c.execute(q):
            # a counter may be reported per XCD / instance: sum them per dispatch
            acc[name][ctr][(db, did)] = acc[name][ctr].get((db, did), 0.0) + float(val)
    out = {}
    for name, ctrs in acc.items():
        out[name] = {ctr: (sum(v.values()) / len(v), len(v)) for ctr, v in ctrs.items()}
    return out


def _pick(table, sub):
    hits = [k for k in table if sub in k]
    if not hits:
        raise SystemExit(f"no kernel matching {sub!r}; kernels: {sorted(table)[:20]}")
    return hits


def traffic(fetch_dirs, write_dirs, kernel_sub, copy_sub=None, copy_bytes=None, per_sub=None):
    """Measured HBM bytes per dispatch (per dispatch of the kernel matching
    ``per_sub`` when given): (report lines, {kernel: bytes})."""
    fetch = counter_means(fetch_dirs)
    write = counter_means(write_dirs)
    ffac = wfac = 1.0
    lines = []
    if copy_sub:
        ck = _pick(fetch, copy_sub)[0]
        fk = fetch[ck]["FETCH_SIZE"][0] * 1024.0
        wk = write[_pick(write, copy_sub)[0]]["WRITE_SIZE"][0] * 1024.0
        ffac, wfac = copy_bytes / fk, copy_bytes / wk
        lines.append(f"calibration kernel: {ck[:100]}")
        lines.append(f"  moves {copy_bytes} B each way; FETCH_SIZE reads {fk:.0f} B (factor {ffac:.3f}), "
                     f"WRITE_SIZE reads {wk:.0f} B (factor {wfac:.3f})")
    launches = None
    if per_sub:
        pk = _pick(fetch, per_sub)[0]
        launches = fetch[pk]["FETCH_SIZE"][1]
        lines.append(f"per launch: {launches} dispatches of {pk[:100]}")
    res = {}
    for k in _pick(fetch, kernel_sub):
        n = fetch[k]["FETCH_SIZE"][1]
        scale = n / launches if launches else 1.0
        f = fetch[k]["FETCH_SIZE"][0] * 1024.0 * ffac * scale
        wks = [w for w in write if w == k]
        w = write[wks[0]]["WRITE_SIZE"][0] * 1024.0 * wfac * scale if wks else float("nan")
        res[k] = f + w
        lines.append(f"kernel: {k[:140]}")
        unit = "per launch" if launches else "per dispatch"
        lines.append(f"  dispatches {n}; fetched {f:.4g} B, written {w:.4g} B, "
                     f"total {f + w:.4g} B {unit} (calibrated)")
    return lines, res


def _short(name, width=110):
    return name if len(name) <= width else name[:width - 3] + "..."


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python3 -m pyop2_amd.profiling")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("dirs", nargs="+")
    s.add_argument("--top", type=int, default=25)
    s.add_argument("--match", action="append", default=[],
                   help="also list the kernels whose name contains this (repeatable)")
    c = sub.add_parser("counters")
    c.add_argument("dirs", nargs="+")
    c.add_argument("--kernel", default="")
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", nargs="+", required=True)
    t.add_argument("--write", nargs="+", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--key", help="substring written on the traffic line (default: --kernel)")
    t.add_argument("--copy")
    t.add_argument("--copy-bytes", type=float)
    t.add_argument("--per", help="report bytes per dispatch of the kernel containing this")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        st = kernel_stats(a.dirs)
        ranked = sorted(st.items(), key=lambda kv: -kv[1]["total_us"])
        rows = ranked[:a.top] + [kv for kv in ranked[a.top:] if any(m in kv[0] for m in a.match)]
        print(f"{'calls':>6} {'avg us':>10} {'med us':>10} {'min us':>10} {'max us':>10} {'grid':>9} {'wg':>4} "
              f"{'vgpr':>4} {'lds':>6}  kernel")
        for name, r in rows:
            print(f"{r['calls']:6d} {r['avg_us']:10.2f} {r['med_us']:10.2f} {r['min_us']:10.2f} {r['max_us']:10.2f} "
                  f"{r['grid']:9d} {r['wg']:4d} {r['vgpr']:4d} {r['lds']:6d}  {_short(name)}")
    elif a.cmd == "counters":
        for name, ctrs in sorted(counter_means(a.dirs).items()):
            if a.kernel not in name:
                continue
            print(_short(name, 160))
            for ctr, (v, n) in sorted(ctrs.items()):
                print(f"    {ctr:28s} {v:16.6g}   ({n} dispatches)")
    else:
        lines, res = traffic(a.fetch, a.write, a.kernel, a.copy, a.copy_bytes, a.per)
        print("\n".join(lines))
        for k, v in res.items():
            print(f"traffic_bytes_per_launch {a.key or a.kernel} {v:.0f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
