"""Autotuner of the hand-written loops' launch configuration.

On the first launch of a loop with ``configuration["strategy"] == "auto"``
the tuner times candidate configurations with HIP events on scratch copies
of the written arguments (so the user's data is untouched) and keeps the
fastest, in memory and in ``<cache_dir>/autotune.json``.  A configuration
is the conflict-resolution strategy plus its launch parameters:

* staged tiles: elements per tile (``tile``), the LDS budget a tile's
  staged values and accumulators must fit (``lds``, which caps the nodes of
  a tile and so sets how many tiles share a CU), the workgroup size
  (``block``) and how the nodes tiles share are reduced (``partials``:
  partial sums added by a second pass, or device atomics) -- the
  reference's plan sizes its mini-partitions to the shared memory the same
  way (``doc/sphinx/source/plan.rst:13-19``);
* coloured staged tiles (``staged_coloured``, the two-level colouring of
  ``doc/sphinx/source/plan.rst:46-51``): tile size, LDS budget and workgroup
  size;
* row gathers: the batch budget (``batch`` = most row incidences, most
  blocks of a batch, which sets its LDS) and the workgroup size;
* split loops: each part's configuration, tuned as a loop of its own (the
  best of each staged strategy kept for the parts that write no matrix),
  then which of those runs beside the assembly and whether it overlaps the
  assembly on a side stream (``overlap``).

Tuning runs in two stages: every applicable strategy with its default
parameters, then the parameter candidates of the fastest strategy.  The key
of a decision is the loop's kernel family and precision, the iteration
set's sizes and a digest of the contents of every map the loop reads, so
two meshes share a decision only when their connectivity is the same.
With ``autotune`` off the first candidate strategy is used with its default
parameters.
"""
import hashlib
import json
import os
import threading

import numpy as np

from pyop2_amd import device as dev
from pyop2_amd.configuration import configuration
from pyop2_amd.types.set import Subset

_lock = threading.Lock()
_memo = {}
_loaded = False

REPEATS = 5

# parameter candidates per strategy (the first is the strategy's default)
STAGED_PARAMS = [
    {"tile": 256, "lds": 65536, "block": 256},
    {"tile": 256, "lds": 40960, "block": 256},
    {"tile": 384, "lds": 65536, "block": 256},
    {"tile": 384, "lds": 40960, "block": 192},
    {"tile": 384, "lds": 40960, "block": 128},
    {"tile": 384, "lds": 65536, "block": 128},
    {"tile": 384, "lds": 65536, "block": 192},
    {"tile": 512, "lds": 65536, "block": 128},
    {"tile": 512, "lds": 54272, "block": 256},
    {"tile": 512, "lds": 65536, "block": 256},
    {"tile": 384, "lds": 40960, "block": 192, "partials": False},
    {"tile": 256, "lds": 65536, "block": 256, "partials": False},
]
# coloured tiles: small tiles give each colour's launch more workgroups
STAGED_COLOURED_PARAMS = [
    {"tile": 256, "lds": 65536, "block": 256},
    {"tile": 128, "lds": 32768, "block": 128},
    {"tile": 384, "lds": 40960, "block": 192},
]
# split loops: the parts in turn on the calling stream, or the parts that
# write no matrix overlapped on a side stream, issued before / after the
# assembly
SPLIT_OVERLAP = [None, "before", "after"]
ROWGATHER_PARAMS = {
    # P2 BSR(3) float64 / float32: (incidences, blocks) of a batch
    "p2": [{"batch": None, "block": 256},
           {"batch": [384, 480], "block": 256},
           {"batch": [320, 400], "block": 256},
           {"batch": [512, 640], "block": 256},
           {"batch": [384, 480], "block": 192},
           {"batch": [768, 880], "block": 256}],
    # P1 CSR
    "p1": [{"batch": None, "block": 256},
           {"batch": [512, 1024], "block": 128},
           {"batch": [2048, 3584], "block": 256}],
}


def _path():
    return os.path.join(configuration["cache_dir"], "autotune.json")


def _load():
    global _loaded
    if _loaded:
        return
    _loaded = True
    try:
        with open(_path()) as f:
            _memo.update(json.load(f))
    except (OSError, ValueError):
        pass


def _save():
    try:
        os.makedirs(configuration["cache_dir"], exist_ok=True)
        tmp = _path() + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(_memo, f)
        os.replace(tmp, _path())
    except OSError:
        pass


def candidates(executor, parloop):
    """Strategies the loop can run, the default first."""
    has_mat = executor.spec["args"][0][0] == "mat"
    if has_mat:
        if isinstance(parloop.iterset, Subset):
            return ["atomic", "coloured"]
        split = ["split"] if "split" in executor.spec else []
        return split + ["rowgather_slots", "rowgather_rows", "rowgather_cells", "rowgather_staged",
                        "rowgather_thread", "rowgather", "atomic"]
    if "staged" in executor.spec and not isinstance(parloop.iterset, Subset):
        return ["staged", "staged_coloured", "atomic", "coloured"]
    return ["atomic", "coloured"]


def param_candidates(executor, strategy):
    """Parameter sets of ``strategy`` for this family (the default first)."""
    if strategy == "staged":
        return STAGED_PARAMS
    if strategy == "staged_coloured":
        return STAGED_COLOURED_PARAMS
    if strategy in ("rowgather_slots", "rowgather_rows"):
        return ROWGATHER_PARAMS["p2" if executor.family.startswith("p2_") else "p1"]
    return [{}]


def default(executor, parloop):
    return {"strategy": candidates(executor, parloop)[0]}


def map_digest(m):
    """Digest of a map's contents (cached on the map: maps are immutable
    once built)."""
    d = m._cache.get("content-digest")
    if d is None:
        t = np.ascontiguousarray(m._table)
        h = hashlib.blake2b(digest_size=16)
        h.update(np.asarray(t.shape, dtype=np.int64).tobytes())
        h.update(t.tobytes())
        d = m._cache["content-digest"] = h.hexdigest()
    return d


def _key(executor, parloop):
    it = parloop.iterset
    maps = [map_digest(m) for a in parloop.arguments for m in a.maps]
    sub = it._plan_key()[-1] if isinstance(it, Subset) else None
    return json.dumps([executor.family, executor.variant, type(it).__name__, it.core_size, it.size,
                       it.total_size, sub, maps])


def choice(executor, parloop):
    """The loop's tuned configuration (tuned now on a first encounter)."""
    if not configuration["autotune"]:
        return default(executor, parloop)
    # the key is fixed by the loop object (set sizes, map contents): built
    # once per loop and family
    held = getattr(parloop, "__dict__", None)
    keys = held.setdefault("_tune_keys", {}) if held is not None else {}
    key = keys.get((executor.family, executor.variant))
    if key is None:
        key = keys[(executor.family, executor.variant)] = _key(executor, parloop)
    with _lock:
        _load()
        hit = _memo.get(key)
    if isinstance(hit, dict):
        return hit
    best = tune(executor, parloop)
    with _lock:
        _memo[key] = best
        _save()
    return best


def decisions():
    """Every decision held (key -> configuration), for reports."""
    with _lock:
        _load()
        return dict(_memo)


def cached(key):
    """Stored result for ``key`` (a JSON-serialisable list), or None."""
    with _lock:
        _load()
        return _memo.get(json.dumps(key))


def store(key, value):
    with _lock:
        _memo[json.dumps(key)] = value
        _save()


def _scratch_outputs(executor, parloop):
    torch = dev.torch()
    outputs = {}
    for k, ((kind, access, _, _), arg) in enumerate(zip(executor.spec["args"], parloop.arguments)):
        if access.name == "INC":
            if kind == "mat":
                outputs["mat"] = torch.zeros_like(arg.data.device_values)
            else:
                outputs[k] = torch.zeros_like(arg.data._device_array())
    return outputs


def _time(executor, parloop, cfg, outputs, repeats):
    """Mean time (ms) of ``repeats`` launches of ``cfg`` (after a warm-up
    that builds its plans); None if the configuration cannot run."""
    from pyop2_amd.exceptions import DeviceError
    torch = dev.torch()
    start, end = 0, parloop.iterset.size
    s = torch.cuda.current_stream()
    try:
        executor.launch(start, end, parloop, cfg, outputs)
    except DeviceError:
        return None
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(s)
    for _ in range(repeats):
        executor.launch(start, end, parloop, cfg, outputs)
    e1.record(s)
    e1.synchronize()
    return e0.elapsed_time(e1) / repeats


def _tune_split(executor, parloop, outputs, repeats, report):
    """Each part of a split loop tuned as a loop of its own (the assembly
    part under the slot-table row gather; the other parts under each staged
    strategy, partials and coloured tiles), on the split loop's scratch
    outputs.  The best configuration of every strategy of a part is kept:
    a part that is slower alone can be the faster one beside the assembly
    on a side stream, so the overlap modes are timed over those
    alternatives."""
    import itertools
    from pyop2_amd.registry import _SubLoop
    alternatives = []
    for sub, idx in executor._split_parts():
        part = _SubLoop(parloop.iterset, [parloop.arguments[k] for k in idx])
        sub_out = {}
        for n, k in enumerate(idx):
            if k in outputs:
                sub_out[n] = outputs[k]
        if "mat" in outputs:
            sub_out["mat"] = outputs["mat"]
        wanted = (["rowgather_slots"] if sub.spec["args"][0][0] == "mat"
                  else ["staged", "staged_coloured"])
        strats = []
        for w in wanted:
            strat = sub.applicable(part, w)
            if strat not in strats:
                strats.append(strat)
        alts = []
        for strat in strats:
            best, best_t = {"strategy": strat}, None
            for params in param_candidates(sub, strat):
                cfg = dict(params, strategy=strat)
                t = _time(sub, part, cfg, sub_out, repeats)
                if report is not None:
                    report[f"{sub.family}:{json.dumps(cfg, sort_keys=True)}"] = t
                if t is not None and (best_t is None or t < best_t):
                    best, best_t = cfg, t
            alts.append(best)
        alternatives.append(alts)
    best, best_t = None, None
    for parts in itertools.product(*alternatives):
        for overlap in SPLIT_OVERLAP:
            cfg = {"strategy": "split", "parts": list(parts)}
            if overlap:
                cfg["overlap"] = overlap
            t = _time(executor, parloop, cfg, outputs, repeats)
            if report is not None:
                names = ",".join(p["strategy"] for p in parts)
                report[f"split:{names}:overlap={overlap}"] = t
            if t is not None and (best_t is None or t < best_t):
                best, best_t = cfg, t
    return best


def tune(executor, parloop, repeats=REPEATS, report=None):
    """Time the candidates on scratch outputs; return the fastest
    configuration."""
    outputs = _scratch_outputs(executor, parloop)
    times = {}
    for strat in candidates(executor, parloop):
        t = _time(executor, parloop, {"strategy": strat}, outputs, repeats)
        if t is not None:
            times[strat] = t
    if report is not None:
        report.update(times)
    strat = min(times, key=times.get)
    if strat == "split":
        best = _tune_split(executor, parloop, outputs, repeats, report)
        return best
    best, best_t = {"strategy": strat}, times[strat]
    for params in param_candidates(executor, strat)[1:]:
        cfg = dict(params, strategy=strat)
        t = _time(executor, parloop, cfg, outputs, repeats)
        if report is not None:
            report[json.dumps(cfg, sort_keys=True)] = t
        if t is not None and t < best_t:
            best, best_t = cfg, t
    return best
