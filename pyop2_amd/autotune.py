"""Autotuner of the conflict-resolution strategy of hand-written loops.

On the first launch of a (kernel family, precision, iteration set, maps)
combination with ``configuration["strategy"] == "auto"``, every applicable
strategy is timed with HIP events on scratch copies of the written
arguments (so the user's data is untouched), and the fastest is kept in
memory and in ``<cache_dir>/autotune.json``.  With ``autotune`` disabled the
first candidate is used: ``rowgather_slots`` for Mat assembly, ``staged``
for Dat INC over a Set.
"""
import json
import os
import threading

from pyop2_amd import device as dev
from pyop2_amd.configuration import configuration
from pyop2_amd.types.set import Subset

_lock = threading.Lock()
_memo = {}
_loaded = False

REPEATS = 3


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
    has_mat = executor.spec["args"][0][0] == "mat"
    if has_mat:
        if isinstance(parloop.iterset, Subset):
            return ["atomic", "coloured"]
        split = ["split"] if "split" in executor.spec else []
        return split + ["rowgather_slots", "rowgather_rows", "rowgather_cells", "rowgather_staged",
                        "rowgather_thread", "rowgather", "atomic"]
    if "staged" in executor.spec and not isinstance(parloop.iterset, Subset):
        return ["staged", "atomic", "coloured"]
    return ["atomic", "coloured"]


def default(executor, parloop):
    return candidates(executor, parloop)[0]


def _key(executor, parloop):
    it = parloop.iterset
    maps = []
    for a in parloop.arguments:
        for m in a.maps:
            maps.append((m.name, m.arity, m.toset.total_size))
    return json.dumps([executor.family, executor.variant, type(it).__name__, it.size, it.core_size, maps])


def choice(executor, parloop):
    if not configuration["autotune"]:
        return default(executor, parloop)
    key = _key(executor, parloop)
    with _lock:
        _load()
        hit = _memo.get(key)
    if hit is not None:
        return hit
    best = tune(executor, parloop)
    with _lock:
        _memo[key] = best
        _save()
    return best


def cached(key):
    """Stored result for ``key`` (a JSON-serialisable list), or None."""
    with _lock:
        _load()
        return _memo.get(json.dumps(key))


def store(key, value):
    with _lock:
        _memo[json.dumps(key)] = value
        _save()


def tune(executor, parloop, repeats=REPEATS, report=None):
    """Time each candidate on scratch outputs; return the fastest."""
    torch = dev.torch()
    outputs = {}
    for k, ((kind, access, _, _), arg) in enumerate(zip(executor.spec["args"], parloop.arguments)):
        if access.name == "INC":
            if kind == "mat":
                outputs["mat"] = torch.zeros_like(arg.data.device_values)
            else:
                outputs[k] = torch.zeros_like(arg.data._device_array())
    start, end = 0, parloop.iterset.size
    times = {}
    s = torch.cuda.current_stream()
    for cand in candidates(executor, parloop):
        executor.launch(start, end, parloop, cand, outputs)   # warm-up, builds plans
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(s)
        for _ in range(repeats):
            executor.launch(start, end, parloop, cand, outputs)
        e1.record(s)
        e1.synchronize()
        times[cand] = e0.elapsed_time(e1) / repeats
    if report is not None:
        report.update(times)
    return min(times, key=times.get)
