"""The split-loop tuner on stub executors with a fake timer (no GPU): each
part is tuned on its own, the best configuration of every strategy of the
part that writes no matrix is kept, and the overlap modes are timed over
all of those alternatives, so a strategy that is slower alone can win
beside the assembly."""
import json

from pyop2_amd import autotune


class _Arg:
    pass


class _Part:
    def __init__(self, family, first_kind):
        self.family = family
        self.variant = "f64"
        self.spec = {"args": [(first_kind, None, None, None)]}

    def applicable(self, parloop, strategy):
        return strategy


class _Split:
    family = "fused"
    variant = "f64"

    def __init__(self):
        self.assembly = _Part("p2_assembly", "mat")
        self.residual = _Part("p2_residual", "dat")

    def _split_parts(self):
        return [(self.assembly, [0]), (self.residual, [1])]


class _Loop:
    iterset = None
    arguments = [_Arg(), _Arg()]


# standalone times (ms) of each part's configurations, and the times of the
# whole split loop by (residual strategy, overlap)
ALONE = {
    ("assembly", "rowgather_slots", (384, 480), 192): 1.66,
    ("residual", "staged", 384, 192): 0.106,
    ("residual", "staged_coloured", 384, 192): 0.165,
}
TOGETHER = {
    ("staged", None): 1.80, ("staged", "before"): 1.78, ("staged", "after"): 1.78,
    ("staged_coloured", None): 1.85, ("staged_coloured", "before"): 1.76,
    ("staged_coloured", "after"): 1.75,
}


def _fake_time(executor, parloop, cfg, outputs, repeats):
    if cfg["strategy"] == "split":
        return TOGETHER[(cfg["parts"][1]["strategy"], cfg.get("overlap"))]
    if executor.family == "p2_assembly":
        return ALONE.get(("assembly", cfg["strategy"], tuple(cfg.get("batch") or ()), cfg.get("block")), 2.0)
    return ALONE.get(("residual", cfg["strategy"], cfg.get("tile"), cfg.get("block")), 1.0)


def test_split_tuner_times_the_overlap_over_every_residual_strategy(monkeypatch):
    monkeypatch.setattr(autotune, "_time", _fake_time)
    report = {}
    best = autotune._tune_split(_Split(), _Loop(), {}, 1, report)
    assert best["strategy"] == "split"
    assert best["overlap"] == "after"
    asm, res = best["parts"]
    assert asm == {"strategy": "rowgather_slots", "batch": [384, 480], "block": 192}
    # slower alone (0.165 against 0.106 ms), faster beside the assembly
    assert res["strategy"] == "staged_coloured"
    assert (res["tile"], res["block"]) == (384, 192)
    # every (residual strategy, overlap) pair was timed
    pairs = [k for k in report if k.startswith("split:")]
    assert len(pairs) == 2 * len(autotune.SPLIT_OVERLAP)
    assert "split:rowgather_slots,staged:overlap=None" in report


def test_split_tuner_keeps_partials_when_they_win_together(monkeypatch):
    together = dict(TOGETHER)
    together[("staged", "before")] = 1.70
    monkeypatch.setattr(autotune, "_time", _fake_time)
    monkeypatch.setitem(globals(), "TOGETHER", together)
    best = autotune._tune_split(_Split(), _Loop(), {}, 1, None)
    assert best["overlap"] == "before"
    assert best["parts"][1]["strategy"] == "staged"
    assert json.loads(json.dumps(best)) == best
