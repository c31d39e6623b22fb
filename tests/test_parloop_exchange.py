"""Halo exchange halves around a loop (no GPU): indirect Dats on a set with a
halo object are exchanged once each, in first-use order, with the access
they are read with; a loop none of whose indirect Dats has a halo calls no
exchange hook at all (one rank)."""
import numpy as np

from pyop2_amd import op2
from pyop2_amd.types.dat import Dat

KERNEL = """
static void copy_in(double *out, const double *a, const double *b) { out[0] = a[0] + b[0]; }
"""


class RecordingHalo:
    """The reference's Halo interface, recording the calls it receives."""

    def __init__(self):
        self.calls = []

    def global_to_local_begin(self, dat, mode):
        self.calls.append(("g2l_begin", dat.name, mode))

    def global_to_local_end(self, dat, mode):
        self.calls.append(("g2l_end", dat.name, mode))

    def local_to_global_begin(self, dat, mode):
        self.calls.append(("l2g_begin", dat.name, mode))

    def local_to_global_end(self, dat, mode):
        self.calls.append(("l2g_end", dat.name, mode))


def _loop(halo):
    cells = op2.Set(4, "cells")
    nodes = op2.Set(5, "nodes", halo=halo)
    m = op2.Map(cells, nodes, 1, np.array([0, 1, 2, 4], dtype=np.int32), "m")
    a = op2.Dat(nodes, np.arange(5, dtype=np.float64), name="a")
    out = op2.Dat(cells, name="out")
    k = op2.Kernel(KERNEL, "copy_in")
    # a is read twice through the same map: exchanged once
    pl = op2.ParLoop(k, cells, out(op2.WRITE), a(op2.READ, m), a(op2.READ, m))
    return pl, a


def test_read_halo_is_exchanged_once_per_dat():
    halo = RecordingHalo()
    pl, a = _loop(halo)
    a.halo_valid = False        # as after a loop that wrote a
    pl._exchange("global_to_local_begin", frozenset((op2.READ,)))
    pl._exchange("global_to_local_end", frozenset((op2.READ,)))
    assert halo.calls == [("g2l_begin", "a", op2.WRITE), ("g2l_end", "a", op2.WRITE)]
    assert a.halo_valid
    # a valid halo is not exchanged again
    pl._exchange("global_to_local_begin", frozenset((op2.READ,)))
    assert len(halo.calls) == 2


def test_a_loop_without_halos_calls_no_exchange_hook(monkeypatch):
    pl, _ = _loop(None)
    seen = []
    for step in ("global_to_local_begin", "global_to_local_end",
                 "local_to_global_begin", "local_to_global_end"):
        monkeypatch.setattr(Dat, step, lambda self, mode, step=step: seen.append(step))
    for step in ("global_to_local_begin", "global_to_local_end",
                 "local_to_global_begin", "local_to_global_end"):
        pl._exchange(step, frozenset((op2.READ, op2.INC)))
    assert seen == []


def test_replacing_an_argument_refreshes_the_exchange_list():
    halo = RecordingHalo()
    pl, a = _loop(None)
    assert pl._static_parts()[2] is False
    nodes = op2.Set(5, "nodes2", halo=halo)
    b = op2.Dat(nodes, np.arange(5, dtype=np.float64), name="b")
    b.halo_valid = False
    pl.replace_data(1, b)
    pl._exchange("global_to_local_begin", frozenset((op2.READ,)))
    assert halo.calls == [("g2l_begin", "b", op2.WRITE)]
