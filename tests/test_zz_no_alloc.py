"""The allocation-free P2 residual entry point against the ordinary one.

``op2hip_wrap_p2_elasticity_residual_noalloc`` runs the staged residual with
the partial sums of shared tile nodes in the library's own scratch buffer.
Through the C ABI (ctypes, as a foreign caller would bind it), on two calls
in a row -- a 16^3 cube, then a smaller 12^3 one -- its result must match the
ordinary entry point's (which uses the plan's workspace) and the oracle's,
within 1e-12 normwise; a plan whose partial slots exceed the scratch is
refused before any launch.
"""
import ctypes

import numpy as np
import pytest

from helpers import TetProblem, oracle_p2, relerr
from pyop2_amd import device as dev
from pyop2_amd.meshes import unit_cube_tet_p2, workload_rng

pytestmark = pytest.mark.gpu


def _staged_exec(p):
    """A staged tile plan with shared-node partials for the P2 residual, as
    the library's Python side builds it (tiles of 256 elements)."""
    from pyop2_amd import plans
    tp = plans.tile_plan([p.c2n, p.c2v], 0, p.m.ncells, 256)
    ex = dev.ExecPlan()
    ex.strategy = dev.STRATEGIES["staged"]
    tp.exec_fields(ex, [3, 0])
    return ex, tp


def _run(fn, p, R, ex, *extra):
    torch = dev.torch()
    rc = fn(0, p.m.ncells, dev.ptr(R), dev.ptr(p.X._device_array()), dev.ptr(p.u._device_array()),
            dev.ptr(p.c2n.device_values), dev.ptr(p.c2v.device_values), ctypes.c_double(1.0),
            ctypes.c_double(0.5), ctypes.byref(ex), *extra, dev.stream())
    torch.cuda.synchronize()
    return rc


@pytest.mark.parametrize("order", ["16-then-12"])
def test_noalloc_residual_matches_the_allocating_entry_point(device, order):
    L = dev.lib()
    torch = dev.torch()
    for n in (16, 12):
        print(f"[noalloc] P2 residual {n}^3", flush=True)
        m = unit_cube_tet_p2(n, rng=workload_rng(3))
        p = TetProblem(m)
        ex, tp = _staged_exec(p)
        slots = int(tp.nslot_total[0])
        assert 0 < slots * 3 * 8 <= 64 << 20
        R0 = torch.zeros((m.nnodes, 3), dtype=torch.float64, device=dev.device())
        R1 = torch.zeros_like(R0)
        assert _run(L.op2hip_wrap_p2_elasticity_residual, p, R0, ex) == 0
        assert _run(L.op2hip_wrap_p2_elasticity_residual_noalloc, p, R1, ex, ctypes.c_int64(slots)) == 0
        a, b = dev.to_host(R0), dev.to_host(R1)
        assert relerr(b, a) < 1e-12
        assert relerr(b, oracle_p2("p2_elasticity_residual", m)) < 1e-12


def test_noalloc_residual_refuses_an_oversized_plan(device):
    L = dev.lib()
    m = unit_cube_tet_p2(4, rng=workload_rng(3))
    p = TetProblem(m)
    ex, _ = _staged_exec(p)
    R = dev.torch().zeros((m.nnodes, 3), dtype=dev.torch().float64, device=dev.device())
    too_many = (64 << 20) // 24 + 1
    rc = _run(L.op2hip_wrap_p2_elasticity_residual_noalloc, p, R, ex, ctypes.c_int64(too_many))
    assert rc == 2   # OP2HIP_ERR_ARG, nothing launched
    assert b"exceed" in L.op2hip_last_error()
    assert float(R.abs().max()) == 0.0
