"""The binding stubs of INTEGRATION.md, run as written.

The guide's ctypes code (the op2hip_exec structure, the P1 and P2 stubs) is
taken from the document and executed against the built library: the P2
fused stub's wrapper, called with the reference's (start, end, *ptrs)
arguments on a 4^3 cube, must give the oracle's Mat and residual (1e-12).
"""
import os

import numpy as np
import pytest

from helpers import TetProblem, oracle_p2, relerr
from pyop2_amd import device as dev, op2
from pyop2_amd.meshes import unit_cube_tet_p2, workload_rng

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _guide_namespace():
    """Every python block of the guide's section 2, executed in order with
    the library path pointed at this build."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text[text.index("## 2. C ABI"):text.index("## 3.")]
    blocks = [b.split("```", 1)[0] for b in sec.split("```python")[1:]]
    code = "\n".join(blocks).replace("/path/to/pyop2_amd/_lib/libop2hip.so", dev._LIB_PATH)
    ns = {}
    exec(compile(code, "INTEGRATION.md", "exec"), ns)
    return ns


def test_p2_fused_stub_as_documented(device):
    ns = _guide_namespace()
    m = unit_cube_tet_p2(4, rng=workload_rng(3))
    p = TetProblem(m)
    A = op2.Mat(p.sparsity(), np.float64)
    R = op2.Dat(p.nodes ** 3, name="R")
    wrap = ns["compile_p2_fused"](1.0, 0.5)
    desc = A.descriptor()
    import ctypes
    wrap(0, m.ncells, ctypes.addressof(desc), dev.ptr(R._device_array()), dev.ptr(p.X._device_array()),
         dev.ptr(p.u._device_array()), dev.ptr(p.c2n.device_values), dev.ptr(p.c2v.device_values))
    R._device_modified()
    dev.torch().cuda.synchronize()
    Aref = oracle_p2("p2_elasticity_matrix", m)
    _, _, v = A._host_csr()
    assert relerr(v, Aref.values) < 1e-12
    assert relerr(R.data_ro, oracle_p2("p2_elasticity_residual", m)) < 1e-12
