"""Adjoint of the P2 residual loop (torch.autograd) on the device.

* the hand-written adjoint loop against the oracle's adjoint kernel (1e-12);
* J = 0.5 ||R(u, X)||^2: gradient against central finite differences
  (h = 1e-6, 8 seeded directions on u and on X, float64, 32^3 mesh).  The
  u directions have N(0, 1) entries (the scale of u itself) and the X
  directions N(0, 1) * 0.1 h_mesh entries: a unit-norm direction over the
  6.4e6 entries of u would make the FD signal smaller than J's rounding.
* the gradient against the oracle's tangent-linear loop run on the CPU:
  dJ = R . TLM(du, dX) for seeded directions.
"""
import numpy as np
import pytest

from helpers import TetProblem, relerr, use_strategy
from oracle import seqwrap
from pyop2_amd import adjoint, device as dev, elements, op2
from pyop2_amd.configuration import configuration
from pyop2_amd.meshes import unit_cube_tet_p2, workload_rng

pytestmark = pytest.mark.gpu


def _jitter(m, rng, amp=0.1):
    h = 1.0 / m.n
    X = m.coords + rng.uniform(-amp * h, amp * h, m.coords.shape)
    return np.ascontiguousarray(X)


@pytest.mark.parametrize("strategy", ["atomic", "coloured", "staged", "staged_atomic", "staged_coloured"])
def test_adjoint_loop_matches_oracle(device, strategy):
    use_strategy(strategy)
    m = unit_cube_tet_p2(6, rng=workload_rng(4))
    rng = np.random.default_rng([2026, 4, 1])
    X = _jitter(m, rng)
    E = rng.standard_normal((m.nnodes, 3))
    p = TetProblem(m)
    p.X.data[:] = X
    Rb = op2.Dat(p.nodes ** 3, E, name="Rbar")
    ub = op2.Dat(p.nodes ** 3, name="ub")
    Xb = op2.Dat(p.verts ** 3, name="Xb")
    op2.par_loop(elements.p2_elasticity_adjoint(), p.cells, ub(op2.INC, p.c2n), Xb(op2.INC, p.c2v),
                 Rb(op2.READ, p.c2n), p.X(op2.READ, p.c2v), p.u(op2.READ, p.c2n))
    rub = np.zeros((m.nnodes, 3))
    rXb = np.zeros((m.nverts, 3))
    seqwrap.run(elements.kernel_source("p2_elasticity_adjoint"), "p2_elasticity_adjoint",
                [seqwrap.DatArg(rub, "INC", m.cell_node), seqwrap.DatArg(rXb, "INC", m.cell_vertex),
                 seqwrap.DatArg(E, "READ", m.cell_node), seqwrap.DatArg(X, "READ", m.cell_vertex),
                 seqwrap.DatArg(m.u, "READ", m.cell_node)], [(0, m.ncells)])
    assert relerr(ub.data_ro, rub) < 1e-12
    assert relerr(Xb.data_ro, rXb) < 1e-12


@pytest.fixture(scope="module")
def problem32():
    m = unit_cube_tet_p2(32, rng=workload_rng(4))
    rng = np.random.default_rng([2026, 4, 2])
    X = _jitter(m, rng)
    p = TetProblem(m)
    p.X.data[:] = X
    prob = adjoint.P2Elasticity(p.cells, p.c2v, p.c2n, p.X, p.u)
    return m, p, prob, X, rng


def test_gradient_finite_differences(device, problem32):
    m, p, prob, X, rng = problem32
    torch = dev.torch()
    u = p.u.device_data_ro.clone().requires_grad_(True)
    x = p.X.device_data_ro.clone().requires_grad_(True)
    J = prob.objective(u, x)
    J.backward()
    gu, gx = u.grad.detach(), x.grad.detach()
    h = 1e-6
    errs = []
    drng = np.random.default_rng([2026, 4, 3])
    with torch.no_grad():
        for k in range(8):
            du = torch.from_numpy(drng.standard_normal(u.shape)).to(u.device)
            dx = torch.from_numpy(drng.standard_normal(x.shape)).to(x.device) * (0.1 / m.n)
            for d_u, d_x in ((du, None), (None, dx)):
                uu = u.detach()
                xx = x.detach()
                if d_u is not None:
                    jp = prob.objective(uu + h * d_u, xx)
                    jm = prob.objective(uu - h * d_u, xx)
                    ad = float((gu * d_u).sum())
                else:
                    jp = prob.objective(uu, xx + h * d_x)
                    jm = prob.objective(uu, xx - h * d_x)
                    ad = float((gx * d_x).sum())
                fd = float(jp - jm) / (2 * h)
                errs.append(abs(fd - ad) / max(abs(ad), 1e-30))
    print("FD relative errors:", ["%.2e" % e for e in errs])
    assert max(errs) < 1e-6


def test_gradient_against_cpu_tangent_linear(device):
    m = unit_cube_tet_p2(8, rng=workload_rng(4))
    rng = np.random.default_rng([2026, 4, 4])
    X = _jitter(m, rng)
    p = TetProblem(m)
    p.X.data[:] = X
    prob = adjoint.P2Elasticity(p.cells, p.c2v, p.c2n, p.X, p.u)
    u = p.u.device_data_ro.clone().requires_grad_(True)
    x = p.X.device_data_ro.clone().requires_grad_(True)
    R = prob.residual(u, x)
    J = 0.5 * (R * R).sum()
    J.backward()
    Rh = R.detach().cpu().numpy()
    for k in range(3):
        du = rng.standard_normal((m.nnodes, 3))
        dX = rng.standard_normal((m.nverts, 3)) * (0.1 / m.n)
        dR = np.zeros((m.nnodes, 3))
        seqwrap.run(elements.kernel_source("p2_elasticity_tlm"), "p2_elasticity_tlm",
                    [seqwrap.DatArg(dR, "INC", m.cell_node), seqwrap.DatArg(X, "READ", m.cell_vertex),
                     seqwrap.DatArg(m.u, "READ", m.cell_node), seqwrap.DatArg(dX, "READ", m.cell_vertex),
                     seqwrap.DatArg(du, "READ", m.cell_node)], [(0, m.ncells)])
        tlm = float((Rh * dR).sum())
        ad = float((u.grad.cpu().numpy() * du).sum() + (x.grad.cpu().numpy() * dX).sum())
        assert abs(tlm - ad) <= 1e-11 * abs(tlm)
