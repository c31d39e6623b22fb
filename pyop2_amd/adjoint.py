"""Differentiable P2 elasticity residual loop (torch.autograd).

``R(u, X)`` is the INC par_loop of ``p2_elasticity_residual`` over the cells
(hand-written HIP).  Its backward is the adjoint par_loop
``p2_elasticity_adjoint``: the transposed scatter/gather -- gather Rbar
through the P2 map, scatter-add ``ubar += K_e^T Rbar_e`` through the same map
and ``Xbar += d(Rbar_e . R_e)/dX_e`` through the vertex map -- with
hand-derived element derivatives (no reference counterpart: PyOP2 has no
differentiation).

    prob = P2Elasticity(cells, c2v, c2n, X, u)
    u_t = u.device_data_ro.clone().requires_grad_()
    J = prob.objective(u_t, X_t)     # 0.5 * ||R(u, X)||^2
    J.backward()                     # u_t.grad, X_t.grad
"""
import numpy as np

from pyop2_amd import device as dev, elements, op2

__all__ = ["P2Elasticity", "residual"]


def _bind(dat, tensor):
    """Make ``dat`` use ``tensor`` (same shape and dtype) as its device copy."""
    t = tensor.detach()
    if not t.is_contiguous():
        t = t.contiguous()
    assert tuple(t.shape) == tuple(dat.shape), (t.shape, dat.shape)
    dat._dev = t
    dat._where = 2   # device copy is the valid one
    return dat


class P2Elasticity:
    """The residual loop and its adjoint on one mesh (float64)."""

    def __init__(self, cells, c2v, c2n, X, u, lam=1.0, mu=0.5):
        self.cells, self.c2v, self.c2n = cells, c2v, c2n
        self.vdset = X.dataset
        self.ndset = u.dataset
        dt = np.float64
        self._u = op2.Dat(self.ndset, dtype=dt, name="u_in")
        self._X = op2.Dat(self.vdset, dtype=dt, name="X_in")
        self._R = op2.Dat(self.ndset, dtype=dt, name="R_out")
        self._Rb = op2.Dat(self.ndset, dtype=dt, name="Rbar")
        self._ub = op2.Dat(self.ndset, dtype=dt, name="ubar")
        self._Xb = op2.Dat(self.vdset, dtype=dt, name="Xbar")
        self._fwd = op2.ParLoop(elements.p2_elasticity_residual(lam, mu), cells,
                                self._R(op2.INC, c2n), self._X(op2.READ, c2v), self._u(op2.READ, c2n))
        self._adj = op2.ParLoop(elements.p2_elasticity_adjoint(lam, mu), cells,
                                self._ub(op2.INC, c2n), self._Xb(op2.INC, c2v), self._Rb(op2.READ, c2n),
                                self._X(op2.READ, c2v), self._u(op2.READ, c2n))

    def forward(self, u, X):
        torch = dev.torch()
        R = torch.zeros_like(u)
        _bind(self._u, u)
        _bind(self._X, X)
        _bind(self._R, R)
        self._fwd()
        return R

    def backward(self, Rbar, u, X):
        torch = dev.torch()
        ub = torch.zeros_like(u)
        Xb = torch.zeros_like(X)
        _bind(self._u, u)
        _bind(self._X, X)
        _bind(self._Rb, Rbar)
        _bind(self._ub, ub)
        _bind(self._Xb, Xb)
        self._adj()
        return ub, Xb

    def residual(self, u, X):
        """Differentiable R(u, X): (nnodes, 3) tensor."""
        return _Residual.apply(u, X, self)

    def objective(self, u, X):
        """J = 0.5 * ||R(u, X)||^2 (a 0-d tensor)."""
        R = self.residual(u, X)
        return 0.5 * (R * R).sum()


def residual(prob, u, X):
    return prob.residual(u, X)


def _function():
    torch = dev.torch()

    class Residual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, u, X, prob):
            ctx.prob = prob
            ctx.save_for_backward(u, X)
            return prob.forward(u.detach().contiguous(), X.detach().contiguous())

        @staticmethod
        def backward(ctx, Rbar):
            u, X = ctx.saved_tensors
            ub, Xb = ctx.prob.backward(Rbar.contiguous(), u.detach().contiguous(), X.detach().contiguous())
            return ub, Xb, None

    return Residual


class _Lazy:
    _fn = None

    def apply(self, *args):
        if _Lazy._fn is None:
            _Lazy._fn = _function()
        return _Lazy._fn.apply(*args)


_Residual = _Lazy()
