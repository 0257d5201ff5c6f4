"""Differentiable SHORE lasso fit: a ``torch.autograd.Function`` over the
device FISTA solve and its implicit-differentiation adjoint.

The forward pass is the FistaSolver fit (src/solver/fista.rs:70-143) of a
batch of signals on one design, with a per-signal L1 weight alpha and an
optional warm start. The backward pass differentiates the solution through
its optimality conditions on the active set S = {k : c_k != 0}: with
G = A^T A and z = sign(c_S),

    G_SS c_S = A_S^T s - N alpha z,

so for an upstream gradient g the products are w = G_SS^-1 g_S,
dL/ds = A_S w and dL/dalpha = -N z^T w (FP64 Cholesky of G_SS on the device,
``shorefit_lasso_adjoint``). The reference has no derivative of its fit; this
is new. The active set is taken as fixed (the derivative of the piecewise
affine solution map inside a piece).
"""

from __future__ import annotations

import ctypes as C
import warnings

import numpy as np
import torch

from . import _lib as L

__all__ = ["LassoDesign", "ShoreLasso", "lasso"]


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


class LassoDesign:
    """A design matrix held on the device (FP64 copy, A^T A, Lipschitz
    constant, FP32 kernel operands) for repeated forward / adjoint calls."""

    def __init__(self, A):
        A = np.ascontiguousarray(A, dtype=np.float64)
        if A.ndim != 2:
            raise ValueError(f"A must be 2-D, got shape {A.shape}")
        self.N, self.K = A.shape
        h = C.c_void_p()
        L.check(L.lib().shorefit_design_create(A.ctypes.data_as(L._pd), self.N, self.K, C.byref(h)))
        self._h = h
        lip = C.c_double()
        L.check(L.lib().shorefit_design_info(self._h, None, None, C.byref(lip)))
        self.lipschitz = lip.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            L.lib().shorefit_design_destroy(h)
            self._h = None

    def _check_signals(self, signals):
        if not (signals.is_cuda and signals.dtype == torch.float32):
            raise ValueError("signals must be a float32 CUDA tensor")
        if signals.ndim != 2 or signals.shape[1] != self.N:
            raise ValueError(f"signals must have shape (V, {self.N}), got {tuple(signals.shape)}")
        return signals.detach().contiguous()

    def forward(self, signals: torch.Tensor, alpha, *, warm=None, max_iter=1000, tol=1e-6, non_negative=False):
        """Solve every row of ``signals`` (V, N); alpha is a float or a (V,)
        tensor. Returns (coefficients (V, K) float32, iterations (V,) int32)."""
        signals = self._check_signals(signals)
        V = signals.shape[0]
        dev = signals.device
        coef = torch.empty((V, self.K), dtype=torch.float32, device=dev)
        its = torch.empty((V,), dtype=torch.int32, device=dev)
        a_ptr, a_scalar = None, 0.0
        if isinstance(alpha, torch.Tensor):
            alpha = alpha.detach().to(device=dev, dtype=torch.float32).contiguous()
            if alpha.shape != (V,):
                raise ValueError(f"alpha must have shape ({V},), got {tuple(alpha.shape)}")
            a_ptr = alpha.data_ptr()
        else:
            a_scalar = float(alpha)
        w_ptr = None
        if warm is not None:
            warm = warm.detach().to(device=dev, dtype=torch.float32).contiguous()
            if warm.shape != (V, self.K):
                raise ValueError(f"warm must have shape ({V}, {self.K}), got {tuple(warm.shape)}")
            w_ptr = warm.data_ptr()
        L.check(L.lib().shorefit_lasso_forward(self._h, signals.data_ptr(), V, a_ptr, a_scalar, w_ptr,
                                               int(max_iter), float(tol), int(bool(non_negative)),
                                               coef.data_ptr(), its.data_ptr(), _stream()))
        return coef, its

    def adjoint(self, coef: torch.Tensor, grad_coef: torch.Tensor, *, want_signal=True, want_alpha=True):
        """(grad_signal (V, N) or None, grad_alpha (V,) or None, status (V,) uint8)."""
        coef = coef.detach().contiguous()
        grad_coef = grad_coef.detach().to(torch.float32).contiguous()
        V = coef.shape[0]
        dev = coef.device
        gs = torch.empty((V, self.N), dtype=torch.float32, device=dev) if want_signal else None
        ga = torch.empty((V,), dtype=torch.float32, device=dev) if want_alpha else None
        st = torch.empty((V,), dtype=torch.uint8, device=dev)
        L.check(L.lib().shorefit_lasso_adjoint(self._h, coef.data_ptr(), grad_coef.data_ptr(), V,
                                               None if gs is None else gs.data_ptr(),
                                               None if ga is None else ga.data_ptr(), st.data_ptr(), _stream()))
        return gs, ga, st


class ShoreLasso(torch.autograd.Function):
    """coef = argmin_c (1/2N)|s - A c|^2 + alpha |c|_1, differentiable in
    the signals and in alpha."""

    @staticmethod
    def forward(ctx, signals, alpha, design: LassoDesign, max_iter=1000, tol=1e-6, warm=None):
        coef, its = design.forward(signals, alpha, warm=warm, max_iter=max_iter, tol=tol)
        ctx.save_for_backward(coef)
        ctx.design = design
        ctx.alpha_is_tensor = isinstance(alpha, torch.Tensor)
        ctx.mark_non_differentiable(its)
        return coef, its

    @staticmethod
    def backward(ctx, grad_coef, _grad_its):
        (coef,) = ctx.saved_tensors
        want_s = ctx.needs_input_grad[0]
        want_a = ctx.alpha_is_tensor and ctx.needs_input_grad[1]
        gs, ga, st = ctx.design.adjoint(coef, grad_coef, want_signal=want_s, want_alpha=want_a)
        bad = st != 0
        n_bad = int(bad.sum())
        if n_bad:
            # G_SS is singular (e.g. more active coefficients than signal
            # samples): the active-set derivative does not exist there
            warnings.warn(f"lasso backward: {n_bad} of {len(st)} signals have an active set whose Gram block "
                          "A_S^T A_S is not positive definite; their gradients are set to zero", RuntimeWarning,
                          stacklevel=2)
            if gs is not None:
                gs[bad] = 0.0
            if ga is not None:
                ga[bad] = 0.0
        return gs, ga, None, None, None, None


def lasso(signals, alpha, design: LassoDesign, *, max_iter=1000, tol=1e-6, warm=None):
    """Differentiable batched lasso fit; returns (coefficients, iterations)."""
    return ShoreLasso.apply(signals, alpha, design, max_iter, tol, warm)
