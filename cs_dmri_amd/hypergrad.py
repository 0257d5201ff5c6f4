"""Per-voxel L1-weight selection on held-out q-points: a hypergradient
descent driven by the lasso adjoint, and the plain grid search it is judged
against.

The held-out loss of voxel v is MSE_v(alpha_v) = mean_i (A_te c_v(alpha_v) -
s_te)^2 over the held-out rows, with c_v the FISTA fit on the training rows.
Its derivative comes from ``ShoreLasso`` (implicit differentiation on the
active set); the descent runs on u = ln(alpha) with a per-voxel Adam step and
warm-started fits. The grid search fits every voxel at each alpha of a
log-spaced list and keeps, per voxel, the alpha with the smallest held-out
MSE. The reference selects alpha by BIC or by its L2-anchored rule
(src/solver/alpha.rs:107-165); neither uses held-out data, so both methods
here are new.
"""

from __future__ import annotations

import math
import time

import numpy as np
import torch

from .autograd import LassoDesign, lasso

__all__ = ["heldout_mse", "grid_search", "hypergradient_search"]


def heldout_mse(coef: torch.Tensor, A_te: torch.Tensor, S_te: torch.Tensor) -> torch.Tensor:
    return ((coef @ A_te.T - S_te) ** 2).mean(dim=1)


def grid_search(design: LassoDesign, S_tr, A_te, S_te, alphas, *, max_iter=500, tol=1e-6):
    """Per-voxel best alpha over ``alphas`` by held-out MSE. Returns a dict
    with the chosen alpha (V,), its MSE (V,), the (n_alpha, V) MSE table and
    the wall time."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    table = []
    for a in alphas:
        coef, _ = design.forward(S_tr, float(a), max_iter=max_iter, tol=tol)
        table.append(heldout_mse(coef, A_te, S_te))
    table = torch.stack(table)
    best = torch.argmin(table, dim=0)
    torch.cuda.synchronize()
    al = torch.as_tensor(np.asarray(alphas, dtype=np.float32), device=S_tr.device)
    return {"alpha": al[best], "mse": table.gather(0, best[None])[0], "table": table,
            "seconds": time.perf_counter() - t0}


def hypergradient_search(design: LassoDesign, S_tr, A_te, S_te, *, alpha0=1e-3, steps=30, lr=0.25,
                         bounds=(1e-5, 1e-1), max_iter=500, tol=1e-6, betas=(0.9, 0.999)):
    """Adam on u = ln(alpha) per voxel, gradients from the lasso adjoint.
    Returns the final alpha and MSE, the best visited alpha and MSE per
    voxel, the mean-MSE history and the wall time."""
    V = S_tr.shape[0]
    dev = S_tr.device
    lo, hi = math.log(bounds[0]), math.log(bounds[1])
    u = torch.full((V,), math.log(alpha0), dtype=torch.float32, device=dev)
    m = torch.zeros_like(u)
    v2 = torch.zeros_like(u)
    b1, b2 = betas
    warm = None
    best_mse = torch.full((V,), float("inf"), device=dev)
    best_alpha = torch.exp(u)
    hist = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for it in range(steps):
        uu = u.clone().requires_grad_(True)
        alpha = torch.exp(uu)
        coef, _ = lasso(S_tr, alpha, design, max_iter=max_iter, tol=tol, warm=warm)
        mse = heldout_mse(coef, A_te, S_te)
        mse.sum().backward()
        g = uu.grad.detach()
        with torch.no_grad():
            better = mse < best_mse
            best_mse = torch.where(better, mse, best_mse)
            best_alpha = torch.where(better, alpha.detach(), best_alpha)
            hist.append(float(mse.mean()))
            g = torch.nan_to_num(g, nan=0.0, posinf=0.0, neginf=0.0)
            m = b1 * m + (1 - b1) * g
            v2 = b2 * v2 + (1 - b2) * g * g
            mh = m / (1 - b1 ** (it + 1))
            vh = v2 / (1 - b2 ** (it + 1))
            u = torch.clamp(u - lr * mh / (torch.sqrt(vh) + 1e-12), lo, hi)
        warm = coef.detach()
    with torch.no_grad():
        alpha = torch.exp(u)
        coef, _ = design.forward(S_tr, alpha, warm=warm, max_iter=max_iter, tol=tol)
        mse = heldout_mse(coef, A_te, S_te)
        better = mse < best_mse
        best_mse = torch.where(better, mse, best_mse)
        best_alpha = torch.where(better, alpha, best_alpha)
    torch.cuda.synchronize()
    return {"alpha": alpha, "mse": mse, "best_alpha": best_alpha, "best_mse": best_mse, "history": hist,
            "seconds": time.perf_counter() - t0}
