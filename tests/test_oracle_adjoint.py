"""The oracle's lasso adjoint (implicit differentiation on the active set),
pinned on the CPU. The reference has no derivative of its fit, so the pins
are mathematical:

* orthogonal design (A^T A = N I): c = soft(A^T s / N, alpha) in closed form,
  so dL/ds = A_S g_S / N and dL/dalpha = -sum_S sign(c) g exactly;
* central finite differences (h = 1e-4 relative) of the exact, KKT-verified
  lasso minimiser (the active-set polish of a converged FISTA solution) on a
  random well-conditioned design and on a SHORE design;
* the polish itself: it reproduces a tightly converged FISTA solution, and a
  planted sign error or a dropped active coefficient is detected.
"""

import numpy as np
import pytest

from cs_dmri_amd import workloads as W


def _fd_check(orc, A, s, alpha, rng, n_dirs=3, rel=1e-4):
    N, K = A.shape
    c0, _ = orc.fista(A, s, alpha, max_iter=200000, tol=1e-13)
    c, st = orc.lasso_polish(A, s, alpha, c0)
    assert st == 0
    g = rng.standard_normal(K)
    gs, ga = orc.lasso_adjoint(A, c[None], g[None])
    # d<g, c>/ds along random directions, and d<g, c>/dalpha
    for _ in range(n_dirs):
        d = rng.standard_normal(N)
        h = rel * np.linalg.norm(s) / np.linalg.norm(d)
        cp, sp = orc.lasso_polish(A, s + h * d, alpha, c)
        cm, sm = orc.lasso_polish(A, s - h * d, alpha, c)
        assert sp == 0 and sm == 0  # the active set does not change within h
        fd = (g @ cp - g @ cm) / (2 * h)
        assert fd == pytest.approx(gs[0] @ d, rel=1e-6, abs=1e-9 * np.abs(gs[0]).sum() * np.abs(d).max())
    h = rel * alpha
    cp, sp = orc.lasso_polish(A, s, alpha + h, c)
    cm, sm = orc.lasso_polish(A, s, alpha - h, c)
    assert sp == 0 and sm == 0
    fd = (g @ cp - g @ cm) / (2 * h)
    assert fd == pytest.approx(ga[0], rel=1e-6)
    return c


def test_orthogonal_design_closed_form(orc):
    rng = np.random.default_rng(11)
    N, K = 40, 12
    Q, _ = np.linalg.qr(rng.standard_normal((N, K)))
    A = Q * np.sqrt(N)  # A^T A = N I
    s = rng.standard_normal(N) * 3
    alpha = 0.4
    u = A.T @ s / N
    c = np.sign(u) * np.maximum(np.abs(u) - alpha, 0)
    assert 0 < np.count_nonzero(c) < K
    g = rng.standard_normal(K)
    gs, ga = orc.lasso_adjoint(A, c[None], g[None])
    S = c != 0
    np.testing.assert_allclose(gs[0], A[:, S] @ g[S] / N, rtol=1e-12, atol=1e-14)
    assert ga[0] == pytest.approx(-np.sum(np.sign(c[S]) * g[S]), rel=1e-12)


def test_finite_differences_random_design(orc):
    rng = np.random.default_rng(12)
    N, K = 60, 20
    A = rng.standard_normal((N, K))
    s = A @ (rng.standard_normal(K) * (rng.random(K) < 0.5)) + 0.1 * rng.standard_normal(N)
    c = _fd_check(orc, A, s, 0.05, rng)
    assert 0 < np.count_nonzero(c) < K


def test_finite_differences_shore_design(orc):
    rng = np.random.default_rng(13)
    gt = W.cartesian_table(100, seed=1)
    A = orc.design_matrix(4, 700.0, gt.bvals, gt.bvecs, gt.big_delta, gt.small_delta)
    s = W.multi_tensor_signal(gt, 1, 1000.0 / 30.0, seed=2)[0].astype(np.float64)
    c = _fd_check(orc, A, s, 1e-3, rng)
    assert 0 < np.count_nonzero(c) <= A.shape[1]


def test_polish_detects_wrong_active_sets(orc):
    rng = np.random.default_rng(14)
    N, K = 60, 20
    A = rng.standard_normal((N, K))
    s = A @ (rng.standard_normal(K) * (rng.random(K) < 0.5)) + 0.1 * rng.standard_normal(N)
    c0, _ = orc.fista(A, s, 0.05, max_iter=200000, tol=1e-13)
    c, st = orc.lasso_polish(A, s, 0.05, c0)
    assert st == 0 and np.abs(c - c0).max() <= 1e-9 * np.abs(c0).max()
    on = np.flatnonzero(c0)
    bad = c0.copy()
    bad[on[0]] = -bad[on[0]]
    assert orc.lasso_polish(A, s, 0.05, bad)[1] in (1, 2)
    drop = c0.copy()
    drop[on[np.argmax(np.abs(c0[on]))]] = 0.0
    assert orc.lasso_polish(A, s, 0.05, drop)[1] in (1, 2)


def test_empty_active_set(orc):
    A = np.eye(5)
    gs, ga = orc.lasso_adjoint(A, np.zeros((1, 5)), np.ones((1, 5)))
    assert not gs.any() and ga[0] == 0.0
