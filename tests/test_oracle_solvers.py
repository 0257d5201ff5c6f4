"""Pins the oracle's FISTA, Tikhonov, Lipschitz and alpha machinery.

Independent checks: the reference's committed sklearn fixtures
(tests/optimization_sklearn.rs:139-205), KKT optimality
(tests/optimization_kkt.rs), closed-form limits (tests/optimization_limits.rs),
numpy.linalg (SVD, lstsq, solve), scipy.linalg.cho_solve, and the unit KATs of
src/solver/{fista,tikhonov,alpha}.rs.
"""

import math

import numpy as np
import pytest
from scipy import linalg as sla

from tests.fixtures import (kkt_lasso, kkt_nonneg_lasso, lasso_fixtures, lasso_objective, make_design,
                            make_ill_conditioned, make_noisy_signal, make_planted_beta, shell_gtab)

FIXTURES = lasso_fixtures()


@pytest.mark.parametrize("fx", FIXTURES, ids=[f["name"] for f in FIXTURES])
def test_fista_matches_sklearn_fixture(orc, fx):
    coef, _ = orc.fista(fx["A"], fx["y"], fx["alpha"], 200_000, 1e-12, fx["positive"])
    rel = np.linalg.norm(coef - fx["beta"]) / max(np.linalg.norm(fx["beta"]), 1.0)
    assert rel < 1e-3
    obj = lasso_objective(fx["A"], fx["y"], coef, fx["alpha"])
    assert abs(obj - fx["objective"]) / max(abs(fx["objective"]), 1.0) < 1e-6
    kkt = (kkt_nonneg_lasso if fx["positive"] else kkt_lasso)(fx["A"], fx["y"], coef, fx["alpha"])
    assert max(kkt) < 1e-5


def test_fixture_alpha_max_header(orc):
    for fx in FIXTURES:
        assert orc.alpha_max(fx["A"], fx["y"]) == pytest.approx(fx["alpha_max"], rel=1e-12)


@pytest.mark.parametrize("m,n,kind,tol", [(200, 50, "well", 1e-6), (30, 80, "under", 1e-6), (400, 20, "over", 1e-6),
                                          (150, 40, "ill", 1e-3)])
def test_fista_kkt(orc, m, n, kind, tol):
    A = make_ill_conditioned(m, n, 1e4, 0xB1) if kind == "ill" else make_design(m, n, 0xA1 + m)
    beta = make_planted_beta(n, 6, True, 0xA2 + n)
    y = make_noisy_signal(A, beta, 0.05, 0xA3)
    coef, _ = orc.fista(A, y, 0.05, 100_000, 1e-10, False)
    assert max(kkt_lasso(A, y, coef, 0.05)) < tol


def test_nonneg_fista_kkt(orc):
    A = make_design(200, 50, 0x11)
    y = make_noisy_signal(A, make_planted_beta(50, 8, True, 0x12), 0.05, 0x13)
    coef, _ = orc.fista(A, y, 0.05, 100_000, 1e-10, True)
    assert coef.min() >= -1e-12
    assert max(kkt_nonneg_lasso(A, y, coef, 0.05)) < 1e-5


def test_alpha_at_and_above_alpha_max_is_exact_zero(orc):
    A = make_design(150, 40, 0xB1)
    y = make_noisy_signal(A, make_planted_beta(40, 5, True, 0xB2), 0.05, 0xB3)
    amax = orc.alpha_max(A, y)
    assert amax == pytest.approx(np.abs(A.T @ y).max() / 150, rel=1e-14)
    for a in (amax, 10 * amax):
        coef, d = orc.fista(A, y, a, 200_000, 1e-12, False)
        assert np.abs(coef).max() < 1e-12
    coef, _ = orc.fista(A, y, 0.5 * amax, 2000, 1e-9, False)
    assert (coef != 0).sum() > 0


def test_alpha_zero_limits(orc):
    # tests/optimization_limits.rs:36-84
    A = make_design(200, 20, 0xA1)
    y = make_noisy_signal(A, make_planted_beta(20, 6, True, 0xA2), 0.05, 0xA3)
    ols = np.linalg.lstsq(A, y, rcond=None)[0]
    coef, _ = orc.fista(A, y, 1e-12, 200_000, 1e-12, False)
    assert np.abs(coef - ols).max() < 1e-3
    rng = np.random.default_rng(0xC1)
    U, _, Vt = np.linalg.svd(rng.standard_normal((25, 25)))
    M = (U * (1.0 + 0.1 * np.arange(25) / 25)) @ Vt
    yv = rng.uniform(-1, 1, 25)
    coef, _ = orc.fista(M, yv, 1e-12, 200_000, 1e-12, False)
    assert np.abs(coef - np.linalg.solve(M, yv)).max() < 1e-4


def test_fista_unit_kats(orc):
    # src/solver/fista.rs:188-230
    m = np.array([1.0, 0.0, 0.5, -0.2, 0.0, 1.0, -0.1, 0.3, -0.3, 0.4, 1.0, 0.0, 0.2, -0.1, 0.0, 1.0,
                  0.5, 0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.5, 0.5, -0.5, 0.5, -0.5, 0.1, 0.2, -0.3, 0.4]).reshape(8, 4)
    s = m @ np.array([2.0, 0.0, -1.5, 0.0])
    c, _ = orc.fista(m, s, 1e-3, 2000, 1e-9, False)
    assert abs(c[0] - 2.0) < 0.05 and abs(c[1]) < 0.05 and abs(c[2] + 1.5) < 0.05 and abs(c[3]) < 0.05
    m = np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 1.0], [1.0, -1.0]])
    c, _ = orc.fista(m, m @ np.array([2.0, -3.0]), 0.0, 5000, 1e-12, False)
    assert abs(c[0] - 2.0) < 1e-3 and abs(c[1] + 3.0) < 1e-3


def test_lipschitz_matches_svd(orc):
    for m, n, seed in ((50, 20, 1), (30, 80, 2), (258, 145, 3)):
        A = make_design(m, n, seed)
        s = np.linalg.svd(A, compute_uv=False)[0]
        assert orc.lipschitz(A) == pytest.approx(s * s / m, rel=1e-12)
    assert orc.lipschitz(np.zeros((4, 3))) == np.finfo(float).eps  # floored at f64::EPSILON


def _fista_textbook(A, s, alpha, iters, L):
    """Beck & Teboulle (2009) FISTA in vector form, fixed iteration count."""
    n = A.shape[0]
    x = np.zeros(A.shape[1])
    y = x.copy()
    t = 1.0
    for _ in range(iters):
        v = y - (A.T @ (A @ y - s) / n) / L
        xn = np.sign(v) * np.maximum(np.abs(v) - alpha / L, 0.0)
        tn = 0.5 * (1 + math.sqrt(1 + 4 * t * t))
        y = xn + (t - 1) / tn * (xn - x)
        x, t = xn, tn
    return x


def test_fista_trajectory_matches_textbook(orc):
    A = make_design(60, 30, 7)
    s = make_noisy_signal(A, make_planted_beta(30, 5, True, 8), 0.05, 9)
    L = orc.lipschitz(A)
    for iters in (1, 2, 7, 50):
        c, d = orc.fista(A, s, 0.01, iters, 0.0, False, L)
        assert d["iterations"] == iters and not d["converged"]
        np.testing.assert_allclose(c, _fista_textbook(A, s, 0.01, iters, L), rtol=1e-10, atol=1e-12)
        assert d["residual_l2"] == pytest.approx(np.linalg.norm(s - A @ c), rel=1e-12)


def test_fista_stopping_rule(orc):
    A = make_design(60, 30, 7)
    s = make_noisy_signal(A, make_planted_beta(30, 5, True, 8), 0.05, 9)
    c, d = orc.fista(A, s, 0.01, 10_000, 1e-6, False)
    assert d["converged"] and d["iterations"] < 10_000
    # one fewer iteration must not satisfy the rule: rel change of the last step
    L = orc.lipschitz(A)
    prev, _ = orc.fista(A, s, 0.01, d["iterations"] - 1, 0.0, False, L)
    rel = np.linalg.norm(c - prev) / max(np.linalg.norm(c), 1e-12)
    assert rel < 1e-6
    # zero signal: converges at the first iteration with exact zeros
    c, d = orc.fista(A, np.zeros(60), 0.01, 100, 1e-6, False)
    assert d["iterations"] == 1 and d["converged"] and not c.any()


def test_fista_warm_start(orc):
    A = make_design(60, 30, 7)
    s = make_noisy_signal(A, make_planted_beta(30, 5, True, 8), 0.05, 9)
    c, d = orc.fista(A, s, 0.01, 100_000, 1e-10, False)
    c2, d2 = orc.fista(A, s, 0.01, 100_000, 1e-10, False, warm=c)
    assert d2["iterations"] < 5
    np.testing.assert_allclose(c2, c, atol=1e-8)


def test_tikhonov_normal_equations_on_shore_design(orc):
    # tests/optimization_kkt.rs:164-218
    for per_shell, lam, seed, bound in ((60, 1e-2, 0x21, 1e-10), (80, 1e-8, 0x31, 1e-9)):
        bvals, bvecs = shell_gtab(per_shell)
        A = orc.design_matrix(6, 700.0, bvals, bvecs, 0.04, 0.012)
        dn, dl = orc.regularization(6)
        y = make_noisy_signal(A, make_planted_beta(A.shape[1], 4, True, seed), 0.01, seed + 1)
        c, r = orc.tikhonov(A, y, dn, dl, lam, lam)
        lhs = (A.T @ A + lam * np.diag(dn) + lam * np.diag(dl)) @ c
        rhs = A.T @ y
        assert np.linalg.norm(lhs - rhs) / np.linalg.norm(rhs) < bound
        ref = sla.cho_solve(sla.cho_factor(A.T @ A + lam * np.diag(dn + dl)), rhs)
        np.testing.assert_allclose(c, ref, rtol=1e-6, atol=1e-9 * np.abs(ref).max())
        assert r == pytest.approx(np.linalg.norm(y - A @ c), rel=1e-10)


def test_tikhonov_normal_equations_order8_cartesian(orc):
    # the whole-brain L2 leg: order 8 (K = 145) on 1 b0 + 257 Cartesian
    # q-points, lambda_n = lambda_l = 1e-2; the reference's 1e-10 bound
    # (tests/optimization_kkt.rs:164-190) and scipy's cho_solve
    from cs_dmri_amd import workloads as W
    gt = W.cartesian_table(257, seed=3)
    A = orc.design_matrix(8, 700.0, gt.bvals, gt.bvecs, gt.big_delta, gt.small_delta)
    dn, dl = orc.regularization(8)
    assert A.shape == (258, 145)
    S = W.multi_tensor_signal(gt, 8, W.S0 / 20.0, seed=2).astype(np.float64)
    C, R = orc.tikhonov(A, S, dn, dl, 1e-2, 1e-2)
    Greg = A.T @ A + 1e-2 * np.diag(dn + dl)
    fac = sla.cho_factor(Greg)
    for s, c, r in zip(S, C, R):
        rhs = A.T @ s
        assert np.linalg.norm(Greg @ c - rhs) / np.linalg.norm(rhs) < 1e-10
        np.testing.assert_allclose(c, sla.cho_solve(fac, rhs), rtol=1e-6, atol=1e-9 * np.abs(c).max())
        assert r == pytest.approx(np.linalg.norm(s - A @ c), rel=1e-10)


def test_tikhonov_limits_and_not_pd(orc):
    A = make_design(200, 30, 0xD1)
    y = make_noisy_signal(A, make_planted_beta(30, 6, True, 0xD2), 0.05, 0xD3)
    c, _ = orc.tikhonov(A, y, np.ones(30), None, 1e-14, 0.0)
    assert np.abs(c - np.linalg.lstsq(A, y, rcond=None)[0]).max() < 1e-6
    c, _ = orc.tikhonov(A, y, np.ones(30), None, 1e10, 0.0)
    assert np.abs(c).max() < 1e-6
    with pytest.raises(ValueError):
        orc.tikhonov(make_design(10, 20, 1), np.ones(10))  # rank-deficient, no regularisation
    # src/solver/tikhonov.rs:85-106
    m = np.array([[1.0, 0.0], [0.0, 1.0], [1.0, 1.0], [1.0, -1.0]])
    c, _ = orc.tikhonov(m, m @ np.array([2.0, -3.0]), np.zeros(2), None, 1e-12, 0.0)
    np.testing.assert_allclose(c, [2.0, -3.0], atol=1e-6)


def test_regularization_diagonals(orc):
    dn, dl = orc.regularization(4)
    idx = orc.shore_indices(4)
    np.testing.assert_array_equal(dn, (idx[:, 0] * (idx[:, 0] + 1.0)) ** 2)
    np.testing.assert_array_equal(dl, (idx[:, 1] * (idx[:, 1] + 1.0)) ** 2)


def test_alpha_path_and_bic(orc):
    p = orc.log_path(10.0, 5, 1e-3)
    assert p[0] == pytest.approx(10.0, abs=1e-12) and p[-1] == pytest.approx(1e-2, abs=1e-12)
    assert np.all(np.diff(p) < 0)
    np.testing.assert_allclose(p, np.geomspace(10.0, 1e-2, 5), rtol=1e-13)
    assert orc.bic(100, 2.5, 7) == pytest.approx(100 * math.log(2.5 / 100) + math.log(100) * 7, rel=1e-14)
    assert np.isfinite(orc.bic(100, 0.0, 3))  # RSS clamped at f64::MIN_POSITIVE
