"""GPU parity: the HIP path (through the C ABI) against the FP64 oracle.

Tolerances
----------
* design matrix (FP64 on the device): 1e-12 relative to max|A|.
* Lipschitz constant (FP64 power iteration vs the oracle's Jacobi
  eigen-solve): 1e-13 relative.
* FISTA, FP32 plain form, equal iteration counts (tol = 0): per-voxel
  coefficient error |c_gpu - c_ref| / |c_ref| <= 1e-5 (north-star bound for
  the FP32 path).
* FISTA with the stopping rule (tol = 1e-6): FP32 rounding moves the
  iteration at which |dx|/|x| first drops below tol, so the bound is on the
  distribution: median <= 1e-5, 99th percentile <= 1e-3, and the per-voxel
  iteration counts agree within 10% for 95% of voxels.
* Tikhonov (FP64 factorisation, FP32 application): 1e-4 relative per voxel.
* sklearn fixtures and the KKT problems: the reference's own bounds
  (tests/optimization_sklearn.rs:159-202, tests/optimization_kkt.rs:39-156).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import cs_dmri_amd as sf  # noqa: E402
from cs_dmri_amd import workloads as W  # noqa: E402
from tests.fixtures import (kkt_lasso, kkt_nonneg_lasso, lasso_fixtures, lasso_objective, make_design,  # noqa: E402
                            make_ill_conditioned, make_noisy_signal, make_planted_beta, shell_gtab)


def _rel(c, ref):
    return np.linalg.norm(c - ref, axis=-1) / np.maximum(np.linalg.norm(ref, axis=-1), 1e-30)


@pytest.fixture(scope="module")
def cfg1(orc):
    wl = W.config1()
    g = wl.gtab
    A = orc.design_matrix(6, 700.0, g.bvals, g.bvecs, g.big_delta, g.small_delta)
    return wl, A


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10])
def test_design_matrix_matches_oracle(orc, order):
    for bvals, bvecs, big, small in [(*shell_gtab(20), 0.04, 0.012),
                                     (W.cartesian_table(257, 3).bvals, W.cartesian_table(257, 3).bvecs,
                                      W.TAU_CARTESIAN, 0.0)]:
        A = sf.design_matrix(bvals, bvecs, radial_order=order, big_delta=big, small_delta=small)
        ref = orc.design_matrix(order, 700.0, bvals, bvecs, big, small)
        assert np.abs(A - ref).max() <= 1e-12 * np.abs(ref).max()


def test_lipschitz_matches_oracle(orc, cfg1):
    wl, A = cfg1
    _, d = sf.fista_batch(A, wl.data[wl.mask][:16], 1e-3, max_iter=2)
    assert d["lipschitz"] == pytest.approx(orc.lipschitz(A), rel=1e-13)


@pytest.mark.parametrize("iters", [1, 7, 50, 500])
def test_fista_trajectory_equal_iterations(orc, cfg1, iters):
    wl, A = cfg1
    S = wl.data[wl.mask]
    V = 2048 if iters <= 50 else 512
    coef, d = sf.fista_batch(A, S[:V], 1e-3, max_iter=iters, tol=0.0)
    assert (d["iterations"] == iters).all() and not d["converged"].any()
    sub = S[:V].reshape(V, 1, 1, -1)
    ref = orc.fit_volume(sub, None, A, alpha_mode="fixed", alpha=1e-3, max_iter=iters, tol=0.0)
    cref = ref["coefficients"].reshape(V, -1).astype(np.float64)
    err = _rel(coef.astype(np.float64), cref)
    assert err.max() <= 1e-5, (err.max(), np.median(err))
    np.testing.assert_allclose(d["residual_l2"], ref["residual_l2"].ravel(), rtol=1e-4)


def test_fista_stopping_rule_config1(orc, cfg1):
    wl, A = cfg1
    out = sf.shore_fit(wl.data, wl.gtab.bvals, wl.gtab.bvecs, wl.mask, big_delta=wl.gtab.big_delta,
                       small_delta=wl.gtab.small_delta, radial_order=6, alpha_mode="fixed", alpha=1e-3,
                       max_iter=500, tol=1e-6, diagnostics=True)
    ref = orc.fit_volume(wl.data, wl.mask, A, alpha_mode="fixed", alpha=1e-3, max_iter=500, tol=1e-6)
    c = out["coefficients"][wl.mask].astype(np.float64)
    cr = ref["coefficients"][wl.mask].astype(np.float64)
    err = _rel(c, cr)
    assert np.median(err) <= 1e-5 and np.quantile(err, 0.99) <= 1e-3, (np.median(err), np.quantile(err, 0.99))
    it, itr = out["iterations"][wl.mask].astype(float), ref["iterations"][wl.mask].astype(float)
    assert np.mean(np.abs(it - itr) <= 0.1 * itr) >= 0.95
    np.testing.assert_allclose(out["r2"][wl.mask], ref["r2"][wl.mask], rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(out["residual_l2"][wl.mask], ref["residual_l2"][wl.mask], rtol=1e-3)
    assert (out["regularization_kind"][wl.mask] == 1).all()
    assert np.all(out["alpha"][wl.mask] == np.float32(1e-3))
    # The relative-change rule stops short of the exact minimiser in either
    # arithmetic, so the device's KKT residual is held to the FP64 oracle's on
    # the same voxel (plus the FP32 floor), not to an absolute bound.
    S = wl.data[wl.mask].astype(np.float64)
    worst = 0.0
    for v in range(0, len(S), 31):
        scale = np.abs(A.T @ S[v]).max() / len(S[v])
        g = max(kkt_lasso(A, S[v], c[v], 1e-3))
        r = max(kkt_lasso(A, S[v], cr[v], 1e-3))
        worst = max(worst, (g - r) / scale)
        assert g <= 2.0 * r + 1e-6 * scale, (v, g, r, scale)
    print(f"worst (device - oracle) KKT residual / |A^T s|_inf/N: {worst:.3e}")


FIXTURES = lasso_fixtures()


@pytest.mark.parametrize("fx", FIXTURES, ids=[f["name"] for f in FIXTURES])
def test_fista_sklearn_fixtures(fx):
    """The reference's settings and bounds (tests/optimization_sklearn.rs:159-202):
    max_iter 200000, tol 1e-12, coefficients 1e-3, objective 1e-6, KKT 1e-5.
    The FP32 device solve lands at objective gaps of 1e-15 and KKT residuals
    of at most 1.1e-8 on these fixtures."""
    coef, d = sf.fista_batch(fx["A"], fx["y"][None], fx["alpha"], max_iter=200_000, tol=1e-12,
                             non_negative=fx["positive"])
    c = coef[0].astype(np.float64)
    rel = np.linalg.norm(c - fx["beta"]) / max(np.linalg.norm(fx["beta"]), 1.0)
    assert rel < 1e-3
    obj = lasso_objective(fx["A"], fx["y"], c, fx["alpha"])
    assert abs(obj - fx["objective"]) / max(abs(fx["objective"]), 1.0) < 1e-6
    active, inactive = (kkt_nonneg_lasso if fx["positive"] else kkt_lasso)(fx["A"], fx["y"], c, fx["alpha"])
    assert active < 1e-5 and inactive < 1e-5, (active, inactive)


def _kkt_case(kind):
    """The problems of tests/optimization_kkt.rs:39-156 (numpy's generator in
    place of ChaCha8, so the draws differ; shapes, kappa, sparsity, noise and
    alpha as there)."""
    if kind == "well":
        A, beta, seed, nn = make_design(200, 50, 0xA1), make_planted_beta(50, 8, True, 0xA2), 0xA3, False
    elif kind == "ill":
        A, beta, seed, nn = make_ill_conditioned(150, 40, 1e4, 0xB1), make_planted_beta(40, 6, True, 0xB2), 0xB3, False
    elif kind == "under":
        A, beta, seed, nn = make_design(30, 80, 0xC1), make_planted_beta(80, 5, True, 0xC2), 0xC3, False
    elif kind == "over":
        A, beta, seed, nn = make_design(400, 20, 0xD1), make_planted_beta(20, 4, True, 0xD2), 0xD3, False
    elif kind == "nonneg":
        A, beta, seed, nn = make_design(180, 45, 0xF1), make_planted_beta(45, 7, False, 0xF2), 0xF3, True
    else:  # nonneg fit of a mixed-sign truth: the constraint binds
        A, beta, seed, nn = make_design(200, 50, 0x11), make_planted_beta(50, 8, True, 0x12), 0x13, True
    return A, make_noisy_signal(A, beta, 0.05, seed), nn


# (case, KKT bound) as in tests/optimization_kkt.rs:39-156
KKT_CASES = [("well", 1e-6), ("ill", 1e-3), ("under", 1e-6), ("over", 1e-6), ("nonneg", 1e-6),
             ("nonneg-binding", 1e-5)]


@pytest.mark.parametrize("kind,bound", KKT_CASES, ids=[k for k, _ in KKT_CASES])
@pytest.mark.parametrize("alpha_rel", [None, 0.1], ids=["alpha0.05", "alpha0.1max"])
def test_fista_kkt_cases(kind, bound, alpha_rel):
    """KKT residuals of the device solve, at the reference's alpha = 0.05 and,
    because with N(0, 1/m) designs 0.05 is above alpha_max for most of these
    draws (a zero solution), also at 0.1 * alpha_max, where the solution has
    an active set. Non-negative fits never go below zero."""
    A, y, nn = _kkt_case(kind)
    amax = np.abs(A.T @ y).max() / A.shape[0]
    alpha = 0.05 if alpha_rel is None else alpha_rel * amax
    coef, d = sf.fista_batch(A, y[None], alpha, max_iter=100_000, tol=1e-10, non_negative=nn)
    c = coef[0].astype(np.float64)
    if alpha_rel is not None:
        assert np.count_nonzero(c) > 0
    if nn:
        assert c.min() >= 0.0
    active, inactive = (kkt_nonneg_lasso if nn else kkt_lasso)(A, y, c, alpha)
    assert active < bound and inactive < bound, (active, inactive, int(d["iterations"][0]))


def test_fista_at_alpha_max_is_exact_zero():
    """tests/optimization_kkt.rs:106-122: at alpha = alpha_max the solution is 0."""
    A = make_design(120, 35, 0xE1)
    y = make_noisy_signal(A, make_planted_beta(35, 5, True, 0xE2), 0.05, 0xE3)
    amax = np.abs(A.T @ y).max() / A.shape[0]
    coef, _ = sf.fista_batch(A, y[None], amax, max_iter=100_000, tol=1e-10)
    assert np.abs(coef).max() < 1e-12


def _small_volume():
    gt = W.cartesian_table(60, seed=21)
    shape = (5, 4, 3)
    mask = np.ones(shape, bool)
    mask[0, 0, 0] = mask[4, 3, 2] = mask[2, 1, 1] = False
    data = np.zeros(shape + (len(gt.bvals),), np.float32)
    data[mask] = W.multi_tensor_signal(gt, int(mask.sum()), W.S0 / 30, seed=22)
    data[1, 2, 0] = 0.0  # all-zero voxel inside the mask
    data[~mask] = 7.0    # garbage outside the mask must not leak into the fit
    return gt, data, mask


@pytest.mark.parametrize("mode", ["fixed", "alpha-ratio"])
def test_volume_layout_and_diagnostics(orc, mode):
    gt, data, mask = _small_volume()
    kw = dict(alpha_mode=mode, alpha=1e-3, alpha_ratio=1e-2, max_iter=300, tol=0.0)
    out = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, big_delta=gt.big_delta, small_delta=gt.small_delta,
                       radial_order=4, diagnostics=True, **kw)
    A = orc.design_matrix(4, 700.0, gt.bvals, gt.bvecs, gt.big_delta, gt.small_delta)
    ref = orc.fit_volume(data, mask, A, **kw)
    assert out["coefficients"].shape == mask.shape + (29,) and out["coefficients"].dtype == np.float32
    assert not out["coefficients"][~mask].any()
    for k in ("r2", "residual_l2", "iterations", "regularization_kind", "alpha", "bic"):
        assert not out[k][~mask].any(), k
    err = _rel(out["coefficients"][mask].astype(float), ref["coefficients"][mask].astype(float))
    nz = np.linalg.norm(ref["coefficients"][mask], axis=-1) > 0
    assert err[nz].max() <= 1e-5
    np.testing.assert_array_equal(out["iterations"], ref["iterations"])
    np.testing.assert_allclose(out["alpha"], ref["alpha"], rtol=1e-6)
    # BIC from the device's own coefficients and residual (src/solver/alpha.rs:99-103)
    n = len(gt.bvals)
    nnz = (out["coefficients"] != 0).sum(-1)
    rss = np.maximum(out["residual_l2"].astype(np.float64) ** 2, np.finfo(float).tiny)
    live = mask & (out["alpha"] > 0)
    np.testing.assert_allclose(out["bic"][live], (n * np.log(rss / n) + np.log(n) * nnz)[live], rtol=1e-5, atol=1e-4)
    np.testing.assert_allclose(out["residual_l2"][mask], ref["residual_l2"][mask], rtol=1e-4)
    np.testing.assert_allclose(out["rmse"], out["residual_l2"] / np.sqrt(np.float32(len(gt.bvals))), rtol=1e-6)
    if mode == "alpha-ratio":
        assert out["alpha_distribution"] is not None
        np.testing.assert_allclose(out["alpha_distribution"], ref["alpha_distribution"], rtol=1e-6)
        assert out["alpha"][1, 2, 0] == 0 and out["iterations"][1, 2, 0] == 0
        assert out["sidecar"]["solver"]["chosen_alpha"]["scope"] == "per-voxel"
    else:
        assert out["alpha_distribution"] is None
        assert out["sidecar"]["solver"]["chosen_alpha"] == {"scope": "global", "alpha": 1e-3}


def test_tikhonov_volume(orc):
    gt, data, mask = _small_volume()
    out = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, big_delta=gt.big_delta, small_delta=gt.small_delta,
                       radial_order=4, regularization="l2", lambda_n=1e-2, lambda_l=1e-2, diagnostics=True)
    A = orc.design_matrix(4, 700.0, gt.bvals, gt.bvecs, gt.big_delta, gt.small_delta)
    dn, dl = orc.regularization(4)
    ref = orc.fit_volume(data, mask, A, regularization="l2", dn=dn, dl=dl, lambda_n=1e-2, lambda_l=1e-2)
    c, cr = out["coefficients"][mask].astype(float), ref["coefficients"][mask].astype(float)
    nz = np.linalg.norm(cr, axis=-1) > 0
    assert _rel(c[nz], cr[nz]).max() <= 1e-4
    assert not out["coefficients"][~mask].any()
    assert (out["iterations"][mask] == 1).all() and (out["regularization_kind"][mask] == 2).all()
    assert "alpha" not in out and out["sidecar"]["solver"]["kind"] == "tikhonov"


def test_tikhonov_batch_normal_equations(orc):
    bvals, bvecs = shell_gtab(60)
    A = orc.design_matrix(6, 700.0, bvals, bvecs, 0.04, 0.012)
    dn, dl = orc.regularization(6)
    rng = np.random.default_rng(5)
    S = (A @ rng.standard_normal((A.shape[1], 64))).T * 100 + rng.standard_normal((64, len(bvals)))
    C, R = sf.tikhonov_batch(A, S, dn, dl, 1e-2, 1e-2)
    Cr, Rr = orc.tikhonov(A, S.astype(np.float32).astype(np.float64), dn, dl, 1e-2, 1e-2)
    assert _rel(C.astype(float), Cr).max() <= 1e-4
    np.testing.assert_allclose(R, Rr, rtol=1e-3)
    with pytest.raises(ValueError, match="PD"):
        sf.tikhonov_batch(A[:20], S[:, :20])  # N < K without regularisation: not positive definite


def test_empty_mask_and_determinism():
    gt, data, mask = _small_volume()
    kw = dict(big_delta=gt.big_delta, small_delta=gt.small_delta, radial_order=4, alpha_mode="fixed", alpha=1e-3,
              max_iter=200, diagnostics=True)
    empty = sf.shore_fit(data, gt.bvals, gt.bvecs, np.zeros(mask.shape, bool), **kw)
    assert not empty["coefficients"].any() and not empty["iterations"].any()
    a = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, **kw)
    b = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, **kw)
    for k in ("coefficients", "iterations", "residual_l2", "r2", "bic"):
        np.testing.assert_array_equal(a[k], b[k])
    one = np.zeros(mask.shape, bool)
    one[3, 2, 1] = True
    s1 = sf.shore_fit(data, gt.bvals, gt.bvecs, one, **kw)
    np.testing.assert_array_equal(s1["coefficients"][3, 2, 1], a["coefficients"][3, 2, 1])


def test_shore_model_front_end(tmp_path):
    # ShoreModel.fit is shore_fit on the same arrays (python/cs_dmri/shore.py:80-95)
    gt, data, mask = _small_volume()
    gtab = sf.GradientTable(gt.bvals, gt.bvecs, big_delta=gt.big_delta, small_delta=gt.small_delta)
    kw = dict(radial_order=4, alpha_mode="fixed", alpha=1e-3, max_iter=200)
    fit = sf.ShoreModel(gtab, **kw).fit(data, mask)
    ref = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, big_delta=gt.big_delta, small_delta=gt.small_delta,
                       diagnostics=True, **kw)
    np.testing.assert_array_equal(fit.coefficients, ref["coefficients"])
    for k in ("r2", "rmse", "residual_l2", "iterations", "regularization_kind", "alpha", "bic"):
        np.testing.assert_array_equal(getattr(fit, k), ref[k], err_msg=k)
    assert fit.rss_l2 is None and fit.frame == "image" and fit.radial_order == 4
    assert fit.sidecar == ref["sidecar"] and fit.sidecar["bvec_frame"] == "image-axis"
    np.testing.assert_allclose(fit.sparsity[mask], (ref["coefficients"][mask] != 0).mean(-1), rtol=1e-6)
    assert "ShoreFit" in repr(fit) and f"voxels={int(mask.sum())}" in repr(fit)
    # world frame: the bvecs go through the affine's rotation
    aff = np.diag([-2.0, 2.0, 2.0, 1.0])
    fw = sf.ShoreModel(gtab, **kw).fit(data, mask, affine=aff)
    rw = sf.shore_fit(data, gt.bvals, gt.bvecs * np.array([-1.0, 1.0, 1.0]), mask, big_delta=gt.big_delta,
                      small_delta=gt.small_delta, bvec_frame="world-ras", diagnostics=True, **kw)
    assert fw.frame == "world" and fw.sidecar["bvec_frame"] == "world-ras"
    np.testing.assert_array_equal(fw.coefficients, rw["coefficients"])
    # without a mask an array gets the b0 fallback mask (src/io/dwi.rs:347-365)
    from cs_dmri_amd.io import default_mask_from_b0
    fb = sf.ShoreModel(gtab, **kw).fit(data)
    np.testing.assert_array_equal(fb.mask, default_mask_from_b0(data, gt.bvals))
    # save / load round trip in the cs-fit layout; existing files are refused first
    p = tmp_path / "fit.nii.gz"
    fw.save(p)
    for suffix in ("_r2", "_rmse", "_alpha", "_bic", "_sparsity"):
        assert (tmp_path / f"fit{suffix}.nii.gz").exists(), suffix
    assert (tmp_path / "fit.json").exists() and not (tmp_path / "fit_rss_l2.nii.gz").exists()
    with pytest.raises(FileExistsError):
        fw.save(p)
    back = sf.ShoreFit.load(p, mask=mask)
    np.testing.assert_array_equal(back.coefficients, fw.coefficients)
    np.testing.assert_array_equal(back.r2, fw.r2)
    np.testing.assert_array_equal(back.bic, fw.bic)
    assert back.frame == "world" and back.sidecar == fw.sidecar and back.iterations is None
    np.testing.assert_allclose(back.affine, aff)
    np.testing.assert_array_equal(back.odf_sh(), fw.odf_sh())


def test_tortoise_deltas_through_the_library(orc):
    """Deltas left out: the library's host code estimates them with TORTOISE's
    heuristic (src/qspace.rs:100-112, 209-227), the same numbers as the
    oracle (whose estimate the CPU tests pin by b = 2 gamma^2 G^2 delta^3),
    reports the source, and fits with them."""
    gt, data, mask = _small_volume()
    out = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, radial_order=4, alpha_mode="fixed", alpha=1e-3,
                       max_iter=200, tol=0.0)
    meta = out["sidecar"]
    big, small = orc.estimate_deltas_tortoise(gt.bvals, 80e-3)
    assert meta["delta_source"] == "estimated_tortoise"
    assert meta["big_delta_seconds"] == pytest.approx(big, rel=1e-15)
    assert meta["small_delta_seconds"] == pytest.approx(small, rel=1e-15)
    assert meta["tau_seconds"] == pytest.approx(big - small / 3.0, rel=1e-14)
    A = orc.design_matrix(4, 700.0, gt.bvals, gt.bvecs, big, small)
    ref = orc.fit_volume(data, mask, A, alpha_mode="fixed", alpha=1e-3, max_iter=200, tol=0.0)
    # FP32 against FP64 over 200 iterations of this design (tau from the
    # estimated deltas): measured median 2e-6, max 1.9e-5
    err = _rel(out["coefficients"][mask].astype(float), ref["coefficients"][mask].astype(float))
    nz = np.linalg.norm(ref["coefficients"][mask], axis=-1) > 0
    assert np.median(err[nz]) <= 5e-6 and err[nz].max() <= 5e-5
