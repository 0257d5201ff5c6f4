"""GPU parity of the gradient-deviation (per-voxel design) fit.

Each deviated voxel's design A_v is built on the device from its effective
table (src/graddev.rs:160-185); the FP64 oracle builds the same table and
design on the CPU and runs the reference FISTA. Tolerances:

* set-up seam: G_v = A_v^T A_v (FP64 accumulation stored as FP32) to 2e-7
  relative to max|G_v| per entry; b_v = A_v^T s (FP64) to 1e-10 relative;
  L = sigma_max(A_v)^2 / N from the device's FP32 Lanczos estimate to 2e-6
  relative of the oracle's Jacobi eigen-solve;
* equal iteration counts (tol = 0): per-voxel coefficient error <= 1e-5
  (the FP32 Gram-form iteration on a perturbation of A_v^T A_v of one FP32
  rounding);
* with the stopping rule (tol = 1e-6): median error <= 1e-5, 99th percentile
  <= 1e-3, iterations within 10% for >= 90% of voxels, residual_l2 rtol
  1e-3, BIC within 0.5 for >= 95% (an exact zero at the soft threshold can
  differ between FP32 and FP64);
* identity voxels (max |T - I| < 1e-6) follow the shared-design path and
  match the plain fit of the same voxels exactly;
* every tile-shape variant and every chunking gives the same fit to 1e-5.
"""

import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import cs_dmri_amd as sf  # noqa: E402
from cs_dmri_amd import workloads as W  # noqa: E402


def _rel(c, ref):
    return np.linalg.norm(c - ref, axis=-1) / np.maximum(np.linalg.norm(ref, axis=-1), 1e-30)


def _volume(shape, seed=0):
    gt = W.cartesian_table(257, seed=3)
    n = int(np.prod(shape))
    sig = W.mixed_snr_signals(gt, n, seed=5 + seed)
    data = sig.reshape(shape + (-1,)).astype(np.float32)
    field = W.graddev_tensors(n, seed=4 + seed).reshape(shape + (3, 3)).astype(np.float32)
    return gt, data, field


def _kw(gt):
    return dict(big_delta=gt.big_delta, small_delta=gt.small_delta)


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10])
def test_gram_seam_matches_oracle(orc, order):
    gt, data, field = _volume((2, 2, 2))
    vox = np.array([0, 3, 5, 7])
    G, b, lip = sf.graddev_gram(data, gt.bvals, gt.bvecs, field, vox, radial_order=order, **_kw(gt))
    flat = data.reshape(-1, data.shape[-1])
    T = field.reshape(-1, 3, 3).astype(np.float64)
    for i, v in enumerate(vox):
        be, d = orc.effective_table(T[v], np.eye(3), gt.bvals, gt.bvecs, gt.bvecs)
        A = orc.design_matrix(order, 700.0, be, d, gt.big_delta, gt.small_delta)
        Gr = A.T @ A
        assert np.abs(G[i] - Gr).max() <= 2e-7 * np.abs(Gr).max(), order
        br = A.T @ flat[v].astype(np.float64)
        assert np.abs(b[i] - br).max() <= 1e-10 * np.abs(br).max(), order
        assert lip[i] == pytest.approx(orc.lipschitz(A), rel=2e-6), order


def _oracle(orc, gt, data, field, mask, order, **kw):
    return orc.fit_graddev(data, mask, field, gt.bvals, gt.bvecs, radial_order=order, **_kw(gt), **kw)


@pytest.mark.parametrize("order,shape", [(6, (4, 4, 4)), (8, (4, 4, 3)), (10, (2, 2, 4))])
def test_equal_iterations(orc, order, shape):
    gt, data, field = _volume(shape)
    out = sf.shore_fit(data, gt.bvals, gt.bvecs, np.ones(shape, bool), radial_order=order, alpha_mode="fixed",
                       alpha=1e-3, max_iter=40, tol=0.0, diagnostics=True, graddev=field, **_kw(gt))
    ref = _oracle(orc, gt, data, field, None, order, max_iter=40, tol=0.0)
    assert (out["iterations"] == 40).all() and (ref["iterations"] == 40).all()
    err = _rel(out["coefficients"].reshape(-1, out["coefficients"].shape[-1]).astype(np.float64),
               ref["coefficients"].reshape(-1, ref["coefficients"].shape[-1]).astype(np.float64))
    print(f"order {order}: max coef err at 40 iterations {err.max():.2e}")
    assert err.max() <= 1e-5


@pytest.mark.parametrize("order,shape", [(6, (6, 4, 4)), (8, (4, 4, 4)), (10, (2, 2, 4))])
def test_stopping_rule_volume(orc, order, shape):
    gt, data, field = _volume(shape, seed=1)
    mask = np.ones(shape, bool)
    mask[0, 0, 0] = False
    mask[1, 0, 1] = False
    field[0, 1, 0] = np.eye(3, dtype=np.float32)                         # identity
    field[0, 1, 1] = np.eye(3, dtype=np.float32) + np.float32(4e-7)       # identity within 1e-6
    out = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, radial_order=order, alpha_mode="fixed", alpha=1e-3,
                       max_iter=500, tol=1e-6, diagnostics=True, graddev=field, **_kw(gt))
    ref = _oracle(orc, gt, data, field, mask, order, max_iter=500, tol=1e-6)
    c = out["coefficients"][mask].astype(np.float64)
    cr = ref["coefficients"][mask].astype(np.float64)
    err = _rel(c, cr)
    it, itr = out["iterations"][mask].astype(float), ref["iterations"][mask].astype(float)
    print(f"order {order}: coef err median {np.median(err):.2e} p99 {np.quantile(err, 0.99):.2e} "
          f"max {err.max():.2e}; iterations {itr.min():.0f}-{itr.max():.0f}")
    assert np.median(err) <= 1e-5 and np.quantile(err, 0.99) <= 1e-3
    assert np.mean(np.abs(it - itr) <= np.maximum(1, 0.1 * itr)) >= 0.9
    np.testing.assert_allclose(out["residual_l2"][mask], ref["residual_l2"][mask], rtol=1e-3)
    np.testing.assert_allclose(out["r2"][mask], ref["r2"][mask], rtol=0, atol=1e-4)
    assert np.mean(np.abs(out["bic"][mask] - ref["bic"][mask]) <= 0.5) >= 0.95
    assert (out["alpha"][mask] == np.float32(1e-3)).all()
    assert (out["regularization_kind"][mask] == 1).all()
    assert not out["coefficients"][~mask].any() and not out["iterations"][~mask].any()
    # identity voxels are the shared-design fit of the same voxels
    plain = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, radial_order=order, alpha_mode="fixed", alpha=1e-3,
                         max_iter=500, tol=1e-6, diagnostics=True, **_kw(gt))
    for idx in [(0, 1, 0), (0, 1, 1)]:
        np.testing.assert_array_equal(out["coefficients"][idx], plain["coefficients"][idx])
        assert out["iterations"][idx] == plain["iterations"][idx]


def test_alpha_ratio_and_zero_voxel(orc):
    gt, data, field = _volume((3, 3, 2), seed=2)
    data[0, 0, 0] = 0.0
    out = sf.shore_fit(data, gt.bvals, gt.bvecs, np.ones(data.shape[:3], bool), radial_order=6,
                       alpha_mode="alpha-ratio", alpha_ratio=1e-3, max_iter=500, tol=1e-6, diagnostics=True,
                       graddev=field, **_kw(gt))
    ref = _oracle(orc, gt, data, field, None, 6, alpha_mode="alpha-ratio", alpha_ratio=1e-3, max_iter=500,
                  tol=1e-6)
    assert out["iterations"][0, 0, 0] == 0 and out["alpha"][0, 0, 0] == 0 and out["bic"][0, 0, 0] == 0
    assert not out["coefficients"][0, 0, 0].any() and out["residual_l2"][0, 0, 0] == 0
    np.testing.assert_allclose(out["alpha"], ref["alpha"], rtol=1e-6)
    err = _rel(out["coefficients"].reshape(-1, 72).astype(np.float64),
               ref["coefficients"].reshape(-1, 72).astype(np.float64))
    assert np.median(err) <= 1e-5 and err.max() <= 1e-3
    assert out["alpha_distribution"] is not None


@pytest.mark.parametrize("order", [4, 6, 8, 10])
def test_variants_and_chunks_agree(order):
    gt, data, field = _volume((4, 4, 2), seed=3)
    mask = np.ones(data.shape[:3], bool)
    kw = dict(radial_order=order, alpha_mode="fixed", alpha=1e-3, max_iter=300, tol=1e-6, diagnostics=True,
              graddev=field, **_kw(gt))
    base = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, graddev_variant=0, **kw)
    for i in range(1, len(sf.graddev_variants(order))):
        other = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, graddev_variant=i, **kw)
        err = _rel(other["coefficients"][mask].astype(np.float64), base["coefficients"][mask].astype(np.float64))
        assert np.median(err) <= 1e-5 and err.max() <= 1e-3, (order, i)
    old = os.environ.get("SHOREFIT_GRADDEV_WORKSPACE_MB")
    os.environ["SHOREFIT_GRADDEV_WORKSPACE_MB"] = "1"  # a few voxels per chunk
    try:
        chunked = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, graddev_variant=0, **kw)
    finally:
        if old is None:
            del os.environ["SHOREFIT_GRADDEV_WORKSPACE_MB"]
        else:
            os.environ["SHOREFIT_GRADDEV_WORKSPACE_MB"] = old
    np.testing.assert_array_equal(chunked["coefficients"], base["coefficients"])
    np.testing.assert_array_equal(chunked["iterations"], base["iterations"])


def test_many_gradients_fall_back_to_a_smaller_chunk(orc):
    """774 gradients at order 10: the default Gram chunk (64 design rows)
    does not fit LDS, so the default falls back to the first variant that
    does; G_v still matches the oracle and the fit matches the explicitly
    chosen 16-row variant."""
    gt, data, field = _volume((2, 1, 1))
    bvals, bvecs = np.tile(gt.bvals, 3), np.tile(gt.bvecs, (3, 1))
    data3 = np.ascontiguousarray(np.tile(data, (1, 1, 1, 3)))
    G, _, _ = sf.graddev_gram(data3, bvals, bvecs, field, np.array([0, 1]), radial_order=10, **_kw(gt))
    T = field.reshape(-1, 3, 3).astype(np.float64)
    for v in (0, 1):
        be, d = orc.effective_table(T[v], np.eye(3), bvals, bvecs, bvecs)
        A = orc.design_matrix(10, 700.0, be, d, gt.big_delta, gt.small_delta)
        Gr = A.T @ A
        assert np.abs(G[v] - Gr).max() <= 2e-7 * np.abs(Gr).max()
    kw = dict(radial_order=10, alpha_mode="fixed", alpha=1e-3, max_iter=30, tol=0.0, graddev=field, **_kw(gt))
    m = np.ones((2, 1, 1), bool)
    a = sf.shore_fit(data3, bvals, bvecs, m, **kw)
    i16 = [v["rows"] for v in sf.graddev_variants(10)].index(16)
    b = sf.shore_fit(data3, bvals, bvecs, m, graddev_variant=i16, **kw)
    np.testing.assert_array_equal(a["coefficients"], b["coefficients"])


def test_autotune_picks_a_variant():
    gt, data, field = _volume((4, 4, 2), seed=4)
    res = sf.graddev_tune(data, gt.bvals, gt.bvecs, np.ones(data.shape[:3], bool), field, radial_order=8,
                          max_iter=200, max_voxels=32, **_kw(gt))
    n = len(sf.graddev_variants(8))
    assert 0 <= res["variant"] < n and len(res["ms"]) == n and all(m > 0 for m in res["ms"])


def test_unsupported_options_raise():
    gt, data, field = _volume((2, 2, 1))
    m = np.ones(data.shape[:3], bool)
    with pytest.raises(ValueError, match="alpha_mode fixed or alpha-ratio"):
        sf.shore_fit(data, gt.bvals, gt.bvecs, m, radial_order=6, graddev=field, **_kw(gt))
    with pytest.raises(ValueError, match="radial orders 2 to 10"):
        sf.shore_fit(data, gt.bvals, gt.bvecs, m, radial_order=12, alpha_mode="fixed", graddev=field, **_kw(gt))
    with pytest.raises(ValueError, match="graddev must have shape"):
        sf.shore_fit(data, gt.bvals, gt.bvecs, m, radial_order=6, alpha_mode="fixed", graddev=field[:1], **_kw(gt))
