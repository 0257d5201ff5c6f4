"""Edge cases of the LDS-staged apply kernels, on the device through the C ABI.

`tikhonov_h16_kernel` (N <= 288, padded K <= 160) runs c = P s and the
residual on the FP16 MFMA with hi/lo splits of both operands; it stages P and
A by chunks of 32 measurements (two 16-sample blocks), a fixed number of
chunks with zeros past the signal, and keeps the signals in registers; the
last sample block of the signal is partial unless 16 divides N. Longer
signals, and padded K above 160 (order 10), go to the FP32 kernels
(`tikhonov_lds_kernel` up to 288 measurements, then the kernels that read the
signals twice). The epilogue of the persistent solvers (residual, R^2, BIC)
is the same split-FP16 residual pass (`epilogue_h16_kernel`, N <= 288) or,
past 288 measurements, the FP32 `epilogue_lds_kernel`, which stages A by
chunks of 16 rows with no bound on N. The cases cover one block (N <= 16), a
block plus one sample, odd and even block counts, the largest register case,
the fallbacks past it, and order 10 on both sides of 288 measurements.

Tolerances: Tikhonov coefficients and residuals against the FP64 oracle
(TikhonovSolver::fit, src/solver/tikhonov.rs:59-76) at 1e-4 relative, as the
order-8 tests; the epilogue's residual, R^2 and BIC against the same maps
recomputed in FP64 (src/fit.rs:120-128, src/solver/alpha.rs:99-103) from the
device's own coefficients, so the check isolates the epilogue from the
solver: residual 1e-5 relative (products of ~22 bits, FP32 sums over N
samples).
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import cs_dmri_amd as sf  # noqa: E402
from cs_dmri_amd import workloads as W  # noqa: E402

from .fixtures import shell_gtab  # noqa: E402


def _rel(c, ref):
    return np.linalg.norm(c - ref, axis=-1) / np.maximum(np.linalg.norm(ref, axis=-1), 1e-30)


@pytest.mark.parametrize("N,order", [(8, 2), (16, 2), (17, 2), (33, 4), (48, 4), (271, 4), (288, 4), (289, 4),
                                     (400, 4), (200, 10), (300, 10)])
def test_tikhonov_batch_chunk_counts(orc, N, order):
    rng = np.random.default_rng(N)
    K = orc.n_coeffs(order)
    A = rng.standard_normal((N, K))
    dn, dl = orc.regularization(order)
    S = (A @ rng.standard_normal((K, 70))).T + 0.1 * rng.standard_normal((70, N))  # 70 voxels: partial tiles
    C, R = sf.tikhonov_batch(A, S, dn, dl, 1e-2, 1e-2)
    Cr, Rr = orc.tikhonov(A, S.astype(np.float32).astype(np.float64), dn, dl, 1e-2, 1e-2)
    err = _rel(C.astype(np.float64), Cr)
    assert err.max() <= 1e-4, (N, err.max())
    np.testing.assert_allclose(R, Rr, rtol=1e-4)


def _gtab(n_total):
    """1 b0 plus golden-spiral shells with n_total measurements in all."""
    per = (n_total - 1) // 3
    bvals, bvecs = shell_gtab(per, shells=(1000.0, 2000.0, 3000.0))
    extra = n_total - len(bvals)
    if extra:
        b2, v2 = shell_gtab(extra, shells=(2500.0,))
        bvals = np.concatenate([bvals, b2[1:]])
        bvecs = np.concatenate([bvecs, v2[1:]])
    assert len(bvals) == n_total
    return bvals, bvecs


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
@pytest.mark.parametrize("N,order", [(16, 2), (17, 2), (49, 4), (301, 6)])
def test_epilogue_maps_chunk_counts(orc, precision, N, order):
    bvals, bvecs = _gtab(N)
    big, small = 0.04, 0.012
    gt = W.GradientTable(bvals, bvecs, big, small, 50.0)
    shape = (5, 4, 3)
    mask = np.ones(shape, bool)
    mask[1, 1, 1] = False
    data = np.zeros(shape + (N,), np.float32)
    data[mask] = W.multi_tensor_signal(gt, int(mask.sum()), W.S0 / 30, seed=N)
    out = sf.shore_fit(data, bvals, bvecs, mask, big_delta=big, small_delta=small, radial_order=order,
                       alpha_mode="fixed", alpha=1e-3, max_iter=150, tol=1e-6, precision=precision,
                       diagnostics=True)
    A = orc.design_matrix(order, 700.0, bvals, bvecs, big, small)
    c = out["coefficients"][mask].astype(np.float64)
    s = data[mask].astype(np.float64)
    rss = ((s - c @ A.T) ** 2).sum(-1)
    sst = ((s - s.mean(-1, keepdims=True)) ** 2).sum(-1)
    np.testing.assert_allclose(out["residual_l2"][mask], np.sqrt(rss), rtol=1e-5)
    np.testing.assert_allclose(out["r2"][mask], 1.0 - rss / sst, rtol=0, atol=1e-5)
    nnz = (out["coefficients"][mask] != 0).sum(-1)
    bic = N * np.log(np.maximum(rss, np.finfo(float).tiny) / N) + np.log(N) * nnz
    np.testing.assert_allclose(out["bic"][mask], bic, rtol=1e-5, atol=1e-2)
    assert not out["residual_l2"][~mask].any() and not out["coefficients"][~mask].any()


@pytest.mark.parametrize("scale", [1e-15, 1e15])
def test_tikhonov_batch_signal_scales(orc, scale):
    # the split-FP16 apply scales each voxel's signals and coefficients by a
    # power of two of their own magnitude: far from 1 the results follow the
    # signals' scale exactly as the FP64 oracle's do
    rng = np.random.default_rng(7)
    N, order = 258, 8
    K = orc.n_coeffs(order)
    A = rng.standard_normal((N, K))
    dn, dl = orc.regularization(order)
    S = ((A @ rng.standard_normal((K, 40))).T + 0.1 * rng.standard_normal((40, N))) * scale
    C, R = sf.tikhonov_batch(A, S, dn, dl, 1e-2, 1e-2)
    Cr, Rr = orc.tikhonov(A, S.astype(np.float32).astype(np.float64), dn, dl, 1e-2, 1e-2)
    assert np.isfinite(C).all() and np.isfinite(R).all()
    assert _rel(C.astype(np.float64), Cr).max() <= 1e-4
    np.testing.assert_allclose(R, Rr, rtol=1e-4)
