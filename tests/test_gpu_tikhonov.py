"""GPU parity of the Tikhonov L2 fit at whole-brain shape (order 8, K = 145).

The workload is the L2 leg of the whole-brain configuration: 1 b0 plus 257
Cartesian q-points (seed 3, bmax 5000), SHORE order 8 (K = 145), zeta 700,
lambda_n = lambda_l = 1e-2, multi-tensor signals at SNR 20. At this size the
device runs the blocked Cholesky with a one-column remainder (145 = 18*8 + 1),
the 16-lane shuffle column solve for P = (A^T A + Lambda)^-1 A^T, and the
split-FP16 apply kernel (padded K 160 = 10 blocks of 16; 258 samples = 17
blocks of 16, an odd count, the last one partial).

Tolerances
----------
* coefficients (FP64 factorisation; apply on FP16 hi/lo splits with FP32
  accumulation, ~22 bits per product) vs the FP64 oracle
  (TikhonovSolver::fit, src/solver/tikhonov.rs:59-76): 1e-4 relative per voxel;
* residual_l2: 1e-4 relative (the reference's residual is |s - A c| in FP64);
* normal equations (tests/optimization_kkt.rs:164-218 restated on the
  device's FP32 coefficients, evaluated in FP64): the relative residual
  |(A^T A + Lambda) c - A^T s| / |A^T s| is below 1e-5, the relative
  scale of FP32 coefficient parity (measured: at most 4.4e-7 with the
  split-FP16 apply and 6.4e-7 with the FP32 one, against 5e-8 for the
  oracle's FP64 solution merely rounded to float32). The oracle's own FP64
  solution meets the reference's 1e-10 on this design
  (tests/test_oracle_solvers.py);
* rss_l2 map of the L2-anchored strategy (src/fit.rs:433-438): 1e-4 relative
  against the oracle's Tikhonov residual squared.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import cs_dmri_amd as sf  # noqa: E402
from cs_dmri_amd import workloads as W  # noqa: E402

LAM = 1e-2


def _rel(c, ref):
    return np.linalg.norm(c - ref, axis=-1) / np.maximum(np.linalg.norm(ref, axis=-1), 1e-30)


def _normal_residual(A, dn, dl, s, c):
    """tests/optimization_kkt.rs `tikhonov_normal_residual`, in FP64."""
    G = A.T @ A + np.diag(LAM * dn + LAM * dl)
    rhs = A.T @ s
    return np.linalg.norm(G @ c - rhs) / np.linalg.norm(rhs)


@pytest.fixture(scope="module")
def cfg2(orc):
    gt = W.cartesian_table(257, seed=3)
    shape = (7, 7, 7)  # 343 voxels less the holes: not a multiple of the 128-voxel workgroup
    mask = np.ones(shape, bool)
    mask[0, :, 0] = False
    mask[3, 3, 3] = False
    data = np.zeros(shape + (len(gt.bvals),), np.float32)
    data[mask] = W.multi_tensor_signal(gt, int(mask.sum()), W.S0 / 20.0, seed=2)
    data[~mask] = 5.0  # values outside the mask must not reach the fit
    A = orc.design_matrix(8, 700.0, gt.bvals, gt.bvecs, gt.big_delta, gt.small_delta)
    dn, dl = orc.regularization(8)
    assert A.shape == (258, 145)
    return gt, data, mask, A, dn, dl


def test_tikhonov_volume_order8(orc, cfg2):
    gt, data, mask, A, dn, dl = cfg2
    out = sf.shore_fit(data, gt.bvals, gt.bvecs, mask, big_delta=gt.big_delta, small_delta=gt.small_delta,
                       radial_order=8, regularization="l2", lambda_n=LAM, lambda_l=LAM, diagnostics=True)
    ref = orc.fit_volume(data, mask, A, regularization="l2", dn=dn, dl=dl, lambda_n=LAM, lambda_l=LAM)
    c, cr = out["coefficients"][mask].astype(np.float64), ref["coefficients"][mask].astype(np.float64)
    err = _rel(c, cr)
    print(f"order-8 Tikhonov: coef err median {np.median(err):.2e} max {err.max():.2e}")
    assert err.max() <= 1e-4
    assert out["coefficients"].shape == mask.shape + (145,)
    assert not out["coefficients"][~mask].any() and not out["residual_l2"][~mask].any()
    np.testing.assert_allclose(out["residual_l2"][mask], ref["residual_l2"][mask], rtol=1e-4)
    np.testing.assert_allclose(out["r2"][mask], ref["r2"][mask], rtol=1e-4, atol=1e-6)
    assert (out["iterations"][mask] == 1).all() and (out["regularization_kind"][mask] == 2).all()
    sc = out["sidecar"]["solver"]
    assert sc["kind"] == "tikhonov" and sc["lambda_n"] == LAM and sc["lambda_l"] == LAM
    # normal equations on the device's coefficients (every 7th voxel)
    S = data[mask].astype(np.float64)
    worst, worst_floor = 0.0, 0.0
    for v in range(0, len(S), 7):
        r = _normal_residual(A, dn, dl, S[v], c[v])
        floor = _normal_residual(A, dn, dl, S[v], cr[v].astype(np.float32).astype(np.float64))
        worst, worst_floor = max(worst, r), max(worst_floor, floor)
        assert r < 1e-5, (v, r, floor)
    print(f"  normal-equation residual max {worst:.2e} (float32 rounding floor {worst_floor:.2e})")


def test_tikhonov_batch_order8(orc, cfg2):
    gt, data, mask, A, dn, dl = cfg2
    S = data[mask]
    C, R = sf.tikhonov_batch(A, S, dn, dl, LAM, LAM)
    Cr, Rr = orc.tikhonov(A, S.astype(np.float64), dn, dl, LAM, LAM)
    err = _rel(C.astype(np.float64), Cr)
    print(f"tikhonov_batch K=145: coef err max {err.max():.2e}")
    assert err.max() <= 1e-4
    np.testing.assert_allclose(R, Rr, rtol=1e-4)
    # one signal, and partial 16-voxel tiles and workgroups
    for n in (1, 17, 33):
        Cn, Rn = sf.tikhonov_batch(A, S[:n], dn, dl, LAM, LAM)
        np.testing.assert_array_equal(Cn, C[:n])
        np.testing.assert_array_equal(Rn, R[:n])


def test_l2_anchored_rss_l2_order8(orc, cfg2):
    # rss_l2 is the Tikhonov reference of the L2-anchored rule, computed per
    # voxel with the spec's lambdas (src/fit.rs:433-438)
    gt, data, mask, A, dn, dl = cfg2
    sub = np.ascontiguousarray(data[:3])
    msub = mask[:3]
    out = sf.shore_fit(sub, gt.bvals, gt.bvecs, msub, big_delta=gt.big_delta, small_delta=gt.small_delta,
                       radial_order=8, alpha_mode="l2-anchored", path_n_alphas=4, max_iter=40, tol=1e-6,
                       lambda_n=LAM, lambda_l=LAM, diagnostics=True)
    _, Rr = orc.tikhonov(A, sub[msub].astype(np.float64), dn, dl, LAM, LAM)
    np.testing.assert_allclose(out["rss_l2"][msub], (Rr ** 2).astype(np.float32), rtol=1e-4)
    assert not out["rss_l2"][~msub].any()
