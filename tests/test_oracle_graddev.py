"""The oracle's per-voxel-design (gradient-deviation) fit, pinned on the CPU.

The fit is a composition of restatements that are pinned elsewhere: the
effective table (src/graddev.rs:160-185; KATs of src/graddev.rs:218-253 in
test_oracle_basis.py), the design matrix (src/basis/shore.rs KATs and scipy
goldens), the Lipschitz constant (numpy SVD) and FISTA (sklearn fixtures,
KKT). These tests pin the composition itself:

* T = I everywhere gives exactly the nominal volume fit (the identity branch,
  src/graddev.rs:35-36, :89-93);
* a deviated voxel equals the single-voxel chain effective table -> design ->
  FISTA on the same f32-rounded tensor;
* an isotropic T = s I scales every b by s^2 and keeps the directions, so its
  design is the nominal builder's at b s^2 (closed form of q = T^T g);
* the per-voxel Lipschitz constant is sigma_max(A_v)^2 / N by numpy's SVD;
* KKT optimality of the deviated voxels' solutions for their own A_v.
"""

import numpy as np
import pytest

from cs_dmri_amd import workloads as W


@pytest.fixture(scope="module")
def block():
    gt, sig = W.config2_signals(24)
    data = sig.reshape(2, 3, 4, -1).astype(np.float32)
    T = W.graddev_tensors(24).reshape(2, 3, 4, 3, 3)
    return gt, data, T


def _fit(orc, gt, data, T, order=6, **kw):
    return orc.fit_graddev(data, None, T, gt.bvals, gt.bvecs, big_delta=gt.big_delta, small_delta=gt.small_delta,
                           radial_order=order, **kw)


def test_identity_field_is_the_nominal_fit(orc, block):
    gt, data, _ = block
    eye = np.broadcast_to(np.eye(3), data.shape[:3] + (3, 3)).copy()
    o = _fit(orc, gt, data, eye, max_iter=300)
    A0 = orc.design_matrix(6, 700.0, gt.bvals, gt.bvecs, gt.big_delta, gt.small_delta)
    ref = orc.fit_volume(data, None, A0, alpha_mode="fixed", alpha=1e-3, max_iter=300)
    for k in ("coefficients", "iterations", "residual_l2", "r2", "bic"):
        np.testing.assert_array_equal(o[k], ref[k], err_msg=k)
    assert np.all(o["lipschitz"] == orc.lipschitz(A0))


def test_deviated_voxel_equals_single_voxel_chain(orc, block):
    gt, data, T = block
    o = _fit(orc, gt, data, T, max_iter=300)
    T32 = T.astype(np.float32).astype(np.float64)
    for idx in [(0, 0, 0), (1, 2, 3), (0, 1, 2)]:
        b, d = orc.effective_table(T32[idx], np.eye(3), gt.bvals, gt.bvecs, gt.bvecs)
        A = orc.design_matrix(6, 700.0, b, d, gt.big_delta, gt.small_delta)
        L = orc.lipschitz(A)
        c, diag = orc.fista(A, data[idx].astype(np.float64), 1e-3, max_iter=300, tol=1e-6, L=L)
        np.testing.assert_array_equal(o["coefficients"][idx], c.astype(np.float32))
        assert o["iterations"][idx] == diag["iterations"]
        assert o["residual_l2"][idx] == np.float32(diag["residual_l2"])
        assert o["lipschitz"][idx] == L
        sv = np.linalg.svd(A, compute_uv=False)
        assert L == pytest.approx(sv[0] ** 2 / A.shape[0], rel=1e-12)


def test_isotropic_scaling_scales_b(orc, block):
    gt, data, _ = block
    s = 1.0625  # exact in float32, so the field carries the same tensor
    T = np.broadcast_to(s * np.eye(3), data.shape[:3] + (3, 3)).copy()
    b, d = orc.effective_table(s * np.eye(3), np.eye(3), gt.bvals, gt.bvecs, gt.bvecs)
    dw = gt.bvals > 50.0
    np.testing.assert_allclose(b[dw], gt.bvals[dw] * s * s, rtol=1e-14)
    np.testing.assert_allclose(d[dw], gt.bvecs[dw], atol=1e-15)
    A = orc.design_matrix(6, 700.0, b, d, gt.big_delta, gt.small_delta)
    bs = np.where(dw, gt.bvals * s * s, gt.bvals)
    A2 = orc.design_matrix(6, 700.0, bs, gt.bvecs, gt.big_delta, gt.small_delta)
    np.testing.assert_allclose(A, A2, rtol=0, atol=1e-14 * np.abs(A2).max())
    o = _fit(orc, gt, data[:1, :1, :2], T[:1, :1, :2], max_iter=200)
    c, _ = orc.fista(A2, data[0, 0, 0].astype(np.float64), 1e-3, max_iter=200, tol=1e-6)
    np.testing.assert_allclose(o["coefficients"][0, 0, 0], c.astype(np.float32), rtol=1e-6,
                               atol=1e-7 * np.abs(c).max())


def test_kkt_of_deviated_voxels(orc, block):
    # converged lasso solutions satisfy |A^T (s - A c)| / N <= alpha (+ slack)
    # with equality on the support (tests/common/mod.rs:127-189 conditions)
    gt, data, T = block
    o = _fit(orc, gt, data, T, max_iter=20000, tol=1e-10)
    T32 = T.astype(np.float32).astype(np.float64)
    for idx in [(0, 0, 0), (1, 1, 1)]:
        b, d = orc.effective_table(T32[idx], np.eye(3), gt.bvals, gt.bvecs, gt.bvecs)
        A = orc.design_matrix(6, 700.0, b, d, gt.big_delta, gt.small_delta)
        c = o["coefficients"][idx].astype(np.float64)
        g = A.T @ (data[idx].astype(np.float64) - A @ c) / A.shape[0]
        on = c != 0
        scale = np.abs(A.T @ data[idx]).max() / A.shape[0]
        assert np.all(np.abs(g[~on]) <= 1e-3 + 1e-4 * scale)
        np.testing.assert_allclose(g[on], 1e-3 * np.sign(c[on]), atol=1e-4 * scale)


def test_alpha_ratio_and_zero_voxel(orc, block):
    gt, data, T = block
    d = data.copy()
    d[0, 0, 0] = 0.0
    o = _fit(orc, gt, d, T, alpha_mode="alpha-ratio", alpha_ratio=1e-3, max_iter=200)
    # src/solver/alpha.rs:178-189: an all-zero signal short-circuits
    assert o["iterations"][0, 0, 0] == 0 and o["alpha"][0, 0, 0] == 0 and not o["coefficients"][0, 0, 0].any()
    assert o["bic"][0, 0, 0] == 0 and o["residual_l2"][0, 0, 0] == 0
    T32 = T.astype(np.float32).astype(np.float64)
    b, dd = orc.effective_table(T32[1, 0, 0], np.eye(3), gt.bvals, gt.bvecs, gt.bvecs)
    A = orc.design_matrix(6, 700.0, b, dd, gt.big_delta, gt.small_delta)
    amax = orc.alpha_max(A, d[1, 0, 0].astype(np.float64))
    assert o["alpha"][1, 0, 0] == np.float32(amax * 1e-3)
