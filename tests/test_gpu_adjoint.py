"""GPU checks of the differentiable lasso fit (forward with per-voxel alpha
and warm starts, the implicit-differentiation adjoint, the autograd
Function, and the held-out alpha search).

Tolerances:
* forward with a per-voxel alpha equal to a scalar alpha: bit-identical;
  the persistent forward kernel against the plain FP32 kernel: bit-identical;
  against the FP64 oracle at equal iteration counts: 1e-5 per voxel;
* adjoint against the FP64 oracle adjoint on the same coefficients: 1e-6
  relative per voxel (both FP64 Cholesky on the same active set);
* adjoint against central finite differences (h = 1e-4 relative) of the
  exact, KKT-verified FP64 minimiser: 1e-4 relative (the FP32 coefficient
  input of the device adjoint is the only approximation);
* autograd gradients are the adjoint's;
* the hypergradient search lowers the mean held-out MSE from alpha0 and
  lands within 5% of the per-voxel best of the 20-point grid.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from cs_dmri_amd import workloads as W  # noqa: E402
from cs_dmri_amd.autograd import LassoDesign, lasso  # noqa: E402
from cs_dmri_amd.hypergrad import grid_search, heldout_mse, hypergradient_search  # noqa: E402

DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def problem(orc):
    gt, sig, tr, te = W.config5(512)
    A = orc.design_matrix(8, 700.0, gt.bvals, gt.bvecs, gt.big_delta, gt.small_delta)
    return (np.ascontiguousarray(A[tr]), np.ascontiguousarray(A[te]), np.ascontiguousarray(sig[:, tr], np.float32),
            np.ascontiguousarray(sig[:, te], np.float32))


def _rel(a, b):
    return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-30)


def test_forward_per_voxel_alpha_and_warm_start(orc, problem):
    A_tr, _, S_tr, _ = problem
    d = LassoDesign(A_tr)
    S = torch.from_numpy(S_tr[:128]).to(DEV)
    a = 2.0 ** -10  # exact in float32
    c1, i1 = d.forward(S, a, max_iter=300, tol=1e-6)
    c2, i2 = d.forward(S, torch.full((128,), a, device=DEV), max_iter=300, tol=1e-6)
    assert torch.equal(c1, c2) and torch.equal(i1, i2)
    # equal iteration counts against the oracle with per-voxel alphas
    alphas = np.float32(2.0 ** -10) * np.float32(2.0) ** np.arange(-4, 4, dtype=np.float32)[np.arange(128) % 8]
    c3, _ = d.forward(S, torch.from_numpy(alphas).to(DEV), max_iter=60, tol=0.0)
    c3 = c3.cpu().numpy().astype(np.float64)
    for v in range(0, 128, 9):
        ref, _ = orc.fista(A_tr, S_tr[v].astype(np.float64), float(alphas[v]), max_iter=60, tol=0.0)
        assert _rel(c3[v], ref) <= 1e-5, v
    # warm start x0 = y0 = warm, t = 1 (src/solver/fista.rs:76-80): equal
    # iteration counts against the oracle's warm-started FISTA
    w0 = c3.astype(np.float32)
    c4, _ = d.forward(S, a, warm=torch.from_numpy(w0).to(DEV), max_iter=25, tol=0.0)
    c4 = c4.cpu().numpy().astype(np.float64)
    for v in range(0, 128, 9):
        ref, _ = orc.fista(A_tr, S_tr[v].astype(np.float64), a, max_iter=25, tol=0.0, warm=w0[v].astype(np.float64))
        assert _rel(c4[v], ref) <= 1e-5, v
    # restarting from a converged fit needs far fewer iterations
    _, i5 = d.forward(S, a, warm=c1, max_iter=300, tol=1e-6)
    assert float(i5.float().median()) < 0.5 * float(i1.float().median())


def test_forward_persistent_matches_plain_kernel(problem):
    """The lasso forward runs the persistent FP32 kernel (slots refilled as
    voxels stop); fista_batch on the same design runs the plain tile kernel.
    Same arithmetic per voxel: identical coefficients and iteration counts."""
    import cs_dmri_amd as sf
    A_tr, _, S_tr, _ = problem
    d = LassoDesign(A_tr)
    S = np.ascontiguousarray(S_tr[:300])
    c_p, i_p = d.forward(torch.from_numpy(S).to(DEV), 1e-3, max_iter=500, tol=1e-6)
    c_f, info = sf.fista_batch(A_tr, S, 1e-3, max_iter=500, tol=1e-6, precision="fp32")
    its = i_p.cpu().numpy().astype(np.int64)
    assert its.min() < its.max()  # uneven convergence: slots were refilled
    np.testing.assert_array_equal(its, info["iterations"].astype(np.int64))
    np.testing.assert_array_equal(c_p.cpu().numpy(), c_f)
    for n in (1, 17):  # fewer voxels than one wave's slots, and one more than a tile
        c_n, i_n = d.forward(torch.from_numpy(S[:n].copy()).to(DEV), 1e-3, max_iter=500, tol=1e-6)
        np.testing.assert_array_equal(c_n.cpu().numpy(), c_f[:n])
        np.testing.assert_array_equal(i_n.cpu().numpy().astype(np.int64), its[:n])


def test_adjoint_matches_fp64_oracle(orc, problem):
    A_tr, _, S_tr, _ = problem
    d = LassoDesign(A_tr)
    S = torch.from_numpy(S_tr[:256]).to(DEV)
    coef, _ = d.forward(S, 1e-3, max_iter=500, tol=1e-6)
    g = torch.randn(coef.shape, generator=torch.Generator().manual_seed(3)).to(DEV)
    gs, ga, st = d.adjoint(coef, g)
    assert (st.cpu().numpy() == 0).all()
    rs, ra = orc.lasso_adjoint(A_tr, coef.cpu().numpy().astype(np.float64), g.cpu().numpy().astype(np.float64))
    err = _rel(gs.cpu().numpy().astype(np.float64), rs)
    print(f"adjoint vs FP64: grad_signal max rel {err.max():.2e}, "
          f"grad_alpha max rel {np.max(np.abs(ga.cpu().numpy() - ra) / np.abs(ra)):.2e}")
    assert err.max() <= 1e-6
    np.testing.assert_allclose(ga.cpu().numpy(), ra, rtol=1e-6)


def test_adjoint_register_and_lds_paths(orc, problem):
    """Active sets of at most 64 coefficients go through the register kernel,
    larger ones through the LDS kernel: both against the FP64 oracle on one
    batch that holds both kinds and an empty active set."""
    A_tr, _, S_tr, _ = problem
    d = LassoDesign(A_tr)
    S = torch.from_numpy(S_tr[:128]).to(DEV)
    c_big, _ = d.forward(S, 1e-3, max_iter=500, tol=1e-6)
    c_small, _ = d.forward(S, 2e-5, max_iter=500, tol=1e-6)
    coef = torch.cat([c_big, c_small, torch.zeros_like(c_big[:1])])
    nnz = (coef != 0).sum(dim=1).cpu().numpy()
    assert nnz[-1] == 0 and (nnz[:-1] <= 64).any() and (nnz > 64).any(), (nnz.min(), nnz.max())
    g = torch.randn(coef.shape, generator=torch.Generator().manual_seed(4)).to(DEV)
    gs, ga, st = d.adjoint(coef, g)
    assert (st.cpu().numpy() == 0).all()
    assert not gs[-1].any() and float(ga[-1]) == 0.0
    rs, ra = orc.lasso_adjoint(A_tr, coef[:-1].cpu().numpy().astype(np.float64),
                               g[:-1].cpu().numpy().astype(np.float64))
    err = _rel(gs[:-1].cpu().numpy().astype(np.float64), rs)
    print(f"active sets {nnz[:-1].min()}-{nnz.max()}: grad_signal max rel {err.max():.2e}")
    assert err.max() <= 1e-6
    np.testing.assert_allclose(ga[:-1].cpu().numpy(), ra, rtol=1e-6)


def test_adjoint_active_set_sizes(orc):
    """Active sets of 1, 2, 63, 64 (the largest register-kernel case), 65 (the
    smallest LDS-kernel case), 100 and all 145 coefficients in one batch, on a
    well-conditioned Gaussian design; the adjoint depends only on the
    support and signs of the coefficients, so they are drawn directly."""
    rng = np.random.default_rng(11)
    A = rng.standard_normal((232, 145))
    d = LassoDesign(A)
    sizes = np.repeat([1, 2, 63, 64, 65, 100, 145], 3)
    coef = np.zeros((len(sizes), 145), np.float32)
    for i, m in enumerate(sizes):
        S = rng.choice(145, m, replace=False)
        coef[i, S] = rng.choice([-1.0, 1.0], m) * rng.uniform(0.1, 1.0, m)
    g = rng.standard_normal(coef.shape).astype(np.float32)
    gs, ga, st = d.adjoint(torch.from_numpy(coef).to(DEV), torch.from_numpy(g).to(DEV))
    assert (st.cpu().numpy() == 0).all()
    rs, ra = orc.lasso_adjoint(A, coef.astype(np.float64), g.astype(np.float64))
    err = _rel(gs.cpu().numpy().astype(np.float64), rs)
    assert err.max() <= 1e-6, dict(zip(sizes.tolist(), err.tolist()))
    np.testing.assert_allclose(ga.cpu().numpy(), ra, rtol=1e-6)


def test_adjoint_against_finite_differences(orc, problem):
    A_tr, _, S_tr, _ = problem
    rng = np.random.default_rng(5)
    d = LassoDesign(A_tr)
    S = torch.from_numpy(S_tr[:64]).to(DEV)
    coef, _ = d.forward(S, 1e-3, max_iter=2000, tol=1e-7)
    checked = 0
    for v in range(64):
        s = S_tr[v].astype(np.float64)
        c, st = orc.lasso_polish(A_tr, s, 1e-3, coef[v].cpu().numpy().astype(np.float64))
        if st != 0:
            continue  # the FP32 active set is not the exact one for this voxel
        g = rng.standard_normal(A_tr.shape[1])
        gs, ga, _ = d.adjoint(torch.from_numpy(c[None].astype(np.float32)).to(DEV),
                              torch.from_numpy(g[None].astype(np.float32)).to(DEV))
        dirn = rng.standard_normal(len(s))
        h = 1e-4 * np.linalg.norm(s) / np.linalg.norm(dirn)
        cp, sp = orc.lasso_polish(A_tr, s + h * dirn, 1e-3, c)
        cm, sm = orc.lasso_polish(A_tr, s - h * dirn, 1e-3, c)
        if sp or sm:
            continue  # the active set changes within h
        g32 = g.astype(np.float32).astype(np.float64)
        fd = (g32 @ cp - g32 @ cm) / (2 * h)
        assert float(gs[0].cpu().numpy().astype(np.float64) @ dirn) == pytest.approx(fd, rel=1e-4), v
        ha = 1e-4 * 1e-3
        ap, sp = orc.lasso_polish(A_tr, s, 1e-3 + ha, c)
        am, sm = orc.lasso_polish(A_tr, s, 1e-3 - ha, c)
        if not (sp or sm):
            assert float(ga[0]) == pytest.approx((g32 @ ap - g32 @ am) / (2 * ha), rel=1e-4), v
        checked += 1
    print(f"finite-difference checks on {checked} of 64 voxels")
    assert checked >= 16


def test_autograd_function(problem):
    A_tr, _, S_tr, _ = problem
    d = LassoDesign(A_tr)
    S = torch.from_numpy(S_tr[:64]).to(DEV).requires_grad_(True)
    alpha = torch.full((64,), 1e-3, device=DEV, requires_grad=True)
    coef, its = lasso(S, alpha, d, max_iter=500, tol=1e-6)
    g = torch.randn(coef.shape, generator=torch.Generator().manual_seed(4)).to(DEV)
    (coef * g).sum().backward()
    gs, ga, _ = d.adjoint(coef.detach(), g)
    assert torch.equal(S.grad, gs) and torch.equal(alpha.grad, ga)
    assert its.dtype == torch.int32 and not its.requires_grad


def test_autograd_singular_active_set():
    # A zero column of A whose coefficient is active (a warm start keeps it
    # away from zero for a few iterations) makes G_SS singular: its pivot is
    # exactly zero. The adjoint reports status 1 for those signals; the
    # autograd backward warns and returns zero gradients for them and leaves
    # the other signals' gradients as the adjoint gives them.
    rng = np.random.default_rng(21)
    A = rng.standard_normal((30, 10)) / np.sqrt(30)
    A[:, 4] = 0.0
    d = LassoDesign(A)
    S_np = (A @ rng.standard_normal((10, 8))).T.astype(np.float32)
    warm = torch.zeros((8, 10), device=DEV)
    warm[:4, 4] = 100.0  # signals 0-3: column 4 active
    S = torch.from_numpy(S_np).to(DEV).requires_grad_(True)
    alpha = torch.full((8,), 1e-3, device=DEV, requires_grad=True)
    coef, _ = lasso(S, alpha, d, max_iter=3, tol=0.0, warm=warm)
    assert bool((coef[:4, 4] != 0).all()) and not coef[4:, 4].any()
    g = torch.ones_like(coef)
    gs, ga, st = d.adjoint(coef.detach(), g)
    assert st.tolist() == [1] * 4 + [0] * 4
    with pytest.warns(RuntimeWarning, match="not positive definite"):
        (coef * g).sum().backward()
    assert not S.grad[:4].any() and not alpha.grad[:4].any()
    assert torch.equal(S.grad[4:], gs[4:]) and torch.equal(alpha.grad[4:], ga[4:])
    assert torch.isfinite(S.grad).all() and torch.isfinite(alpha.grad).all()


def test_hypergradient_vs_grid(problem):
    A_tr, A_te, S_tr, S_te = problem
    d = LassoDesign(A_tr)
    S = torch.from_numpy(S_tr).to(DEV)
    St = torch.from_numpy(S_te).to(DEV)
    At = torch.from_numpy(A_te.astype(np.float32)).to(DEV)
    base, _ = d.forward(S, 1e-3, max_iter=500, tol=1e-6)
    mse0 = float(heldout_mse(base, At, St).mean())
    grid = grid_search(d, S, At, St, np.logspace(-5, -1, 20), max_iter=500)
    hg = hypergradient_search(d, S, At, St, alpha0=1e-3, steps=20)
    mg, mh = float(grid["mse"].mean()), float(hg["best_mse"].mean())
    print(f"held-out MSE: alpha0 {mse0:.2f}, grid {mg:.2f}, hypergradient {mh:.2f} (final {float(hg['mse'].mean()):.2f})")
    assert mh < mse0 and mh <= 1.05 * mg
