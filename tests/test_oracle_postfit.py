"""Pinning the oracle's post-fit functions (src/odf.rs, src/scalars/microstructure.rs,
src/io/microstructure.rs, src/synth.rs) on the CPU.

* SHORE -> Tournier SH projection: the reference's own KATs (scipy radial
  factors, index examples, shape, the l = 0 rows) and the reference's
  per-column identity <M e_k, Y_T(u)> = radial(n, l) Y_BS[k](u), with the
  Tournier SH evaluated here independently from scipy.special.lpmv.
* hyp2f1 / binomial against scipy.special.
* RTOP, QIV, RTAP, RTPP against quadrature of the fitted signal E(q) itself
  (RTOP = int E d^3q, QIV = 1 / int E q^2 d^3q, RTAP = int over the plane
  normal to u of E d^2q, RTPP = int along u of E dq) and MSD against the
  finite-difference Laplacian -lap E(0) / (4 pi^2): the physical definitions
  the closed forms stand for.
* NG closed cases, the alpha factor KAT, zero-direction NaN, the rejection
  rule (src/io/microstructure.rs:303-326) and the unit factors.
* Synthesis: the target design matrix (pinned in test_oracle_basis) times c.
"""

import math

import numpy as np
import pytest
from scipy import integrate, special

ZETA = 700.0


def _tournier_sh(lmax, theta, phi):
    # mrtrix / Tournier real SH, index l(l+1)/2 + m (src/odf.rs:120-146 conventions)
    out = np.zeros((lmax + 1) * (lmax + 2) // 2)
    for ell in range(0, lmax + 1, 2):
        for m in range(-ell, ell + 1):
            am = abs(m)
            norm = math.sqrt((2 * ell + 1) / (4 * math.pi) * math.factorial(ell - am) / math.factorial(ell + am))
            p = norm * special.lpmv(am, ell, math.cos(phi))
            if m < 0:
                v = math.sqrt(2) * p * math.sin(am * theta)
            elif m == 0:
                v = p
            else:
                v = math.sqrt(2) * p * math.cos(am * theta)
            out[ell * (ell + 1) // 2 + m] = v
    return out


def test_radial_factor_scipy_values(orc):
    # src/odf.rs:159-180
    cases = [(0, 0, 2.766998550654e-2), (2, 0, 3.788868806943e-2), (2, 2, -4.286615722265e-2),
             (4, 0, 4.340694526519e-2), (4, 2, -4.391643925880e-2), (4, 4, 5.400628174864e-2),
             (6, 0, 4.738453459969e-2), (6, 2, -4.722289470498e-2), (6, 4, 4.949860498183e-2),
             (6, 6, -6.322725024349e-2)]
    for n, ell, want in cases:
        assert orc.odf_radial_factor(n, ell, ZETA) == pytest.approx(want, abs=1e-10)


def test_brainsuite_to_tournier_examples(orc):
    # src/odf.rs:191-203
    for k, idx, sign in [(0, 5, 1.0), (1, 4, 1.0), (2, 3, 1.0), (3, 2, 1.0), (4, 1, -1.0)]:
        assert orc.brainsuite_to_tournier(2, k) == (idx, sign)


def test_projection_shape_and_l0_rows(orc):
    M = orc.tournier_matrix(6, ZETA, 6)
    assert M.shape == (28, 72)
    idx = np.array([(n, l) for n in range(7) for l in range(0, n + 1, 2) for _ in range(2 * l + 1)])
    for col in np.nonzero(idx[:, 1] == 0)[0]:
        assert M[0, col] != 0 and not M[1:, col].any()


@pytest.mark.parametrize("order,lmax", [(6, 6), (8, 8), (8, 4)])
def test_projection_matches_direct_sh_evaluation(orc, order, lmax):
    M = orc.tournier_matrix(order, ZETA, lmax)
    dirs = [(0.3, 0.7), (1.4, 0.5), (-1.1, 1.2), (2.7, 1.5), (0.0, 0.0)]
    col = 0
    for n in range(order + 1):
        for ell in range(0, n + 1, 2):
            bs = 2 * ell + 1
            rad = orc.odf_radial_factor(n, ell, ZETA) if ell <= lmax else 0.0
            for theta, phi in dirs:
                yt = _tournier_sh(lmax, theta, phi)
                ybs = orc.brainsuite_sh_block(ell, theta, phi)
                for k in range(bs):
                    assert M[:, col + k] @ yt == pytest.approx(rad * ybs[k], abs=1e-12)
            col += bs


def test_hyp2f1_and_binomial_against_scipy(orc):
    for n in range(0, 11):
        assert orc.hyp2f1(-n, 2.5, 1.5, 2.0) == pytest.approx(special.hyp2f1(-n, 2.5, 1.5, 2.0), rel=1e-12, abs=1e-12)
        for ell in range(0, n + 1, 2):
            a, b, c = -n + ell, ell / 2 + 1.5, ell + 1.5
            assert orc.hyp2f1(a, b, c, 2.0) == pytest.approx(special.hyp2f1(a, b, c, 2.0), rel=1e-12, abs=1e-12)
    for n in (0.5, 2.5, 7.5, 11.5):
        for k in range(0, 6):
            assert orc.binomial_float(n, k) == pytest.approx(special.binom(n, k), rel=1e-12)


def _coefs(orc, order, seed):
    # a physically shaped coefficient vector: the SHORE fit of a two-fibre signal
    from cs_dmri_amd import workloads as W
    gt = W.cartesian_table(120, seed=seed)
    S = W.multi_tensor_signal(gt, 1, 0.0, seed=seed + 1)[0]
    A = orc.design_matrix(order, ZETA, gt.bvals, gt.bvecs, gt.big_delta, gt.small_delta)
    dn, dl = orc.regularization(order)
    c, _ = orc.tikhonov(A, S.astype(np.float64), dn, dl, 1e-4, 1e-4)
    return c


def _signal(orc, order, c, qvecs):
    # E(q) = sum_i c_i R_i(|q|) Y_i(q/|q|) with q in 1/mm: tau = 1/(4 pi^2) makes q = sqrt(b)
    q = np.linalg.norm(qvecs, axis=1)
    b = q * q
    dirs = np.where(q[:, None] > 0, qvecs / np.maximum(q[:, None], 1e-300), [[0.0, 0.0, 1.0]])
    A = orc.design_matrix(order, ZETA, b, dirs, 1 / (4 * math.pi ** 2), 0.0, b0_threshold=-1.0)
    return A @ c


def _e0(orc, order, c):
    return float(_signal(orc, order, c, np.zeros((1, 3)))[0])


@pytest.mark.parametrize("order,seed", [(4, 11), (6, 12), (8, 13)])
def test_rtop_qiv_msd_against_integrals(orc, order, seed):
    c = _coefs(orc, order, seed)
    raw = orc.microstructure_raw(order, ZETA, c)
    e0 = _e0(orc, order, c)
    # spherical mean of E(q) over directions: only l = 0 survives; use a 2-point
    # radial quadrature along +-z averaged with +-x, +-y is unnecessary for l = 0,
    # so integrate 4 pi q^2 E0(q) with E0(q) the l = 0 part at |q|.
    K = len(c)
    l0 = np.zeros(K)
    col = 0
    for n in range(order + 1):
        for ell in range(0, n + 1, 2):
            if ell == 0:
                l0[col] = c[col]
            col += 2 * ell + 1

    def E0(q):
        return float(_signal(orc, order, l0, np.array([[0.0, 0.0, q]]))[0]) / e0

    lim = 40 * math.sqrt(ZETA)
    rtop_q = integrate.quad(lambda q: 4 * math.pi * q * q * E0(q), 0, lim, limit=400, epsabs=0, epsrel=1e-11)[0]
    assert raw[0] == pytest.approx(max(rtop_q, 0.0), rel=1e-7)
    m2 = integrate.quad(lambda q: 4 * math.pi * q ** 4 * E0(q), 0, lim, limit=400, epsabs=0, epsrel=1e-11)[0]
    assert raw[4] == pytest.approx(1.0 / m2, rel=1e-7)
    # MSD = -lap E(0) / (4 pi^2); for a radial f(q) = f0 + a q^2 + ..., lap f(0) = 6 a
    hs = [0.05 * math.sqrt(ZETA) / 2 ** k for k in range(5)]
    est = [(E0(h) - 1.0) / (h * h) for h in hs]
    for _ in range(3):  # Richardson on the h^2 error term
        est = [(4 * est[i + 1] - est[i]) / 3 for i in range(len(est) - 1)]
    a = est[-1]
    assert raw[3] == pytest.approx(max(-6 * a / (4 * math.pi ** 2), 0.0), rel=1e-6)


@pytest.mark.parametrize("order,seed", [(4, 21), (6, 22)])
def test_rtap_rtpp_against_integrals(orc, order, seed):
    c = _coefs(orc, order, seed)
    e0 = _e0(orc, order, c)
    u = np.array([0.3, -0.5, 0.81])
    u /= np.linalg.norm(u)
    raw = orc.microstructure_raw(order, ZETA, c, u)
    # RTPP: E along the axis through the origin
    lim = 40 * math.sqrt(ZETA)

    def along(t):
        return float(_signal(orc, order, c, (t * u)[None, :])[0]) / e0

    rtpp_q = integrate.quad(along, -lim, lim, limit=400, epsabs=0, epsrel=1e-11, points=[0.0])[0]
    assert raw[2] == pytest.approx(max(rtpp_q, 0.0), rel=1e-7)
    # RTAP: E over the plane normal to u (polar coordinates; periodic angle by the
    # trapezoid rule, which is spectrally accurate for these trigonometric polynomials)
    e1 = np.cross(u, [1.0, 0.0, 0.0])
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(u, e1)
    nb = 4 * (order + 2)
    betas = 2 * math.pi * np.arange(nb) / nb
    ring = np.cos(betas)[:, None] * e1 + np.sin(betas)[:, None] * e2

    def ring_mean(rho):
        return float(np.mean(_signal(orc, order, c, rho * ring))) / e0

    rtap_q = integrate.quad(lambda r: 2 * math.pi * r * ring_mean(r), 0, lim, limit=400, epsabs=0, epsrel=1e-11)[0]
    assert raw[1] == pytest.approx(max(rtap_q, 0.0), rel=1e-7)


def test_ng_alpha_factor_and_direction_cases(orc):
    K = orc.n_coeffs(4)
    c = np.zeros(K)
    c[0] = 1.0
    assert orc.microstructure_raw(4, ZETA, c)[5] == pytest.approx(0.0, abs=1e-12)  # microstructure.rs:338-345
    c[1] = 1.0  # two equal coefficients: 1 - c0^2 / (c0^2 + c1^2) = 1/2
    assert orc.microstructure_raw(4, ZETA, c)[5] == pytest.approx(math.sqrt(0.5), rel=1e-14)
    assert orc.alpha_factor(0, 0, ZETA) == pytest.approx(math.sqrt(2 * math.pi * ZETA ** 1.5 * special.gamma(1.5)),
                                                         abs=1e-9)
    raw = orc.microstructure_raw(4, ZETA, np.ones(K), np.zeros(3))  # microstructure.rs:355-361
    assert np.isnan(raw[1]) and np.isnan(raw[2])
    assert np.isnan(orc.microstructure_raw(4, ZETA, np.zeros(K))).all()  # E(0) = 0


def test_compute_microstructure_units_and_rejection(orc):
    rng = np.random.default_rng(5)
    base = _coefs(orc, 4, 31).astype(np.float32)
    V = 200
    C = np.repeat(base[None, :], V, 0) * (1 + 0.02 * rng.standard_normal((V, 1))).astype(np.float32)
    C[17, 1:] *= 40.0  # a degenerate fit: huge higher-order terms
    dirs = rng.standard_normal((V, 3)).astype(np.float32)
    dirs[3] = 0.0
    dirs[5, 1] = np.nan
    mm = orc.compute_microstructure(4, ZETA, C, dirs, units="mm", outlier_factor=None)
    um = orc.compute_microstructure(4, ZETA, C, dirs, units="um", outlier_factor=None)
    for name, f in zip(("rtop", "rtap", "rtpp", "msd", "qiv", "ng"), (1e-9, 1e-6, 1e-3, 1e6, 1e15, 1.0)):
        np.testing.assert_array_equal(um[name], mm[name] * np.float32(f))
    for v in (3, 5):  # missing peak (zero or non-finite direction)
        assert np.isnan(mm["rtap"][v]) and np.isnan(mm["rtpp"][v])
    rej = orc.compute_microstructure(4, ZETA, C, dirs, units="mm", outlier_factor=10.0)
    for name in ("rtop", "rtap", "rtpp", "msd", "qiv", "ng"):
        x = mm[name]
        fin = np.sort(x[np.isfinite(x)])
        if not len(fin):
            continue
        p99 = fin[min(int(math.floor(0.99 * (len(fin) - 1) + 0.5)), len(fin) - 1)]
        want = np.where(np.isfinite(x) & (x > np.float32(10.0) * p99), np.nan, x)
        np.testing.assert_array_equal(rej[name], want)


def test_synthesize_is_design_times_coefficients(orc):
    from cs_dmri_amd import workloads as W
    gt = W.cartesian_table(40, seed=9)
    A = orc.design_matrix(6, ZETA, gt.bvals, gt.bvecs, gt.big_delta, gt.small_delta)
    rng = np.random.default_rng(3)
    C = rng.standard_normal((5, A.shape[1])).astype(np.float32)
    C[2] = 0.0
    out = orc.synthesize(A, C)
    np.testing.assert_allclose(out, (C.astype(np.float64) @ A.T).astype(np.float32), rtol=1e-6, atol=1e-7)
    assert not out[2].any()
