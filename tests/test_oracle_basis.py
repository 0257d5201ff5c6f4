"""Pins the oracle's special functions, gradient table and SHORE design matrix.

Independent checks: the reference's own KATs (src/math.rs:168-247,
src/basis/shore.rs:205-273, src/qspace.rs:230-370) and scipy.special
(genlaguerre, lpmv, gamma), which evaluate the same functions by other
algorithms.
"""

import math

import numpy as np
import pytest
from scipy import special

from tests.fixtures import shell_gtab


def test_factorial_and_gamma_kats(orc):
    assert orc.factorial(0) == 1.0 and orc.factorial(5) == 120.0 and orc.factorial(10) == 3628800.0
    for n in range(1, 11):
        assert abs(orc.gamma(n) - orc.factorial(n - 1)) < 1e-6
    assert orc.factorial(25) == pytest.approx(math.factorial(25), rel=1e-12)
    for x in (0.3, 1.5, 2.5, 7.5, 10.5, 17.25):
        assert orc.gamma(x) == pytest.approx(special.gamma(x), rel=1e-13)


def test_legendre_kats(orc):
    assert orc.assoc_legendre(0, 0, 0.3) == pytest.approx(1.0, abs=1e-12)
    assert orc.assoc_legendre(1, 0, 0.3) == pytest.approx(0.3, abs=1e-12)
    assert orc.assoc_legendre(1, 1, 0.3) == pytest.approx(-math.sqrt(1 - 0.09), abs=1e-12)
    assert orc.assoc_legendre(2, 0, 0.4) == pytest.approx(0.5 * (3 * 0.16 - 1), abs=1e-12)
    assert orc.assoc_legendre(2, 2, 0.4) == pytest.approx(3 * (1 - 0.16), abs=1e-12)
    for l in range(0, 11):
        for m in range(0, l + 1):
            for x in (-0.9, -0.2, 0.0, 0.35, 0.99):
                ref = special.lpmv(m, l, x)  # Condon-Shortley phase included
                assert orc.assoc_legendre(l, m, x) == pytest.approx(ref, rel=1e-11, abs=1e-12)


def test_laguerre_kats(orc):
    x, a = 0.7, 0.5
    assert orc.gen_laguerre(0, a, x) == 1.0
    assert orc.gen_laguerre(1, a, x) == pytest.approx(1 + a - x, abs=1e-12)
    assert orc.gen_laguerre(2, a, x) == pytest.approx((1 + a) * (2 + a) / 2 - (2 + a) * x + x * x / 2, abs=1e-12)
    for n in range(0, 11):
        for a in (0.5, 2.5, 4.5, 10.5):
            for x in (0.0, 0.3, 2.0, 7.1):
                assert orc.gen_laguerre(n, a, x) == pytest.approx(special.eval_genlaguerre(n, a, x), rel=1e-11,
                                                                  abs=1e-11)


def test_coefficient_counts(orc):
    # src/basis/shore.rs:211-219; the radial orders 6, 8 and 10 of the fits.
    assert orc.n_coeffs(6) == 72 and orc.n_coeffs(2) == 8
    assert orc.n_coeffs(8) == 145 and orc.n_coeffs(10) == 256
    idx = orc.shore_indices(4)
    assert [tuple(r) for r in idx[:3]] == [(0, 0, 0), (1, 0, 0), (2, 0, 0)]
    assert tuple(idx[3]) == (2, 2, -2) and tuple(idx[7]) == (2, 2, 2)


def test_brainsuite_sh_kats(orc):
    v = orc.brainsuite_sh_block(0, 0.3, 0.7)
    assert v.shape == (1,) and v[0] == pytest.approx(1 / math.sqrt(4 * math.pi), abs=1e-12)
    v = orc.brainsuite_sh_block(2, 0.4, 1.1)
    c = math.cos(1.1)
    assert v.shape == (5,)
    assert v[2] == pytest.approx(math.sqrt(5 / (4 * math.pi)) * 0.5 * (3 * c * c - 1), abs=1e-10)


def _sh_block_scipy(ell, theta, phi):
    """BrainSuite real SH block from scipy.special.lpmv, following the ordering
    documented at src/basis/shore.rs:94-98."""
    out = np.zeros(2 * ell + 1)
    p = [math.sqrt((2 * ell + 1) / (4 * math.pi) * math.factorial(ell - k) / math.factorial(ell + k))
         * special.lpmv(k, ell, math.cos(phi)) for k in range(ell + 1)]
    for k in range(1, ell + 1):
        out[ell - k] = math.sqrt(2) * p[k] * math.cos(k * theta)
        out[ell + k] = math.sqrt(2) * (1.0 if k % 2 == 1 else -1.0) * p[k] * math.sin(k * theta)
    out[ell] = p[0]
    return out


def design_matrix_scipy(order, zeta, bvals, bvecs, big_delta, small_delta, b0=50.0):
    tau = big_delta - small_delta / 3.0
    q = np.where(bvals <= b0, 0.0, np.sqrt(np.maximum(bvals, 0) / (4 * math.pi ** 2 * tau)))
    cols = []
    for n in range(order + 1):
        for ell in range(0, n + 1, 2):
            kap = math.sqrt(2 * math.factorial(n - ell) / (zeta ** 1.5 * special.gamma(n + 1.5)))
            block = np.zeros((len(bvals), 2 * ell + 1))
            for i, (qi, g) in enumerate(zip(q, bvecs)):
                x, y, z = g * qi
                r = math.sqrt(x * x + y * y + z * z)
                th = math.atan2(y, x)
                ph = math.acos(max(-1.0, min(1.0, z / r))) if r > 0 else 0.0
                u = r * r / zeta
                rad = kap * special.eval_genlaguerre(n - ell, ell + 0.5, u) * math.exp(-u / 2) * u ** (ell / 2)
                block[i] = rad * _sh_block_scipy(ell, th, ph)
            cols.append(block)
    return np.hstack(cols)


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_design_matrix_matches_scipy(orc, order):
    bvals, bvecs = shell_gtab(12)
    A = orc.design_matrix(order, 700.0, bvals, bvecs, 0.04, 0.012)
    ref = design_matrix_scipy(order, 700.0, bvals, bvecs, 0.04, 0.012)
    assert A.shape == ref.shape == (len(bvals), orc.n_coeffs(order))
    np.testing.assert_allclose(A, ref, rtol=1e-10, atol=1e-13 * np.abs(ref).max())


def test_design_matrix_q0_pattern(orc):
    # src/basis/shore.rs:238-273: a b0 row is zero outside l = 0 and
    # A(0, 0) = kappa(zeta, 0, 0) / sqrt(4 pi).
    bvecs = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 0], [1, 0, 1], [0, 1, 1]], float)
    n = np.maximum(np.linalg.norm(bvecs, axis=1), 1e-12)
    bvecs = bvecs / n[:, None]
    bvals = np.array([0, 1000, 1000, 1000, 2000, 2000, 2000], float)
    A = orc.design_matrix(4, 700.0, bvals, bvecs, 0.043, 0.011)
    idx = orc.shore_indices(4)
    assert A.shape == (7, orc.n_coeffs(4))
    for col, (_, ell, _) in enumerate(idx):
        if ell > 0:
            assert abs(A[0, col]) < 1e-12
    assert A[0, 0] == pytest.approx(orc.kappa(700.0, 0, 0) / math.sqrt(4 * math.pi), abs=1e-10)


def test_qvals_and_deltas(orc):
    # src/qspace.rs:234-282
    q = orc.qvals([0.0, 1000.0], 0.043, 0.011)
    assert q[0] == 0.0 and q[1] > 0.0
    tau = 0.043 - 0.011 / 3
    assert q[1] == pytest.approx(math.sqrt(1000.0 / (4 * math.pi ** 2 * tau)), rel=1e-15)
    big, small = orc.estimate_deltas_tortoise([0.0, 3000.0], 80e-3)
    temp = 3000.0 / 267.51532e6 / 267.51532e6 / 80e-3 / 80e-3 / 2.0 * 1e6
    assert small == pytest.approx(temp ** (1 / 3) * 1000 * 1e-3, abs=1e-15)
    assert big == pytest.approx(3 * small, rel=1e-12)
    assert 5e-3 < small < 30e-3


def test_tortoise_deltas_physics(orc):
    """TORTOISE's (Delta, delta) heuristic (src/qspace.rs:209-227) pinned by the
    relation it encodes rather than by its own arithmetic: in SI units
    (b in s/m^2 = 1e6 x s/mm^2, gamma in rad/s/T, G in T/m) the estimate
    satisfies b = 2 gamma^2 G^2 delta^3 with Delta = 3 delta; it scales as
    b^(1/3) and G^(-2/3); b = 3000 s/mm^2 at 80 mT/m gives 14.85 ms (the
    reference test's 13-15 ms range, src/qspace.rs:251-266); with no
    diffusion weighting it falls back to delta = 1 ms, Delta = 3 ms."""
    gamma = 267.51532e6  # src/qspace.rs:18
    for b in (300.0, 1000.0, 3000.0, 10000.0):
        for g in (40e-3, 80e-3, 300e-3):
            big, small = orc.estimate_deltas_tortoise([0.0, b, 0.5 * b], g)
            assert 2.0 * gamma ** 2 * g ** 2 * small ** 3 == pytest.approx(b * 1e6, rel=1e-13)
            assert big == pytest.approx(3.0 * small, rel=1e-15)
    s1 = orc.estimate_deltas_tortoise([1000.0], 80e-3)[1]
    assert orc.estimate_deltas_tortoise([8000.0], 80e-3)[1] == pytest.approx(2.0 * s1, rel=1e-13)
    assert orc.estimate_deltas_tortoise([1000.0], 640e-3)[1] == pytest.approx(s1 / 4.0, rel=1e-13)
    # only the largest b-value counts
    assert orc.estimate_deltas_tortoise([0.0, 5.0, 1000.0, 200.0], 80e-3)[1] == s1
    small = orc.estimate_deltas_tortoise([0.0, 3000.0], 80e-3)[1]
    assert small == pytest.approx(14.8504e-3, abs=1e-7)
    assert orc.estimate_deltas_tortoise([0.0, 0.0], 80e-3) == pytest.approx((3e-3, 1e-3), rel=1e-15)


def test_shells(orc):
    # src/qspace.rs:318-354 and src/fit.rs:615-617
    s = orc.dwi_shells([0.0, 1000.0, 1010.0, 990.0, 0.0, 3000.0, 3000.0])
    assert len(s) == 2 and abs(s[0] - 1000.0) < 10 and abs(s[1] - 3000.0) < 1e-9
    assert orc.dwi_shells([1000.0, 1000.0, 3000.0]) == [1000.0, 3000.0]
    assert orc.dwi_shells([0.0, 10.0]) == []


def test_effective_table_kats(orc):
    # src/graddev.rs:218-253
    lps = np.array([[-1.0, 0, 0], [0, -1.0, 0], [0, 0, 1.0]])
    bvals = np.array([0.0, 1000.0, 1000.0])
    gv = np.array([[0, 0, 0], [1.0, 0, 0], [0, 1.0, 0]])
    gn = gv @ lps.T
    b, d = orc.effective_table(np.eye(3), lps, bvals, gv, gn)
    np.testing.assert_array_equal(b, bvals)
    np.testing.assert_allclose(d, gn, atol=1e-12)
    T = np.diag([1.1, 1.0, 1.0])
    b, d = orc.effective_table(T, lps, bvals, gv, gn)
    assert b[0] == 0.0 and abs(b[1] - 1210.0) < 1e-9 and abs(b[2] - 1000.0) < 1e-9
    assert abs(d[1, 0] + 1.0) < 1e-12 and abs(d[2, 1] + 1.0) < 1e-12
    T = np.array([[1.0, 0, 0], [0.1, 1.0, 0], [0, 0, 1.0]])
    b, d = orc.effective_table(T, lps, bvals, gv, gn)
    q = np.array([0.1, 1.0, 0.0])
    assert abs(b[2] - 1000.0 * q @ q) < 1e-9
    np.testing.assert_allclose(d[2], lps @ (q / np.linalg.norm(q)), atol=1e-12)
