"""The precision report's measuring instruments (CPU): the 362-point sphere
and the ODF peak finder."""

import math

import numpy as np

from cs_dmri_amd.precision import _neighbours, _tournier_basis, find_peaks, icosphere, odf_values


def test_icosphere_362():
    v, t = icosphere(6)
    assert v.shape == (362, 3) and t.shape == (720, 3)
    np.testing.assert_allclose(np.linalg.norm(v, axis=1), 1.0, rtol=1e-12)
    edges = {tuple(sorted(e)) for a, b, c in t for e in ((a, b), (b, c), (a, c))}
    assert len(v) - len(edges) + len(t) == 2  # Euler characteristic of the sphere
    # antipodally symmetric: every vertex's opposite is a vertex
    assert max(np.min(np.linalg.norm(v + u, axis=1)) for u in v) < 1e-9
    # nearly uniform: neighbour spacing within 20% of the mean
    d = [np.linalg.norm(v[a] - v[b]) for a, b in edges]
    assert max(d) / min(d) < 1.4


def _gauss_sphere(n=24):
    # product rule: Gauss-Legendre in cos(polar) x uniform azimuth (exact for
    # polynomials of degree < 2n in each variable)
    x, w = np.polynomial.legendre.leggauss(n)
    az = 2 * math.pi * np.arange(2 * n) / (2 * n)
    ct, a = np.meshgrid(x, az, indexing="ij")
    st = np.sqrt(1 - ct ** 2)
    pts = np.stack([st * np.cos(a), st * np.sin(a), ct], -1).reshape(-1, 3)
    wts = (np.repeat(w, 2 * n) * (2 * math.pi / (2 * n)))
    return pts, wts


def test_tournier_basis_is_orthonormal_on_the_sphere():
    pts, wts = _gauss_sphere()
    Y = _tournier_basis(8, pts)
    G = (Y * wts[:, None]).T @ Y
    even = [i for i in range(Y.shape[1]) if np.any(Y[:, i] != 0)]
    assert len(even) == sum(2 * l + 1 for l in range(0, 9, 2))
    np.testing.assert_allclose(G[np.ix_(even, even)], np.eye(len(even)), atol=1e-12)


def test_peaks_of_two_fibre_odf():
    v, t = icosphere(6)
    nb = _neighbours(len(v), t)
    u1 = np.array([1.0, 0.2, 0.1])
    u1 /= np.linalg.norm(u1)
    u2 = np.array([0.1, 1.0, -0.3])
    u2 /= np.linalg.norm(u2)
    pts, wts = _gauss_sphere(32)
    f = np.exp(20 * ((pts @ u1) ** 2 - 1)) + 0.8 * np.exp(20 * ((pts @ u2) ** 2 - 1))
    Y = _tournier_basis(8, pts)
    sh = (Y * wts[:, None]).T @ f  # projection onto the orthonormal SH
    p = find_peaks(odf_values(sh, 8, v), v, nb)
    assert len(p) == 2
    ang = lambda a, b: math.degrees(math.acos(min(1.0, abs(float(a @ b)))))
    # within the mesh spacing (edges 9-12 degrees) plus the l <= 8 truncation shift
    assert ang(p[0], u1) < 8 and ang(p[1], u2) < 8
    assert len(find_peaks(np.zeros(len(v)), v, nb)) == 0
