"""Precision report of the device SHORE fit against an FP64 reference fit.

For each device arithmetic (FP32; BF16 / FP16 MFMA increments with an exact
refresh every k iterations) this compares, voxel by voxel, with the FP64 fit
of the same data:

* coefficients (relative L2 error),
* the fitted signal A c on the fit's own gradient table,
* RTOP, MSD and QIV (src/scalars/microstructure.rs:79-179, mm units),
* ODF peaks on a 362-point sphere (a frequency-6 geodesic icosphere): the
  angle between the principal peaks and the agreement of the peak counts.
  The reference extracts peaks on a 321-vertex hemisphere of an external
  crate; this sphere is a measuring instrument for the error, not a parity
  target.

``report`` takes the FP64 coefficients from the caller (the oracle in the
tests and the benchmark), so this module does not depend on them.
"""

from __future__ import annotations

import math

import numpy as np

from .models import shore_fit, shore_microstructure, shore_odf_sh, shore_predict

__all__ = ["icosphere", "odf_values", "find_peaks", "report"]


def icosphere(frequency: int = 6):
    """Vertices (V, 3) and triangles (F, 3) of a geodesic icosphere: each
    icosahedron face split into frequency^2 triangles and projected to the
    unit sphere. frequency 6 gives 10 f^2 + 2 = 362 vertices."""
    t = (1 + 5 ** 0.5) / 2
    base = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                     [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float)
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
             (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    verts, index, tris = [], {}, []

    def vid(p):
        p = p / np.linalg.norm(p)
        key = tuple(np.round(p, 9))
        if key not in index:
            index[key] = len(verts)
            verts.append(p)
        return index[key]

    f = frequency
    for a, b, c in faces:
        A, B, Cc = base[a], base[b], base[c]
        grid = {}
        for i in range(f + 1):
            for j in range(f + 1 - i):
                k = f - i - j
                grid[i, j] = vid((i * A + j * B + k * Cc) / f)
        for i in range(f):
            for j in range(f - i):
                tris.append((grid[i, j], grid[i + 1, j], grid[i, j + 1]))
                if i + j < f - 1:
                    tris.append((grid[i + 1, j], grid[i + 1, j + 1], grid[i, j + 1]))
    return np.array(verts), np.array(tris)


def _tournier_basis(lmax: int, dirs: np.ndarray) -> np.ndarray:
    """Tournier (MRtrix) real SH at unit vectors, index l(l+1)/2 + m."""
    from scipy import special
    x, y, z = dirs[:, 0], dirs[:, 1], dirs[:, 2]
    theta = np.arctan2(y, x)
    phi = np.arccos(np.clip(z, -1, 1))
    out = np.zeros((len(dirs), (lmax + 1) * (lmax + 2) // 2))
    for ell in range(0, lmax + 1, 2):
        for m in range(-ell, ell + 1):
            am = abs(m)
            norm = math.sqrt((2 * ell + 1) / (4 * math.pi) * math.factorial(ell - am) / math.factorial(ell + am))
            p = norm * special.lpmv(am, ell, np.cos(phi))
            if m < 0:
                v = math.sqrt(2) * p * np.sin(am * theta)
            elif m == 0:
                v = p
            else:
                v = math.sqrt(2) * p * np.cos(am * theta)
            out[:, ell * (ell + 1) // 2 + m] = v
    return out


def odf_values(sh: np.ndarray, lmax: int, verts: np.ndarray) -> np.ndarray:
    """ODF amplitudes (..., V) from Tournier SH coefficients (..., n_sh)."""
    return sh @ _tournier_basis(lmax, verts).T


def _neighbours(nv: int, tris: np.ndarray):
    nb = [set() for _ in range(nv)]
    for a, b, c in tris:
        nb[a] |= {b, c}
        nb[b] |= {a, c}
        nb[c] |= {a, b}
    return [np.array(sorted(s)) for s in nb]


def find_peaks(values: np.ndarray, verts: np.ndarray, nbrs, npeaks: int = 5, relative_threshold: float = 0.5,
               min_separation_deg: float = 25.0):
    """Local maxima of one ODF on the sphere mesh, strongest first, above
    relative_threshold x the maximum, antipodal duplicates and peaks closer
    than min_separation_deg to a stronger one removed."""
    vmax = values.max()
    if not np.isfinite(vmax) or vmax <= 0:
        return np.zeros((0, 3))
    cand = [i for i in range(len(values)) if values[i] >= values[nbrs[i]].max() and values[i] >= relative_threshold * vmax]
    cand.sort(key=lambda i: -values[i])
    cos_sep = math.cos(math.radians(min_separation_deg))
    out = []
    for i in cand:
        u = verts[i]
        if all(abs(float(u @ w)) < cos_sep for w in out):
            out.append(u)
        if len(out) == npeaks:
            break
    return np.array(out)


def _angle_deg(u, w):
    return math.degrees(math.acos(min(1.0, abs(float(np.dot(u, w))))))


def _stats(x):
    x = np.asarray(x, float)
    x = x[np.isfinite(x)]
    if not len(x):
        return {"median": None, "p99": None, "max": None}
    return {"median": float(np.median(x)), "p99": float(np.quantile(x, 0.99)), "max": float(x.max())}


def report(data, mask, gtab, reference_coefficients, *, radial_order=6, zeta=700.0, alpha=1e-3, max_iter=500,
           tol=1e-6, settings=(("fp32", 10), ("bf16", 5), ("bf16", 10), ("bf16", 20), ("fp16", 5), ("fp16", 10),
                                ("fp16", 20))) -> dict:
    """Errors of the device fits against reference_coefficients (the FP64 fit
    of the same data, (X, Y, Z, K)), one entry per (precision, refresh)."""
    lmax = radial_order - radial_order % 2
    verts, tris = icosphere(6)
    nbrs = _neighbours(len(verts), tris)
    mask = np.asarray(mask, bool)
    ref = np.asarray(reference_coefficients, np.float32)
    kw = dict(big_delta=gtab.big_delta, small_delta=gtab.small_delta)

    def derived(C):
        sig = shore_predict(C, radial_order, zeta, gtab.bvals, gtab.bvecs, **kw)[mask]
        ms = shore_microstructure(C, radial_order, zeta, mask, None, units="mm", outlier_factor=None)
        sh = shore_odf_sh(C, radial_order, zeta, lmax, mask)[mask]
        return sig, ms, sh

    rsig, rms, rsh = derived(ref)
    rodf = odf_values(rsh.astype(np.float64), lmax, verts)
    rpk = [find_peaks(v, verts, nbrs) for v in rodf]
    out = {"voxels": int(mask.sum()), "sphere_vertices": int(len(verts)), "entries": []}
    for prec, k in settings:
        fit = shore_fit(data, gtab.bvals, gtab.bvecs, mask, radial_order=radial_order, zeta=zeta,
                        alpha_mode="fixed", alpha=alpha, max_iter=max_iter, tol=tol, precision=prec,
                        refresh_every=k, **kw)
        C = fit["coefficients"]
        sig, ms, sh = derived(C)
        c, cr = C[mask].astype(np.float64), ref[mask].astype(np.float64)
        cerr = np.linalg.norm(c - cr, axis=1) / np.maximum(np.linalg.norm(cr, axis=1), 1e-30)
        serr = np.linalg.norm(sig - rsig, axis=1) / np.maximum(np.linalg.norm(rsig, axis=1), 1e-30)
        scal = {}
        for name in ("rtop", "msd", "qiv"):
            a, b = ms[name][mask].astype(np.float64), rms[name][mask].astype(np.float64)
            scal[name] = _stats(np.abs(a - b) / np.maximum(np.abs(b), 1e-30))
        odf = odf_values(sh.astype(np.float64), lmax, verts)
        angles, same_count = [], 0
        for v, p_ref in zip(odf, rpk):
            p = find_peaks(v, verts, nbrs)
            same_count += int(len(p) == len(p_ref))
            if len(p) and len(p_ref):
                angles.append(_angle_deg(p[0], p_ref[0]))
        out["entries"].append({
            "precision": prec, "refresh_every": k if prec != "fp32" else None,
            "coefficients": _stats(cerr), "fitted_signal": _stats(serr), **scal,
            "principal_peak_angle_deg": _stats(angles),
            "same_peak_count_frac": same_count / max(len(rpk), 1),
        })
    return out
