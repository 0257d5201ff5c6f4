"""``ShoreModel`` / ``ShoreFit``: the Python front end of the SHORE fit, with
the constructor keywords, ``fit`` signature, frame handling and result
attributes of the reference's python/cs_dmri/shore.py:25-137.
"""

from __future__ import annotations

import warnings

import numpy as np

from .models import shore_fit, shore_microstructure, shore_odf_sh, shore_predict

__all__ = ["GradientTable", "ShoreModel", "ShoreFit"]

_REGULARIZATIONS = ("l1", "l2", "nonneg")
_ALPHA_MODES = ("l2-anchored", "path-bic", "alpha-ratio", "fixed")
_FRAMES = ("auto", "world", "image")


class GradientTable:
    """b-values and directions (python/cs_dmri/gradients.py:33-80)."""

    def __init__(self, bvals, bvecs, *, b0_threshold: float = 50.0, big_delta: float | None = None,
                 small_delta: float | None = None):
        bvals = np.ascontiguousarray(np.asarray(bvals, dtype=np.float64).ravel())
        bvecs = np.asarray(bvecs, dtype=np.float64)
        n = len(bvals)
        if bvecs.shape == (3, n) and n != 3:
            bvecs = bvecs.T
        if bvecs.shape != (n, 3):
            raise ValueError(f"bvecs must have shape ({n}, 3) or (3, {n}), got {bvecs.shape}")
        self.bvals = bvals
        self.bvecs = np.ascontiguousarray(bvecs)
        self.b0_threshold = float(b0_threshold)
        self.big_delta = big_delta
        self.small_delta = small_delta


def as_gradient_table(gtab) -> GradientTable:
    if isinstance(gtab, GradientTable):
        return gtab
    if isinstance(gtab, tuple) and len(gtab) == 2:
        return GradientTable(*gtab)
    return GradientTable(gtab.bvals, gtab.bvecs, b0_threshold=getattr(gtab, "b0_threshold", 50.0),
                         big_delta=getattr(gtab, "big_delta", None), small_delta=getattr(gtab, "small_delta", None))


def affine_rotation(affine) -> np.ndarray:
    """Column-normalised 3x3 of an affine (src/qspace.rs:49-60)."""
    a = np.asarray(affine, dtype=np.float64)
    if a.shape != (4, 4):
        raise ValueError(f"affine must be 4x4, got {a.shape}")
    r = a[:3, :3].copy()
    n = np.linalg.norm(r, axis=0)
    return r * np.where(n > 0, 1.0 / np.where(n > 0, n, 1.0), 0.0)[None, :]


def fit_frame(bvec_frame: str, affine):
    """python/cs_dmri/_frame.py:13-29."""
    if bvec_frame not in _FRAMES:
        raise ValueError(f"bvec_frame must be one of {_FRAMES}, got {bvec_frame!r}")
    if bvec_frame == "auto":
        bvec_frame = "world" if affine is not None else "image"
    if bvec_frame == "world":
        if affine is None:
            raise ValueError("bvec_frame='world' needs an affine (fit a DWI, or pass affine=)")
        return "world", affine_rotation(affine)
    return "image", np.eye(3)


def rotate(vectors, R) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(vectors, dtype=np.float64) @ R.T)


class ShoreModel:
    """3D-SHORE model; keywords and defaults as python/cs_dmri/shore.py:58-78,
    plus ``precision`` / ``refresh_every`` for the device arithmetic."""

    def __init__(self, gtab, *, radial_order: int = 6, zeta: float = 700.0, regularization: str = "l1",
                 alpha_mode: str = "l2-anchored", alpha: float = 1.0, alpha_ratio: float = 1e-3,
                 path_n_alphas: int = 20, path_eps: float | None = None, slack: float = 0.05,
                 max_iter: int = 1000, tol: float = 1e-6, non_negative: bool = False,
                 lambda_n: float = 1e-8, lambda_l: float = 1e-8, nonneg_max_iter: int = 200,
                 nonneg_tol: float = 1e-9, nonneg_epsilon: float = 1e-10, bvec_frame: str = "auto",
                 allow_single_shell: bool = False, precision: str = "fp32", refresh_every: int = 10):
        if regularization not in _REGULARIZATIONS:
            raise ValueError(f"regularization must be one of {_REGULARIZATIONS}, got {regularization!r}")
        if alpha_mode not in _ALPHA_MODES:
            raise ValueError(f"alpha_mode must be one of {_ALPHA_MODES}, got {alpha_mode!r}")
        self.gtab = as_gradient_table(gtab)
        self.radial_order = radial_order
        self.zeta = zeta
        self.regularization = regularization
        self.bvec_frame = bvec_frame
        self._solver = dict(
            alpha_mode=alpha_mode, alpha=alpha, alpha_ratio=alpha_ratio, path_n_alphas=path_n_alphas,
            path_eps=path_eps, slack=slack, max_iter=max_iter, tol=tol, non_negative=non_negative,
            lambda_n=lambda_n, lambda_l=lambda_l, nonneg_max_iter=nonneg_max_iter, nonneg_tol=nonneg_tol,
            nonneg_epsilon=nonneg_epsilon, allow_single_shell=allow_single_shell, precision=precision,
            refresh_every=refresh_every)

    def fit(self, data, mask=None, *, affine=None, diagnostics: bool = True, n_threads: int | None = None):
        """Fit every voxel in ``mask`` (python/cs_dmri/shore.py:80-95)."""
        vol = np.asarray(data, dtype=np.float32)
        if vol.ndim != 4:
            raise ValueError(f"expected a 4-D (X, Y, Z, N) array, got shape {vol.shape}")
        if mask is None:
            raise ValueError("a mask is required")
        m = np.asarray(mask)
        if m.dtype != bool:
            m = m > 0
        frame, R = fit_frame(self.bvec_frame, None if affine is None else np.asarray(affine, np.float64))
        out = shore_fit(
            vol, self.gtab.bvals, rotate(self.gtab.bvecs, R), m,
            big_delta=self.gtab.big_delta, small_delta=self.gtab.small_delta,
            b0_threshold=self.gtab.b0_threshold,
            bvec_frame="world-ras" if frame == "world" else "image-axis",
            radial_order=self.radial_order, zeta=self.zeta, regularization=self.regularization,
            diagnostics=diagnostics, n_threads=n_threads, **self._solver)
        for message in out.pop("warnings", ()):
            warnings.warn(message, stacklevel=2)
        return ShoreFit(out, model=self, mask=m, affine=affine, frame=frame, rotation=R)


class ShoreFit:
    """Result of :meth:`ShoreModel.fit` (python/cs_dmri/shore.py:98-137)."""

    def __init__(self, arrays: dict, *, model=None, mask=None, affine=None, frame="image", rotation=None):
        self.coefficients = arrays["coefficients"]
        self.sidecar = arrays["sidecar"]
        self.alpha_distribution = arrays.get("alpha_distribution")
        for k in ("r2", "rmse", "residual_l2", "iterations", "regularization_kind", "alpha", "bic", "rss_l2"):
            setattr(self, k, arrays.get(k))
        self.model = model
        self.mask = mask if mask is not None else np.any(self.coefficients != 0, axis=-1)
        self.affine = affine
        self.frame = frame
        self._rotation = np.eye(3) if rotation is None else rotation
        basis = self.sidecar["basis"]
        self.radial_order = int(basis["radial_order"])
        self.zeta = float(basis["zeta"])

    @property
    def dwi_shells(self):
        shells = self.sidecar.get("dwi_shells")
        return None if shells is None else [float(b) for b in shells]

    @property
    def single_shell(self) -> bool:
        return self.dwi_shells is not None and len(self.dwi_shells) < 2

    @property
    def sparsity(self) -> np.ndarray:
        out = np.zeros(self.coefficients.shape[:3], np.float32)
        out[self.mask] = (self.coefficients[self.mask] != 0).mean(-1)
        return out

    def odf_sh(self, lmax: int | None = None, *, n_threads=None) -> np.ndarray:
        """Tournier (MRtrix) SH coefficients of the ODF, (X, Y, Z, n_sh)
        (python/cs_dmri/shore.py:152-157); ``lmax`` defaults to the largest
        even value <= the radial order."""
        lmax = self.radial_order - self.radial_order % 2 if lmax is None else lmax
        return shore_odf_sh(self.coefficients, self.radial_order, self.zeta, lmax, self.mask, n_threads=n_threads)

    def predict(self, gtab=None, *, clip: bool = True, n_threads=None) -> np.ndarray:
        """Signal for ``gtab`` (default: the fit's own), (X, Y, Z, N), as in
        python/cs_dmri/shore.py:159-181: the table is in the image's voxel-axis
        frame and is rotated into the fit frame."""
        if gtab is None:
            if self.model is None:
                raise ValueError("pass a gradient table: this fit has no model")
            gtab = self.model.gtab
        gtab = as_gradient_table(gtab)
        if self.single_shell:
            fitted = self.dwi_shells[0]
            dw = gtab.bvals[gtab.bvals > gtab.b0_threshold]
            if np.any(np.abs(dw - fitted) > 50):
                warnings.warn(
                    f"this fit used a single shell (b = {fitted:.0f}); signals at other b-values are "
                    "extrapolated by the regularization, not the measurements", stacklevel=2)
        out = shore_predict(self.coefficients, self.radial_order, self.zeta, gtab.bvals,
                            rotate(gtab.bvecs, self._rotation), big_delta=self.sidecar["big_delta_seconds"],
                            small_delta=self.sidecar["small_delta_seconds"], n_threads=n_threads)
        return np.maximum(out, 0) if clip else out

    def microstructure(self, directions=None, *, units: str = "um", outlier_factor: float | None = 10.0,
                       n_threads=None) -> dict:
        """RTOP, RTAP, RTPP, MSD, QIV and NG maps (python/cs_dmri/shore.py:183-205).
        RTAP / RTPP need a fibre direction per voxel, ``directions`` (X, Y, Z, 3)
        in the image's voxel-axis frame; without them they are NaN."""
        if self.single_shell:
            raise ValueError(
                "propagator-derived scalars are not determined by single-shell data "
                f"(this fit used b = {self.dwi_shells[0]:.0f} only)")
        if directions is not None:
            directions = rotate(directions, self._rotation).astype(np.float32)
        return shore_microstructure(self.coefficients, self.radial_order, self.zeta, self.mask, directions,
                                    units=units, outlier_factor=outlier_factor, n_threads=n_threads)

    def __repr__(self) -> str:
        r2 = "" if self.r2 is None else f", mean R² {float(self.r2[self.mask].mean()):.3f}"
        return f"ShoreFit(radial_order={self.radial_order}, {int(self.mask.sum())} voxels{r2}, frame={self.frame})"
