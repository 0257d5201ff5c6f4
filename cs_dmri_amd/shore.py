"""Object front end of the SHORE fit: ``ShoreModel`` holds a gradient table
and the fit settings, ``ShoreModel.fit`` runs the device fit, and
``ShoreFit`` holds the fitted volume and the maps derived from it.

What is kept from the reference is its interface: the keywords and defaults
of ``ShoreModel`` (python/cs_dmri/shore.py:58-78), the signature of ``fit``
(:80-95), the names of the result's attributes and methods (:98-205, ``save``
/ ``load`` :243-294), the frame rules (python/cs_dmri/_frame.py:13-34) and the
user-facing error texts. The structure is this package's own: the settings
are one validated dataclass, the frame is a small value type, the result keeps
its per-voxel maps in one mapping, and every computation is a call into
``libshorefit.so``.
"""

from __future__ import annotations

import dataclasses
import json
import warnings
from pathlib import Path
from typing import NamedTuple

import numpy as np

from . import io as sio
from . import nifti as nf
from .models import shore_fit, shore_microstructure, shore_odf_sh, shore_predict

__all__ = ["GradientTable", "FitSettings", "ShoreModel", "ShoreFit"]


class GradientTable:
    """b-values (s/mm^2), unit directions and optional timing of a DWI
    acquisition (the fields of python/cs_dmri/gradients.py:33-80). Directions
    may be given as (N, 3) or (3, N)."""

    def __init__(self, bvals, bvecs, *, b0_threshold: float = 50.0, big_delta: float | None = None,
                 small_delta: float | None = None):
        b = np.asarray(bvals, dtype=np.float64).reshape(-1)
        g = np.asarray(bvecs, dtype=np.float64)
        if g.ndim == 2 and g.shape[0] == 3 and g.shape[1] == b.size and b.size != 3:
            g = g.T
        if g.shape != (b.size, 3):
            raise ValueError(f"bvecs must have shape ({b.size}, 3) or (3, {b.size}), got {g.shape}")
        self.bvals = np.ascontiguousarray(b)
        self.bvecs = np.ascontiguousarray(g)
        self.b0_threshold = float(b0_threshold)
        self.big_delta = big_delta
        self.small_delta = small_delta

    @classmethod
    def coerce(cls, gtab) -> "GradientTable":
        """This class, a ``(bvals, bvecs)`` pair, or any object with ``bvals``
        and ``bvecs`` (a dipy table)."""
        if isinstance(gtab, cls):
            return gtab
        if isinstance(gtab, tuple) and len(gtab) == 2:
            return cls(*gtab)
        return cls(gtab.bvals, gtab.bvecs, b0_threshold=getattr(gtab, "b0_threshold", 50.0),
                   big_delta=getattr(gtab, "big_delta", None), small_delta=getattr(gtab, "small_delta", None))


@dataclasses.dataclass(frozen=True)
class FitSettings:
    """Every solver keyword of ``ShoreModel`` with its default; validated on
    construction. ``precision`` and ``refresh_every`` select the device
    arithmetic of the L1 solve ("fp32", or "fp16" / "bf16" increments with an
    exact refresh every ``refresh_every`` iterations; None takes the
    arithmetic's default)."""

    regularization: str = "l1"
    alpha_mode: str = "l2-anchored"
    alpha: float = 1.0
    alpha_ratio: float = 1e-3
    path_n_alphas: int = 20
    path_eps: float | None = None
    slack: float = 0.05
    max_iter: int = 1000
    tol: float = 1e-6
    non_negative: bool = False
    lambda_n: float = 1e-8
    lambda_l: float = 1e-8
    nonneg_max_iter: int = 200
    nonneg_tol: float = 1e-9
    nonneg_epsilon: float = 1e-10
    allow_single_shell: bool = False
    precision: str = "fp32"
    refresh_every: int | None = None

    REGULARIZATIONS = ("l1", "l2", "nonneg")
    ALPHA_MODES = ("l2-anchored", "path-bic", "alpha-ratio", "fixed")

    def __post_init__(self):
        if self.regularization not in self.REGULARIZATIONS:
            raise ValueError(f"regularization must be one of {self.REGULARIZATIONS}, got {self.regularization!r}")
        if self.alpha_mode not in self.ALPHA_MODES:
            raise ValueError(f"alpha_mode must be one of {self.ALPHA_MODES}, got {self.alpha_mode!r}")

    def as_keywords(self) -> dict:
        return {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}


class Frame(NamedTuple):
    """The frame a fit's directions live in: ``"world"`` (RAS, through the
    affine's column-normalised rotation) or ``"image"`` (the voxel axes)."""

    name: str
    rotation: np.ndarray

    CHOICES = ("auto", "world", "image")

    @classmethod
    def resolve(cls, requested: str, affine) -> "Frame":
        if requested not in cls.CHOICES:
            raise ValueError(f"bvec_frame must be one of {cls.CHOICES}, got {requested!r}")
        world = requested == "world" or (requested == "auto" and affine is not None)
        if not world:
            return cls("image", np.eye(3))
        if affine is None:
            raise ValueError("bvec_frame='world' needs an affine (fit a DWI, or pass affine=)")
        return cls("world", sio.affine_rotation(np.asarray(affine, dtype=np.float64)))

    @property
    def sidecar_tag(self) -> str:
        return "world-ras" if self.name == "world" else "image-axis"

    def apply(self, vectors) -> np.ndarray:
        """Rotate image-axis row vectors (..., 3) into this frame."""
        return np.ascontiguousarray(np.asarray(vectors, dtype=np.float64) @ self.rotation.T)


def _volume_and_mask(data, mask, gtab: GradientTable):
    """(volume, mask, affine) from an array or a loaded ``io.DwiData``. An
    array without a mask gets the b0 fallback mask."""
    if isinstance(data, sio.DwiData):
        vol = data.data if isinstance(data.data, np.ndarray) else data.data.cpu().numpy()
        return vol, (data.mask if mask is None else _as_mask(mask, vol.shape)), data.affine
    vol = np.asarray(data, dtype=np.float32)
    if vol.ndim != 4:
        raise ValueError(f"expected a 4-D (X, Y, Z, N) array, got shape {vol.shape}")
    if mask is None:
        return vol, sio.default_mask_from_b0(vol, gtab.bvals, gtab.b0_threshold), None
    return vol, _as_mask(mask, vol.shape), None


def _as_mask(mask, shape) -> np.ndarray:
    m = np.asarray(mask)
    m = m if m.dtype == bool else m > 0
    if m.shape != tuple(shape[:3]):
        raise ValueError(f"mask shape {m.shape} does not match data spatial shape {tuple(shape[:3])}")
    return m


class ShoreModel:
    """3D-SHORE model: radial order (even; 6 gives 72 coefficients), zeta,
    the frame of the fit's directions (``bvec_frame``: "auto" is world RAS when
    an affine is known, else the voxel axes) and the solver keywords of
    ``FitSettings``. The fit refuses data with fewer than two non-zero shells
    unless ``allow_single_shell``."""

    def __init__(self, gtab, *, radial_order: int = 6, zeta: float = 700.0, bvec_frame: str = "auto", **solver):
        self.settings = FitSettings(**solver)
        self.gtab = GradientTable.coerce(gtab)
        self.radial_order = radial_order
        self.zeta = zeta
        self.bvec_frame = bvec_frame

    @property
    def regularization(self) -> str:
        return self.settings.regularization

    def fit(self, data, mask=None, *, affine=None, diagnostics: bool = True,
            n_threads: int | None = None) -> "ShoreFit":
        """Fit every voxel of ``mask`` (default: a DWI's mask, or the b0
        fallback mask of an array). ``data`` is (X, Y, Z, N) or an
        ``io.DwiData``; ``affine`` overrides the DWI's."""
        vol, m, dwi_affine = _volume_and_mask(data, mask, self.gtab)
        if affine is None:
            affine = dwi_affine
        frame = Frame.resolve(self.bvec_frame, affine)
        g = self.gtab
        result = shore_fit(vol, g.bvals, frame.apply(g.bvecs), m, big_delta=g.big_delta, small_delta=g.small_delta,
                           b0_threshold=g.b0_threshold, bvec_frame=frame.sidecar_tag,
                           radial_order=self.radial_order, zeta=self.zeta, diagnostics=diagnostics,
                           n_threads=n_threads, **self.settings.as_keywords())
        for text in result.pop("warnings", ()):
            warnings.warn(text, stacklevel=2)
        return ShoreFit(result, model=self, mask=m, affine=affine, frame=frame)


class ShoreFit:
    """A fitted volume: ``coefficients`` (X, Y, Z, K) float32, ``sidecar``
    (the fit metadata written beside a coefficient file),
    ``alpha_distribution``, ``mask``, ``affine``, ``frame``, and the per-voxel
    maps named in ``MAPS`` (``None`` when the fit did not produce them)."""

    MAPS = ("r2", "rmse", "residual_l2", "iterations", "regularization_kind", "alpha", "bic", "rss_l2")
    SAVED_MAPS = ("r2", "rmse", "alpha", "bic", "rss_l2")

    def __init__(self, result: dict, *, model=None, mask=None, affine=None, frame: Frame | str = "image"):
        self.coefficients = result["coefficients"]
        self.sidecar = result["sidecar"]
        self.alpha_distribution = result.get("alpha_distribution")
        self._maps = {k: result.get(k) for k in self.MAPS}
        self.model = model
        self.mask = np.any(self.coefficients != 0, axis=-1) if mask is None else mask
        self.affine = affine
        if isinstance(frame, str):
            frame = Frame("world", sio.affine_rotation(affine)) if frame == "world" else Frame("image", np.eye(3))
        self._frame = frame
        self.radial_order = int(self.sidecar["basis"]["radial_order"])
        self.zeta = float(self.sidecar["basis"]["zeta"])

    def __getattr__(self, name):
        maps = self.__dict__.get("_maps")
        if maps is not None and name in maps:
            return maps[name]
        raise AttributeError(f"{type(self).__name__!s} has no attribute {name!r}")

    @property
    def frame(self) -> str:
        return self._frame.name

    @property
    def dwi_shells(self) -> list[float] | None:
        s = self.sidecar.get("dwi_shells")
        return None if s is None else [float(b) for b in s]

    @property
    def single_shell(self) -> bool:
        s = self.dwi_shells
        return s is not None and len(s) < 2

    @property
    def sparsity(self) -> np.ndarray:
        """Fraction of non-zero coefficients per masked voxel (src/fit.rs:810-827)."""
        return sio.sparsity_map(np.asarray(self.coefficients), self.mask)

    def _refuse_single_shell(self):
        if self.single_shell:
            raise ValueError("propagator-derived scalars are not determined by single-shell data "
                             f"(this fit used b = {self.dwi_shells[0]:.0f} only)")

    def odf_sh(self, lmax: int | None = None, *, n_threads=None) -> np.ndarray:
        """Tournier (MRtrix) SH coefficients of the ODF, (X, Y, Z, n_sh);
        ``lmax`` defaults to the largest even value not above the radial order."""
        if lmax is None:
            lmax = self.radial_order // 2 * 2
        return shore_odf_sh(self.coefficients, self.radial_order, self.zeta, lmax, self.mask, n_threads=n_threads)

    def predict(self, gtab=None, *, clip: bool = True, n_threads=None) -> np.ndarray:
        """Signal for ``gtab`` (voxel-axis directions; default: the model's
        table), (X, Y, Z, N); negative values are clipped unless ``clip`` is
        false. A single-shell fit warns when asked for other b-values."""
        if gtab is None and self.model is None:
            raise ValueError("pass a gradient table: this fit has no model")
        table = GradientTable.coerce(self.model.gtab if gtab is None else gtab)
        if self.single_shell:
            fitted = self.dwi_shells[0]
            dw = table.bvals[table.bvals > table.b0_threshold]
            if (np.abs(dw - fitted) > 50).any():
                warnings.warn(f"this fit used a single shell (b = {fitted:.0f}); signals at other b-values are "
                              "extrapolated by the regularization, not the measurements", stacklevel=2)
        sig = shore_predict(self.coefficients, self.radial_order, self.zeta, table.bvals,
                            self._frame.apply(table.bvecs), big_delta=self.sidecar["big_delta_seconds"],
                            small_delta=self.sidecar["small_delta_seconds"], n_threads=n_threads)
        return np.maximum(sig, 0) if clip else sig

    def microstructure(self, directions=None, *, units: str = "um", outlier_factor: float | None = 10.0,
                       n_threads=None) -> dict:
        """RTOP, RTAP, RTPP, MSD, QIV and NG maps. RTAP / RTPP take a fibre
        direction per voxel (``directions`` (X, Y, Z, 3), voxel axes) and are
        NaN without one. Refused for a single-shell fit."""
        self._refuse_single_shell()
        dirs = None if directions is None else self._frame.apply(directions).astype(np.float32)
        return shore_microstructure(self.coefficients, self.radial_order, self.zeta, self.mask, dirs, units=units,
                                    outlier_factor=outlier_factor, n_threads=n_threads)

    # ------------------------------------------------------------ files --
    def _outputs(self, path: Path) -> dict:
        """Every file ``save`` writes, mapped to its array (None = sidecar)."""
        files = {path: self.coefficients, sio.sidecar_path_for(path): None}
        for name in self.SAVED_MAPS:
            if self._maps.get(name) is not None:
                files[sio.sibling_path(path, f"_{name}.nii.gz")] = self._maps[name]
        if self._maps.get("r2") is not None:
            files[sio.sibling_path(path, "_sparsity.nii.gz")] = self.sparsity
        return files

    def save(self, path, *, overwrite: bool = False) -> None:
        """Coefficient NIfTI, sidecar JSON and diagnostic maps (``_r2``,
        ``_rmse``, ``_alpha``, ``_bic``, ``_rss_l2``, ``_sparsity``) in the
        layout ``cs-fit`` writes; every path is checked before anything is
        written."""
        if self.affine is None:
            raise ValueError("saving needs an affine")
        path = Path(path)
        files = self._outputs(path)
        if not overwrite:
            taken = [str(p) for p in files if p.exists()]
            if taken:
                raise FileExistsError(f"refusing to overwrite {taken}; pass overwrite=True")
        grid = nf.new_header(self.coefficients.shape[:3], self.affine)
        sio.write_coefficients(path, self.coefficients, grid, self.sidecar, overwrite=True)
        for p, arr in files.items():
            if arr is not None and p != path:
                sio.write_map(p, arr, grid, overwrite=True)

    @classmethod
    def load(cls, path, mask=None) -> "ShoreFit":
        """A coefficient file written by ``save`` or ``cs-fit``, with the maps
        found beside it."""
        path = Path(path)
        coef, sidecar = sio.read_coefficients(path)
        result = {"coefficients": coef, "sidecar": sidecar}
        for name in cls.SAVED_MAPS:
            p = sio.sibling_path(path, f"_{name}.nii.gz")
            if p.exists():
                result[name] = sio.to_voxel_major(nf.load(p))[..., 0]
        affine = nf.load(path).affine()
        frame = "world" if sidecar.get("bvec_frame") == "world-ras" else "image"
        return cls(result, mask=mask, affine=affine, frame=frame)

    def __repr__(self) -> str:
        n = int(np.count_nonzero(self.mask))
        extra = ""
        if self._maps.get("r2") is not None:
            extra = f" r2_mean={float(np.mean(self._maps['r2'][self.mask])):.3f}"
        return (f"<ShoreFit order={self.radial_order} K={self.coefficients.shape[-1]} voxels={n} "
                f"frame={self.frame}{extra}>")
