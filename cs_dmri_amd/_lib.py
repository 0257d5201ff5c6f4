"""ctypes binding of ``libshorefit.so`` (declarations in include/shorefit.h).

The library is the only compute path: importing this module raises if the
built ``libshorefit.so`` is missing, and every call raises ``ValueError`` /
``RuntimeError`` when the library reports an error.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

# SHOREFIT_LIB selects another build of the library (the host-sanitizer build)
LIB_PATH = (Path(os.environ["SHOREFIT_LIB"]) if os.environ.get("SHOREFIT_LIB")
            else Path(__file__).resolve().parent / "libshorefit.so")

OK, E_INVALID, E_FIT, E_NOT_PD, E_UNSUPPORTED, E_DEVICE = range(6)
REG = {"l1": 1, "l2": 2, "nonneg": 3}
ALPHA_MODE = {"fixed": 0, "alpha-ratio": 1, "path-bic": 2, "l2-anchored": 3}
PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2}

UNITS_UM, UNITS_MM = 0, 1

EXPORTS = (
    "shorefit_spec_default", "shorefit_fit", "shorefit_n_coeffs", "shorefit_design_matrix",
    "shorefit_fista_batch", "shorefit_tikhonov_batch", "shorefit_last_error", "shorefit_version",
    "shorefit_tournier_matrix", "shorefit_odf_sh", "shorefit_predict", "shorefit_microstructure",
    "shorefit_fit_graddev", "shorefit_graddev_variants", "shorefit_graddev_tune", "shorefit_graddev_gram",
    "shorefit_design_create", "shorefit_design_destroy", "shorefit_design_info", "shorefit_lasso_forward",
    "shorefit_lasso_adjoint", "shorefit_volume_to_voxel_major", "shorefit_volume_to_file_order",
    "shorefit_lasso_forward_noalloc", "shorefit_release_cache", "shorefit_build_info", "shorefit_set_progress",
)

PROGRESS_FN = C.CFUNCTYPE(None, C.c_int64, C.c_int64, C.c_void_p)


class ShorefitSpec(C.Structure):
    _fields_ = [
        ("radial_order", C.c_uint32), ("zeta", C.c_double), ("regularization", C.c_int32),
        ("alpha_mode", C.c_int32), ("alpha", C.c_double), ("alpha_ratio", C.c_double),
        ("path_n_alphas", C.c_uint32), ("path_eps", C.c_double), ("slack", C.c_double),
        ("max_iter", C.c_uint32), ("tol", C.c_double), ("non_negative", C.c_int32),
        ("lambda_n", C.c_double), ("lambda_l", C.c_double), ("compute_diagnostics", C.c_int32),
        ("allow_single_shell", C.c_int32), ("precision", C.c_int32), ("refresh_every", C.c_int32),
        ("reserved", C.c_int32 * 6),
    ]


class ShorefitGradients(C.Structure):
    _fields_ = [
        ("bvals", C.POINTER(C.c_double)), ("bvecs", C.POINTER(C.c_double)), ("n_grads", C.c_int64),
        ("big_delta", C.c_double), ("small_delta", C.c_double), ("b0_threshold", C.c_double),
    ]


class ShorefitVolume(C.Structure):
    _fields_ = [
        ("data", C.c_void_p), ("mask", C.c_void_p), ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("on_device", C.c_int32), ("reserved", C.c_int32),
    ]


class ShorefitOutput(C.Structure):
    _fields_ = [
        ("coefficients", C.c_void_p), ("r2", C.c_void_p), ("residual_l2", C.c_void_p),
        ("iterations", C.c_void_p), ("regularization_kind", C.c_void_p), ("alpha", C.c_void_p),
        ("bic", C.c_void_p), ("rss_l2", C.c_void_p),
        ("n_coeffs", C.c_int32), ("has_alpha_distribution", C.c_int32), ("alpha_distribution", C.c_double * 3),
        ("big_delta", C.c_double), ("small_delta", C.c_double), ("tau", C.c_double),
        ("delta_source", C.c_int32), ("n_shells", C.c_int32), ("dwi_shells", C.c_double * 32),
        ("n_warnings", C.c_int32), ("warnings", C.c_char * 2048), ("n_voxels", C.c_int64),
        ("lipschitz", C.c_double), ("fit_ms", C.c_double), ("graddev_ms", C.c_double * 3),
    ]

    def warning_list(self) -> list[str]:
        raw = bytes(self.warnings)
        parts = raw.split(b"\0")
        return [p.decode("utf-8") for p in parts[: self.n_warnings]]


class ShorefitGradDev(C.Structure):
    _fields_ = [
        ("field", C.c_void_p), ("bvecs_vox", C.POINTER(C.c_double)), ("frame_rotation", C.c_double * 9),
        ("variant", C.c_int32), ("reserved", C.c_int32),
    ]


_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)


def _torch_runtime_first() -> None:
    """Starts torch's HIP runtime, when torch and a GPU are present, before
    this library's. torch ships its own libamdhip64 beside the one the
    library links; a process whose first GPU use is the library's runtime
    then cannot start torch's ("No HIP GPUs are available"), while the other
    order works. Loading the library therefore starts torch's first."""
    try:
        import torch
    except ImportError:
        return
    try:
        if torch.cuda.is_available():
            torch.cuda.init()
    except RuntimeError:
        pass


def _load() -> C.CDLL:
    _torch_runtime_first()
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -m cs_dmri_amd.build` (hipcc, gfx950). "
            "There is no CPU fallback for the SHORE fit.")
    lib = C.CDLL(str(LIB_PATH))
    lib.shorefit_spec_default.argtypes = [C.POINTER(ShorefitSpec)]
    lib.shorefit_spec_default.restype = None
    lib.shorefit_fit.argtypes = [C.POINTER(ShorefitGradients), C.POINTER(ShorefitVolume), C.POINTER(ShorefitSpec),
                                 C.POINTER(ShorefitOutput), C.c_void_p]
    lib.shorefit_fit.restype = C.c_int
    lib.shorefit_n_coeffs.argtypes = [C.c_uint32]
    lib.shorefit_n_coeffs.restype = C.c_int
    lib.shorefit_design_matrix.argtypes = [C.POINTER(ShorefitGradients), C.c_uint32, C.c_double, _pd]
    lib.shorefit_design_matrix.restype = C.c_int
    lib.shorefit_fista_batch.argtypes = [_pd, C.c_int32, C.c_int32, _pf, C.c_int64, C.c_double, C.c_uint32,
                                         C.c_double, C.c_int32, C.c_int32, C.c_int32, _pf,
                                         C.POINTER(C.c_uint32), C.POINTER(C.c_uint8), _pf, _pd]
    lib.shorefit_fista_batch.restype = C.c_int
    lib.shorefit_tikhonov_batch.argtypes = [_pd, C.c_int32, C.c_int32, _pd, _pd, C.c_double, C.c_double, _pf,
                                            C.c_int64, _pf, _pf]
    lib.shorefit_tikhonov_batch.restype = C.c_int
    lib.shorefit_last_error.argtypes = []
    lib.shorefit_last_error.restype = C.c_char_p
    lib.shorefit_version.argtypes = []
    lib.shorefit_version.restype = C.c_char_p
    lib.shorefit_release_cache.argtypes = []
    lib.shorefit_release_cache.restype = C.c_int
    lib.shorefit_build_info.argtypes = []
    lib.shorefit_build_info.restype = C.c_char_p
    lib.shorefit_set_progress.argtypes = [PROGRESS_FN, C.c_void_p, C.c_double]
    lib.shorefit_set_progress.restype = C.c_int
    _vp = C.c_void_p
    lib.shorefit_tournier_matrix.argtypes = [C.c_uint32, C.c_double, C.c_uint32, _pd]
    lib.shorefit_tournier_matrix.restype = C.c_int
    lib.shorefit_odf_sh.argtypes = [_vp, _vp, C.c_int64, C.c_int64, C.c_int64, C.c_uint32, C.c_double, C.c_uint32,
                                    _vp, C.c_int32, _vp]
    lib.shorefit_odf_sh.restype = C.c_int
    lib.shorefit_predict.argtypes = [_vp, C.c_int64, C.c_int64, C.c_int64, C.c_uint32, C.c_double,
                                     C.POINTER(ShorefitGradients), _vp, C.c_int32, _vp]
    lib.shorefit_predict.restype = C.c_int
    lib.shorefit_microstructure.argtypes = [_vp, _vp, C.c_int64, C.c_int64, C.c_int64, C.c_uint32, C.c_double, _vp,
                                            C.c_int32, C.c_float] + [_vp] * 6 + [C.c_int32, _vp]
    lib.shorefit_microstructure.restype = C.c_int
    _pg, _pv, _pdv = C.POINTER(ShorefitGradients), C.POINTER(ShorefitVolume), C.POINTER(ShorefitGradDev)
    lib.shorefit_fit_graddev.argtypes = [_pg, _pv, _pdv, C.POINTER(ShorefitSpec), C.POINTER(ShorefitOutput), _vp]
    lib.shorefit_fit_graddev.restype = C.c_int
    _pi = C.POINTER(C.c_int32)
    lib.shorefit_graddev_variants.argtypes = [C.c_uint32, C.c_int32, _pi, _pi, _pi, _pi]
    lib.shorefit_graddev_variants.restype = C.c_int
    lib.shorefit_graddev_tune.argtypes = [_pg, _pv, _pdv, C.POINTER(ShorefitSpec), C.c_int64, _pd, _vp]
    lib.shorefit_graddev_tune.restype = C.c_int
    lib.shorefit_graddev_gram.argtypes = [_pg, _pv, _pdv, C.c_uint32, C.c_double, C.POINTER(C.c_int64), C.c_int64,
                                          _pf, _pd, _pd]
    lib.shorefit_graddev_gram.restype = C.c_int
    lib.shorefit_design_create.argtypes = [_pd, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
    lib.shorefit_design_create.restype = C.c_int
    lib.shorefit_design_destroy.argtypes = [_vp]
    lib.shorefit_design_destroy.restype = None
    lib.shorefit_design_info.argtypes = [_vp, _pi, _pi, _pd]
    lib.shorefit_design_info.restype = C.c_int
    lib.shorefit_lasso_forward.argtypes = [_vp, _vp, C.c_int64, _vp, C.c_double, _vp, C.c_uint32, C.c_double,
                                           C.c_int32, _vp, _vp, _vp]
    lib.shorefit_lasso_forward.restype = C.c_int
    lib.shorefit_lasso_forward_noalloc.argtypes = lib.shorefit_lasso_forward.argtypes
    lib.shorefit_lasso_forward_noalloc.restype = C.c_int
    lib.shorefit_lasso_adjoint.argtypes = [_vp, _vp, _vp, C.c_int64, _vp, _vp, _vp, _vp]
    lib.shorefit_lasso_adjoint.restype = C.c_int
    lib.shorefit_volume_to_voxel_major.argtypes = [_vp, C.c_int32, C.c_int64, C.c_int64, C.c_int64, C.c_int64,
                                                   C.c_float, C.c_float, _vp, C.c_int32, _vp]
    lib.shorefit_volume_to_voxel_major.restype = C.c_int
    lib.shorefit_volume_to_file_order.argtypes = [_vp, C.c_int64, C.c_int64, C.c_int64, C.c_int64, _vp, C.c_int32,
                                                  _vp]
    lib.shorefit_volume_to_file_order.restype = C.c_int
    return lib


_LIB: C.CDLL | None = None


def lib() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = _load()
    return _LIB


class ShorefitError(ValueError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


def check(code: int) -> None:
    if code != OK:
        msg = lib().shorefit_last_error().decode("utf-8", "replace")
        if code == E_DEVICE:
            raise RuntimeError(f"shorefit device error: {msg}")
        raise ShorefitError(code, msg)


def ptr(a: np.ndarray | None, t=C.c_void_p):
    if a is None:
        return None if t is C.c_void_p else C.cast(None, t)
    if t is C.c_void_p:
        return C.c_void_p(a.ctypes.data)
    return a.ctypes.data_as(t)


def gradients(bvals, bvecs, big_delta, small_delta, b0_threshold):
    """Build a ShorefitGradients; returns (struct, keep-alive arrays)."""
    b = np.ascontiguousarray(bvals, dtype=np.float64).ravel()
    v = np.ascontiguousarray(bvecs, dtype=np.float64)
    if v.ndim != 2 or v.shape[1] != 3:
        raise ValueError(f"bvecs must have shape (N, 3), got {v.shape}")
    if len(b) != v.shape[0]:
        raise ValueError(f"bvals ({len(b)}) and bvecs ({v.shape[0]}) must have the same length")
    g = ShorefitGradients(ptr(b, _pd), ptr(v, _pd), len(b),
                          float("nan") if big_delta is None else float(big_delta),
                          float("nan") if small_delta is None else float(small_delta), float(b0_threshold))
    return g, (b, v)
