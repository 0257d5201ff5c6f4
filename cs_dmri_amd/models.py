"""``shore_fit``: the SHORE-fit entry point with the keyword set of the
reference's ``_cs_dmri.shore_fit`` (python/src/models.rs:70-198), computed by
``libshorefit.so`` on the GPU, plus the solver-level helpers the reference's
tests drive (FistaSolver / TikhonovSolver / ShoreBasis::design_matrix).
"""

from __future__ import annotations

import contextlib
import ctypes as C

import numpy as np

from . import _lib as L

TORTOISE_DEFAULT_GMAX = 80e-3


def release_device_cache() -> None:
    """Return the library's cached device blocks of this thread to the driver
    (e.g. when torch's allocator runs short of memory)."""
    L.check(L.lib().shorefit_release_cache())


def n_coeffs(radial_order: int) -> int:
    k = L.lib().shorefit_n_coeffs(int(radial_order))
    if k < 0:
        raise ValueError(f"radial_order must be a positive even number, got {radial_order}")
    return k


def design_matrix(bvals, bvecs, *, radial_order=6, zeta=700.0, big_delta=None, small_delta=None,
                  b0_threshold=50.0) -> np.ndarray:
    """ShoreBasis::design_matrix (src/basis/shore.rs:122-178), FP64, evaluated on the device."""
    g, keep = L.gradients(bvals, bvecs, big_delta, small_delta, b0_threshold)
    A = np.zeros((g.n_grads, n_coeffs(radial_order)), dtype=np.float64)
    L.check(L.lib().shorefit_design_matrix(C.byref(g), int(radial_order), float(zeta), L.ptr(A, L._pd)))
    return A


def fista_batch(A, signals, alpha, *, max_iter=1000, tol=1e-6, non_negative=False, precision="fp32",
                refresh_every=None):
    """FistaSolver::fit (src/solver/fista.rs:38-150) for every row of ``signals`` against design ``A``.

    Returns ``(coef (V,K) f32, dict(iterations, converged, residual_l2, lipschitz))``.
    """
    A = np.ascontiguousarray(A, dtype=np.float64)
    S = np.ascontiguousarray(np.atleast_2d(signals), dtype=np.float32)
    N, K = A.shape
    if S.shape[1] != N:
        raise ValueError(f"signals have {S.shape[1]} samples but the design has {N} rows")
    V = S.shape[0]
    coef = np.zeros((V, K), np.float32)
    its = np.zeros(V, np.uint32)
    conv = np.zeros(V, np.uint8)
    res = np.zeros(V, np.float32)
    lip = C.c_double()
    L.check(L.lib().shorefit_fista_batch(
        L.ptr(A, L._pd), N, K, L.ptr(S, L._pf), V, float(alpha), int(max_iter), float(tol), int(non_negative),
        L.PRECISION[precision], 0 if refresh_every is None else int(refresh_every), L.ptr(coef, L._pf), its.ctypes.data_as(C.POINTER(C.c_uint32)),
        conv.ctypes.data_as(C.POINTER(C.c_uint8)), L.ptr(res, L._pf), C.byref(lip)))
    return coef, {"iterations": its, "converged": conv.astype(bool), "residual_l2": res, "lipschitz": lip.value}


def tikhonov_batch(A, signals, d_n=None, d_l=None, lambda_n=0.0, lambda_l=0.0):
    """TikhonovSolver::fit (src/solver/tikhonov.rs:26-76) for every row of ``signals``."""
    A = np.ascontiguousarray(A, dtype=np.float64)
    S = np.ascontiguousarray(np.atleast_2d(signals), dtype=np.float32)
    N, K = A.shape
    V = S.shape[0]
    dn = None if d_n is None else np.ascontiguousarray(d_n, dtype=np.float64)
    dl = None if d_l is None else np.ascontiguousarray(d_l, dtype=np.float64)
    coef = np.zeros((V, K), np.float32)
    res = np.zeros(V, np.float32)
    L.check(L.lib().shorefit_tikhonov_batch(L.ptr(A, L._pd), N, K, L.ptr(dn, L._pd), L.ptr(dl, L._pd),
                                            float(lambda_n), float(lambda_l), L.ptr(S, L._pf), V,
                                            L.ptr(coef, L._pf), L.ptr(res, L._pf)))
    return coef, res


def _spec(radial_order, zeta, regularization, alpha_mode, alpha, alpha_ratio, path_n_alphas, path_eps, slack,
          max_iter, tol, non_negative, lambda_n, lambda_l, diagnostics, allow_single_shell, precision,
          refresh_every) -> L.ShorefitSpec:
    if regularization not in L.REG:
        raise ValueError(f"regularization must be l1, l2 or nonneg, got {regularization!r}")
    if regularization == "l1" and alpha_mode not in L.ALPHA_MODE:
        raise ValueError(f"alpha_mode must be fixed, alpha-ratio, path-bic or l2-anchored, got {alpha_mode!r}")
    if precision not in L.PRECISION:
        raise ValueError(f"precision must be one of {tuple(L.PRECISION)}, got {precision!r}")
    s = L.ShorefitSpec()
    L.lib().shorefit_spec_default(C.byref(s))
    s.radial_order = int(radial_order)
    s.zeta = float(zeta)
    s.regularization = L.REG[regularization]
    s.alpha_mode = L.ALPHA_MODE.get(alpha_mode, 0)
    s.alpha = float(alpha)
    s.alpha_ratio = float(alpha_ratio)
    s.path_n_alphas = int(path_n_alphas)
    s.path_eps = 0.0 if path_eps is None else float(path_eps)
    s.slack = float(slack)
    s.max_iter = int(max_iter)
    s.tol = float(tol)
    s.non_negative = int(bool(non_negative))
    s.lambda_n = float(lambda_n)
    s.lambda_l = float(lambda_l)
    s.compute_diagnostics = int(bool(diagnostics))
    s.allow_single_shell = int(bool(allow_single_shell))
    s.precision = L.PRECISION[precision]
    s.refresh_every = 0 if refresh_every is None else int(refresh_every)
    return s


def _sidecar(out: L.ShorefitOutput, spec: L.ShorefitSpec, regularization, alpha_mode, alpha, alpha_ratio,
             path_n_alphas, path_eps, slack, max_iter, tol, non_negative, lambda_n, lambda_l, bvec_frame,
             n_coeffs_: int, radial_order, zeta, alpha_distribution) -> dict:
    """SidecarMetadata dict (src/io/coeffs.rs:28-88), as the binding returns it."""
    if regularization == "l1":
        eps = path_eps if path_eps is not None else (1e-4 if alpha_mode == "l2-anchored" else 1e-3)
        strat = {
            "fixed": {"kind": "fixed", "alpha": alpha},
            "alpha-ratio": {"kind": "alpha-max-ratio", "ratio": alpha_ratio},
            "path-bic": {"kind": "path-bic", "path": {"n": path_n_alphas, "eps": eps}},
            "l2-anchored": {"kind": "path-l2-anchored", "path": {"n": path_n_alphas, "eps": eps}, "slack": slack},
        }[alpha_mode]
        if alpha_mode == "fixed":
            chosen = {"scope": "global", "alpha": alpha}
        else:
            med, p10, p90 = alpha_distribution
            chosen = {"scope": "per-voxel", "median": med, "p10": p10, "p90": p90}
        solver = {"kind": "fista", "alpha_strategy": strat, "chosen_alpha": chosen, "non_negative": non_negative,
                  "max_iter": max_iter, "tol": tol}
    else:
        solver = {"kind": "tikhonov", "lambda_n": lambda_n, "lambda_l": lambda_l}
    return {
        "basis": {"name": "shore", "radial_order": radial_order, "zeta": zeta},
        "big_delta_seconds": out.big_delta,
        "small_delta_seconds": out.small_delta,
        "tau_seconds": out.tau,
        "delta_source": "estimated_tortoise" if out.delta_source else "user_supplied",
        "gmax_tesla_per_meter": TORTOISE_DEFAULT_GMAX,
        "solver": solver,
        "n_coefficients": n_coeffs_,
        "bvec_frame": bvec_frame,
        "dwi_shells": [float(out.dwi_shells[i]) for i in range(min(out.n_shells, 32))],
    }


def _is_device_tensor(x) -> bool:
    """True for a torch tensor on a GPU (the fit then runs on it in place)."""
    return hasattr(x, "data_ptr") and bool(getattr(x, "is_cuda", False))


def shore_fit(data, bvals, bvecs, mask, *, big_delta=None, small_delta=None, b0_threshold=50.0,
              bvec_frame="image-axis", radial_order=6, zeta=700.0, regularization="l1", alpha_mode="l2-anchored",
              alpha=1.0, alpha_ratio=1e-3, path_n_alphas=20, path_eps=None, slack=0.05, max_iter=1000, tol=1e-6,
              non_negative=False, lambda_n=1e-8, lambda_l=1e-8, nonneg_max_iter=200, nonneg_tol=1e-9,
              nonneg_epsilon=1e-10, diagnostics=False, allow_single_shell=False, n_threads=None,
              precision="fp32", refresh_every=None, graddev=None, bvecs_vox=None, frame_rotation=None,
              graddev_variant=-1, progress=None, progress_interval=1.0) -> dict:
    """Drop-in for ``_cs_dmri.shore_fit`` (python/src/models.rs:70-198).

    Same keywords, defaults, validation messages (``ValueError``) and returned
    keys: ``coefficients`` (X,Y,Z,K) float32, ``warnings``, ``sidecar``,
    ``alpha_distribution`` and, with ``diagnostics``, ``rmse``, ``r2``,
    ``residual_l2``, ``iterations``, ``regularization_kind`` (+ ``alpha``,
    ``bic``, ``rss_l2`` where the solver has them). ``nonneg_*`` and
    ``n_threads`` are accepted for signature parity; ``precision`` and
    ``refresh_every`` select the device arithmetic.

    ``graddev`` (X,Y,Z,3,3) or (X,Y,Z,9), a row-major gradient-deviation
    tensor T per voxel with g_eff = T^T g (src/graddev.rs:1-24), fits every
    voxel whose T is not the identity with its own effective table and design
    (src/graddev.rs:160-185). ``bvecs_vox`` are the voxel-axis directions
    (default ``bvecs``) and ``frame_rotation`` R maps them to ``bvecs``
    (default identity). L1 with ``alpha_mode`` fixed or alpha-ratio only.

    ``progress(done, total)`` is called with the number of masked voxels
    finished: 0 when the solve starts, every ``progress_interval`` seconds
    while it runs, and ``total`` at the end (the ``on_voxel`` hook of
    src/fit.rs:69-73 behind cs-fit's heartbeat, polled on this thread).
    """
    if bvec_frame not in ("image-axis", "world-ras"):
        raise ValueError(f"bvec_frame must be 'image-axis' or 'world-ras', got {bvec_frame!r}")
    spec = _spec(radial_order, zeta, regularization, alpha_mode, alpha, alpha_ratio, path_n_alphas, path_eps,
                 slack, max_iter, tol, non_negative, lambda_n, lambda_l, diagnostics, allow_single_shell,
                 precision, refresh_every)
    on_dev = _is_device_tensor(data)
    if on_dev:
        import torch
        vol = data
        if vol.dim() != 4:
            raise ValueError(f"data must be 4-D (X, Y, Z, N), got shape {tuple(vol.shape)}")
        vol = vol.to(torch.float32).contiguous()
        dev = vol.device
    else:
        vol = np.asarray(data)
        if vol.ndim != 4:
            raise ValueError(f"data must be 4-D (X, Y, Z, N), got shape {vol.shape}")
        vol = np.ascontiguousarray(vol, dtype=np.float32)
    g, keep = L.gradients(bvals, bvecs, big_delta, small_delta, b0_threshold)
    if vol.shape[3] != g.n_grads:
        raise ValueError(f"data has {vol.shape[3]} volumes but the gradient table has {g.n_grads}")
    if on_dev:
        m = mask if _is_device_tensor(mask) else torch.from_numpy(np.asarray(mask))
        if tuple(m.shape) != tuple(vol.shape[:3]):
            raise ValueError(f"mask shape {tuple(m.shape)} does not match data spatial shape {tuple(vol.shape[:3])}")
        m = m.to(device=dev, dtype=torch.uint8).contiguous()
    else:
        m = np.asarray(mask)
        if m.shape != vol.shape[:3]:
            raise ValueError(f"mask shape {m.shape} does not match data spatial shape {vol.shape[:3]}")
        m = np.ascontiguousarray(m.astype(np.uint8))
    nx, ny, nz, _ = vol.shape
    K = n_coeffs(radial_order) if radial_order % 2 == 0 and radial_order > 0 else 0
    shp = (nx, ny, nz)

    def zeros(shape, dt):
        if on_dev:
            tdt = {np.float32: torch.float32, np.uint32: torch.int32, np.uint8: torch.uint8}[dt]
            return torch.zeros(shape, dtype=tdt, device=dev)
        return np.zeros(shape, dt)

    def addr(a):
        return a.data_ptr() if on_dev else a.ctypes.data

    coefs = zeros(shp + (K,), np.float32)
    diag = bool(diagnostics)
    l1 = regularization == "l1"
    r2 = zeros(shp, np.float32) if diag else None
    res = zeros(shp, np.float32) if diag else None
    its = zeros(shp, np.uint32) if diag else None
    kind = zeros(shp, np.uint8) if diag else None
    amap = zeros(shp, np.float32) if (diag and l1) else None
    bic = zeros(shp, np.float32) if (diag and l1) else None
    # the Tikhonov reference RSS map of the L2-anchored strategy (src/fit.rs:384-386)
    rss_l2 = zeros(shp, np.float32) if (diag and l1 and alpha_mode == "l2-anchored") else None
    out = L.ShorefitOutput()
    out.coefficients = addr(coefs)
    for name, arr in (("r2", r2), ("residual_l2", res), ("iterations", its), ("regularization_kind", kind),
                      ("alpha", amap), ("bic", bic), ("rss_l2", rss_l2)):
        setattr(out, name, None if arr is None else addr(arr))
    v = L.ShorefitVolume(addr(vol), addr(m), nx, ny, nz, 1 if on_dev else 0, 0)
    stream = torch.cuda.current_stream(dev).cuda_stream if on_dev else None
    if graddev is not None and on_dev:
        raise ValueError("graddev fits take host arrays")
    with _progress_hook(progress, progress_interval):
        if graddev is None:
            L.check(L.lib().shorefit_fit(C.byref(g), C.byref(v), C.byref(spec), C.byref(out), stream))
        else:
            gd, _keep_gd = _graddev_struct(graddev, (nx, ny, nz), g.n_grads, bvecs_vox, frame_rotation,
                                           graddev_variant)
            L.check(L.lib().shorefit_fit_graddev(C.byref(g), C.byref(v), C.byref(gd), C.byref(spec),
                                                 C.byref(out), None))
    dist = tuple(out.alpha_distribution) if out.has_alpha_distribution else None
    result = {
        "coefficients": coefs,
        "warnings": out.warning_list(),
        "sidecar": _sidecar(out, spec, regularization, alpha_mode, alpha, alpha_ratio, path_n_alphas, path_eps,
                            slack, max_iter, tol, bool(non_negative), lambda_n, lambda_l, bvec_frame, K,
                            radial_order, zeta, dist),
        "alpha_distribution": dist,
    }
    if diag:
        # src/fit.rs:810-822 — rmse = residual_l2 / sqrt(n_grads)
        denom = np.float32(np.sqrt(np.float32(g.n_grads)))
        result["rmse"] = res / float(denom) if on_dev else (res / denom).astype(np.float32)
        result["r2"] = r2
        result["residual_l2"] = res
        result["iterations"] = its
        result["regularization_kind"] = kind
        if l1:
            result["alpha"] = amap
            result["bic"] = bic
        if rss_l2 is not None:
            result["rss_l2"] = rss_l2
    return result


@contextlib.contextmanager
def _progress_hook(progress, interval):
    """Registers ``progress`` with the library for the calls inside the block."""
    if progress is None:
        yield
        return
    if not callable(progress):
        raise ValueError("progress must be callable(done, total)")
    interval = float(interval)
    if not interval > 0.0:
        raise ValueError(f"progress_interval must be positive, got {interval}")

    def relay(done, total, _user):
        progress(int(done), int(total))

    fn = L.PROGRESS_FN(relay)  # held here until the library lets go of it
    L.check(L.lib().shorefit_set_progress(fn, None, interval))
    try:
        yield
    finally:
        L.lib().shorefit_set_progress(L.PROGRESS_FN(), None, 0.0)


def _graddev_struct(graddev, shape, n_grads, bvecs_vox=None, frame_rotation=None, variant=-1):
    f = np.asarray(graddev)
    if f.shape[:3] != tuple(shape) or f.shape[3:] not in ((3, 3), (9,)):
        raise ValueError(f"graddev must have shape {tuple(shape)} + (3, 3) or (9,), got {f.shape}")
    f = np.ascontiguousarray(f.reshape(tuple(shape) + (9,)), dtype=np.float32)
    gv = None
    if bvecs_vox is not None:
        gv = np.ascontiguousarray(bvecs_vox, dtype=np.float64)
        if gv.shape != (n_grads, 3):
            raise ValueError(f"bvecs_vox must have shape ({n_grads}, 3), got {gv.shape}")
    R = np.eye(3) if frame_rotation is None else np.asarray(frame_rotation, dtype=np.float64)
    if R.shape != (3, 3):
        raise ValueError(f"frame_rotation must be 3x3, got {R.shape}")
    gd = L.ShorefitGradDev()
    gd.field = f.ctypes.data
    gd.bvecs_vox = L.ptr(gv, L._pd)
    gd.frame_rotation[:] = [float(x) for x in R.ravel()]
    gd.variant = int(variant)
    return gd, (f, gv)


def graddev_variants(radial_order) -> list[dict]:
    """Tile-shape variants of the gradient-deviation kernels at radial_order."""
    out = []
    n = L.lib().shorefit_graddev_variants(int(radial_order), -1, None, None, None, None)
    for i in range(n):
        v = [C.c_int32() for _ in range(4)]
        L.lib().shorefit_graddev_variants(int(radial_order), i, *[C.byref(x) for x in v])
        out.append({"kp": v[0].value, "rows": v[1].value, "gram_waves": v[2].value, "fista_waves": v[3].value})
    return out


def graddev_tune(data, bvals, bvecs, mask, graddev, *, big_delta=None, small_delta=None, b0_threshold=50.0,
                 radial_order=6, zeta=700.0, alpha_mode="fixed", alpha=1e-3, alpha_ratio=1e-3, max_iter=1000,
                 tol=1e-6, bvecs_vox=None, frame_rotation=None, max_voxels=16384) -> dict:
    """Autotune the gradient-deviation kernels on (a sample of) this volume:
    returns {"variant": chosen index, "ms": per-variant device time}; later
    fits with the same radial order and gradient count use the choice."""
    vol = np.ascontiguousarray(data, dtype=np.float32)
    g, _keep = L.gradients(bvals, bvecs, big_delta, small_delta, b0_threshold)
    m = np.ascontiguousarray(np.asarray(mask).astype(np.uint8))
    nx, ny, nz, _ = vol.shape
    spec = _spec(radial_order, zeta, "l1", alpha_mode, alpha, alpha_ratio, 20, None, 0.05, max_iter, tol, False,
                 1e-8, 1e-8, False, False, "fp32", 10)
    gd, _keep_gd = _graddev_struct(graddev, (nx, ny, nz), g.n_grads, bvecs_vox, frame_rotation)
    n = L.lib().shorefit_graddev_variants(int(radial_order), -1, None, None, None, None)
    ms = np.zeros(max(n, 1))
    v = L.ShorefitVolume(vol.ctypes.data, m.ctypes.data, nx, ny, nz, 0, 0)
    rc = L.lib().shorefit_graddev_tune(C.byref(g), C.byref(v), C.byref(gd), C.byref(spec), int(max_voxels),
                                       L.ptr(ms, L._pd), None)
    if rc < 0:
        L.check(-rc)
    return {"variant": rc, "ms": ms[:n].tolist()}


def graddev_gram(data, bvals, bvecs, graddev, voxels, *, big_delta=None, small_delta=None, b0_threshold=50.0,
                 radial_order=6, zeta=700.0, bvecs_vox=None, frame_rotation=None):
    """Per-voxel G = A_v^T A_v (float32), b = A_v^T s and Lipschitz constants
    of the gradient-deviation design for linear voxel indices ``voxels``."""
    vol = np.ascontiguousarray(data, dtype=np.float32)
    g, _keep = L.gradients(bvals, bvecs, big_delta, small_delta, b0_threshold)
    nx, ny, nz, _ = vol.shape
    gd, _keep_gd = _graddev_struct(graddev, (nx, ny, nz), g.n_grads, bvecs_vox, frame_rotation)
    vox = np.ascontiguousarray(voxels, dtype=np.int64).ravel()
    K = n_coeffs(radial_order)
    G = np.zeros((len(vox), K, K), np.float32)
    b = np.zeros((len(vox), K))
    lip = np.zeros(len(vox))
    v = L.ShorefitVolume(vol.ctypes.data, None, nx, ny, nz, 0, 0)
    L.check(L.lib().shorefit_graddev_gram(C.byref(g), C.byref(v), C.byref(gd), int(radial_order), float(zeta),
                                          vox.ctypes.data_as(C.POINTER(C.c_int64)), len(vox), L.ptr(G, L._pf),
                                          L.ptr(b, L._pd), L.ptr(lip, L._pd)))
    return G, b, lip


def _coef_volume(coefficients, radial_order):
    c = np.ascontiguousarray(np.asarray(coefficients), dtype=np.float32)
    if c.ndim != 4:
        raise ValueError(f"coefficients must be 4-D (X, Y, Z, K), got shape {c.shape}")
    K = n_coeffs(radial_order)
    if c.shape[3] != K:
        raise ValueError(f"coefficients have {c.shape[3]} channels but a radial-order-{radial_order} SHORE basis "
                         f"has {K}")
    return c


def tournier_matrix(radial_order, zeta, lmax) -> np.ndarray:
    """SHORE -> Tournier SH projection (src/odf.rs:79-102), (n_sh, K) float64."""
    K = n_coeffs(radial_order)
    M = np.zeros(((lmax + 1) * (lmax + 2) // 2, K))
    if lmax % 2 != 0 or lmax > radial_order:
        raise ValueError(f"lmax must be even and ≤ radial_order ({radial_order}), got {lmax}")
    L.check(L.lib().shorefit_tournier_matrix(radial_order, zeta, lmax, L.ptr(M, L._pd)))
    return M


def shore_odf_sh(coefficients, radial_order, zeta, lmax, mask=None, *, n_threads=None) -> np.ndarray:
    """Drop-in for ``_cs_dmri.shore_odf_sh`` (python/src/models.rs:203-248)."""
    if lmax % 2 != 0 or lmax > radial_order:
        raise ValueError(f"lmax must be even and ≤ radial_order ({radial_order}), got {lmax}")
    c = _coef_volume(coefficients, radial_order)
    nx, ny, nz, _ = c.shape
    m = None if mask is None else np.ascontiguousarray(np.asarray(mask), dtype=np.uint8)
    if m is not None and m.shape != (nx, ny, nz):
        raise ValueError("mask does not match the coefficient grid")
    out = np.zeros((nx, ny, nz, (lmax + 1) * (lmax + 2) // 2), np.float32)
    L.check(L.lib().shorefit_odf_sh(c.ctypes.data, None if m is None else m.ctypes.data, nx, ny, nz, radial_order,
                                    zeta, lmax, out.ctypes.data, 0, None))
    return out


def shore_predict(coefficients, radial_order, zeta, bvals, bvecs, *, big_delta=None, small_delta=None,
                  n_threads=None) -> np.ndarray:
    """Drop-in for ``_cs_dmri.shore_predict`` (python/src/models.rs:255-278)."""
    c = _coef_volume(coefficients, radial_order)
    nx, ny, nz, _ = c.shape
    g, _keep = L.gradients(bvals, bvecs, big_delta, small_delta, 50.0)
    out = np.zeros((nx, ny, nz, g.n_grads), np.float32)
    L.check(L.lib().shorefit_predict(c.ctypes.data, nx, ny, nz, radial_order, zeta, C.byref(g), out.ctypes.data, 0,
                                     None))
    return out


def shore_microstructure(coefficients, radial_order, zeta, mask, directions=None, *, units="um",
                         outlier_factor=10.0, n_threads=None) -> dict:
    """Drop-in for ``_cs_dmri.shore_microstructure`` (python/src/models.rs:286-331): RTOP, RTAP, RTPP,
    MSD, QIV and NG maps (zero outside ``mask``; NaN for rejected fits and, for RTAP / RTPP, missing
    directions)."""
    c = _coef_volume(coefficients, radial_order)
    nx, ny, nz, _ = c.shape
    m = np.ascontiguousarray(np.asarray(mask), dtype=np.uint8)
    if m.shape != (nx, ny, nz):
        raise ValueError("mask does not match the coefficient grid")
    if units not in ("um", "mm"):
        raise ValueError(f"units must be 'um' or 'mm', got {units!r}")
    d = None
    if directions is not None:
        d = np.ascontiguousarray(np.asarray(directions), dtype=np.float32)
        if d.shape != (nx, ny, nz, 3):
            raise ValueError(f"directions must have shape {(nx, ny, nz, 3)}, got {d.shape}")
    names = ("rtop", "rtap", "rtpp", "msd", "qiv", "ng")
    maps = {k: np.zeros((nx, ny, nz), np.float32) for k in names}
    L.check(L.lib().shorefit_microstructure(c.ctypes.data, m.ctypes.data, nx, ny, nz, radial_order, zeta,
                                            None if d is None else d.ctypes.data,
                                            {"um": L.UNITS_UM, "mm": L.UNITS_MM}[units],
                                            -1.0 if outlier_factor is None else float(outlier_factor),
                                            *[maps[k].ctypes.data for k in names], 0, None))
    return maps
