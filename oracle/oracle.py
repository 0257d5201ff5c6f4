"""ctypes front end of ``liboracle_shore.so`` (built from shore_oracle.cpp).

Test infrastructure only: see the package docstring. Each method names the
reference statement the C++ restates (paths into PennLINC/cs_dmri 0.3.1).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# ORACLE_LIB selects another build of the same source (the sanitizer build)
_LIB = Path(os.environ["ORACLE_LIB"]) if os.environ.get("ORACLE_LIB") else _HERE / "liboracle_shore.so"

_d = C.c_double
_i = C.c_int
_u = C.c_uint
_pd = C.POINTER(C.c_double)
_pf = C.POINTER(C.c_float)
_pu8 = C.POINTER(C.c_uint8)
_pu32 = C.POINTER(C.c_uint32)
_pi32 = C.POINTER(C.c_int32)

STRATEGIES = {"fixed": 0, "alpha-ratio": 1, "path-bic": 2, "l2-anchored": 3}


def build() -> None:
    """Compile the oracle with its Makefile (g++)."""
    subprocess.run(["make", "-s", "-C", str(_HERE)], check=True)


def _ptr(a: np.ndarray | None, t):
    if a is None:
        return C.cast(None, t)
    return a.ctypes.data_as(t)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    def __init__(self, path: Path = _LIB):
        if not path.exists():
            build()
        lib = C.CDLL(str(path))
        sig = {
            "orc_n_coeffs": (_i, [_u]),
            "orc_shore_indices": (None, [_u, _pi32]),
            "orc_factorial": (_d, [_u]),
            "orc_gamma": (_d, [_d]),
            "orc_assoc_legendre": (_d, [_u, _u, _d]),
            "orc_gen_laguerre": (_d, [_u, _d, _d]),
            "orc_kappa": (_d, [_d, _u, _u]),
            "orc_brainsuite_sh_block": (None, [_u, _d, _d, _pd]),
            "orc_estimate_deltas_tortoise": (None, [_pd, _i, _d, _pd, _pd]),
            "orc_qvals": (None, [_pd, _i, _d, _d, _d, _pd]),
            "orc_dwi_shells": (_i, [_pd, _i, _d, _d, _pd, _i]),
            "orc_design_matrix": (None, [_u, _d, _pd, _pd, _i, _d, _d, _d, _pd]),
            "orc_regularization": (None, [_u, _pd, _pd]),
            "orc_lipschitz": (_d, [_pd, _i, _i]),
            "orc_sym_max_eig": (_d, [_pd, _i]),
            "orc_fista": (None, [_pd, _i, _i, _pd, _d, C.c_uint32, _d, _i, _d, _pd, _pd, _pd]),
            "orc_tikhonov": (_i, [_pd, _i, _i, _pd, _pd, _d, _d, _pd, _i, _pd, _pd]),
            "orc_alpha_max": (_d, [_pd, _i, _i, _pd]),
            "orc_log_path": (None, [_d, _i, _d, _pd]),
            "orc_bic": (_d, [_i, _d, _i]),
            "orc_fit_volume": (_i, [_pf, _pu8, _i, _i, _i, _i, _pd, _i, _i, _i, _d, _d, _i, _d, _d,
                                    C.c_uint32, _d, _i, _pd, _pd, _d, _d, _i, _pf, _pf, _pf, _pu32, _pu8,
                                    _pf, _pf, _pf, _pd]),
            "orc_effective_table": (None, [_pd, _pd, _pd, _pd, _pd, _i, _d, _pd, _pd]),
            "orc_lasso_adjoint": (_i, [_pd, _i, _i, _pd, _pd, _i, _pd, _pd]),
            "orc_lasso_polish": (_i, [_pd, _i, _i, _pd, _d, _pd, _pd]),
            "orc_fit_graddev": (_i, [_pf, _pu8, _i, _i, _i, _i, _pd, _pd, _pd, _pd, _d, _d, _d, _u, _d, _pf,
                                     _i, _d, _d, C.c_uint32, _d, _i, _i, _pf, _pf, _pf, _pu32, _pu8, _pf, _pf,
                                     _pd]),
            "orc_hyp2f1": (_d, [_d, _d, _d, _d]),
            "orc_binomial_float": (_d, [_d, _u]),
            "orc_odf_radial_factor": (_d, [_u, _u, _d]),
            "orc_brainsuite_to_tournier": (None, [_u, _i, C.POINTER(C.c_int), _pd]),
            "orc_tournier_matrix": (None, [_u, _d, _u, _pd]),
            "orc_alpha_factor": (_d, [_u, _u, _d]),
            "orc_microstructure_raw": (None, [_u, _d, _pd, _pd, _pd]),
            "orc_compute_microstructure": (None, [_u, _d, _pf, _pf, C.c_int64, _i, C.c_float, _pf]),
            "orc_synthesize": (None, [_pd, _i, _i, _pf, C.c_int64, _pf]),
            "orc_project": (None, [_pd, _i, _i, _pf, C.c_int64, _pf]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        self.lib = lib

    # -- math (src/math.rs) ---------------------------------------------
    def factorial(self, n: int) -> float:
        return self.lib.orc_factorial(n)

    def gamma(self, x: float) -> float:
        return self.lib.orc_gamma(x)

    def assoc_legendre(self, l: int, m: int, x: float) -> float:
        return self.lib.orc_assoc_legendre(l, m, x)

    def gen_laguerre(self, n: int, a: float, x: float) -> float:
        return self.lib.orc_gen_laguerre(n, a, x)

    def kappa(self, zeta: float, n: int, ell: int) -> float:
        return self.lib.orc_kappa(zeta, n, ell)

    def brainsuite_sh_block(self, ell: int, theta: float, phi: float) -> np.ndarray:
        out = np.zeros(2 * ell + 1)
        self.lib.orc_brainsuite_sh_block(ell, theta, phi, _ptr(out, _pd))
        return out

    # -- gradient table / basis (src/qspace.rs, src/basis/shore.rs) ----
    def n_coeffs(self, radial_order: int) -> int:
        return self.lib.orc_n_coeffs(radial_order)

    def shore_indices(self, radial_order: int) -> np.ndarray:
        k = self.n_coeffs(radial_order)
        out = np.zeros((k, 3), dtype=np.int32)
        self.lib.orc_shore_indices(radial_order, _ptr(out, _pi32))
        return out

    def estimate_deltas_tortoise(self, bvals, gmax: float = 80e-3):
        b = _f64(bvals)
        big, small = C.c_double(), C.c_double()
        self.lib.orc_estimate_deltas_tortoise(_ptr(b, _pd), len(b), gmax, C.byref(big), C.byref(small))
        return big.value, small.value

    def qvals(self, bvals, big_delta, small_delta, b0_threshold=50.0) -> np.ndarray:
        b = _f64(bvals)
        q = np.zeros_like(b)
        self.lib.orc_qvals(_ptr(b, _pd), len(b), big_delta, small_delta, b0_threshold, _ptr(q, _pd))
        return q

    def dwi_shells(self, bvals, b0_threshold=50.0, tol=50.0) -> list[float]:
        b = _f64(bvals)
        out = np.zeros(max(1, len(b)))
        n = self.lib.orc_dwi_shells(_ptr(b, _pd), len(b), b0_threshold, tol, _ptr(out, _pd), len(out))
        return [float(v) for v in out[:n]]

    def design_matrix(self, radial_order, zeta, bvals, bvecs, big_delta, small_delta,
                      b0_threshold=50.0) -> np.ndarray:
        b = _f64(bvals)
        g = _f64(bvecs).reshape(-1, 3)
        A = np.zeros((len(b), self.n_coeffs(radial_order)))
        self.lib.orc_design_matrix(radial_order, zeta, _ptr(b, _pd), _ptr(g, _pd), len(b), big_delta,
                                   small_delta, b0_threshold, _ptr(A, _pd))
        return A

    def regularization(self, radial_order: int):
        k = self.n_coeffs(radial_order)
        dn, dl = np.zeros(k), np.zeros(k)
        self.lib.orc_regularization(radial_order, _ptr(dn, _pd), _ptr(dl, _pd))
        return dn, dl

    # -- solvers (src/solver/) -----------------------------------------
    def lipschitz(self, A) -> float:
        A = _f64(A)
        return self.lib.orc_lipschitz(_ptr(A, _pd), A.shape[0], A.shape[1])

    def sym_max_eig(self, S) -> float:
        S = _f64(S)
        return self.lib.orc_sym_max_eig(_ptr(S, _pd), S.shape[0])

    def fista(self, A, s, alpha, max_iter=1000, tol=1e-6, non_negative=False, L=None, warm=None):
        """One FistaSolver fit; returns (coef, dict(iterations, residual_l2, converged, alpha))."""
        A, s = _f64(A), _f64(s)
        N, K = A.shape
        coef = np.zeros(K)
        d = np.zeros(4)
        w = None if warm is None else _f64(warm)
        self.lib.orc_fista(_ptr(A, _pd), N, K, _ptr(s, _pd), alpha, max_iter, tol, int(non_negative),
                           -1.0 if L is None else float(L), _ptr(w, _pd), _ptr(coef, _pd), _ptr(d, _pd))
        return coef, {"iterations": int(d[0]), "residual_l2": d[1], "converged": bool(d[2]), "alpha": d[3]}

    def tikhonov(self, A, S, dn=None, dl=None, lambda_n=0.0, lambda_l=0.0):
        """TikhonovSolver fits of rows of S; returns (C, residual_l2) or raises if not PD."""
        A = _f64(A)
        S = _f64(S)
        single = S.ndim == 1
        S = np.atleast_2d(S)
        N, K = A.shape
        Cm = np.zeros((S.shape[0], K))
        r = np.zeros(S.shape[0])
        dn = None if dn is None else _f64(dn)
        dl = None if dl is None else _f64(dl)
        rc = self.lib.orc_tikhonov(_ptr(A, _pd), N, K, _ptr(dn, _pd), _ptr(dl, _pd), lambda_n, lambda_l,
                                   _ptr(S, _pd), S.shape[0], _ptr(Cm, _pd), _ptr(r, _pd))
        if rc != 0:
            raise ValueError("normal-equations matrix is not positive definite")
        return (Cm[0], r[0]) if single else (Cm, r)

    def alpha_max(self, A, s) -> float:
        A, s = _f64(A), _f64(s)
        return self.lib.orc_alpha_max(_ptr(A, _pd), A.shape[0], A.shape[1], _ptr(s, _pd))

    def log_path(self, amax, n, eps) -> np.ndarray:
        out = np.zeros(n)
        self.lib.orc_log_path(amax, n, eps, _ptr(out, _pd))
        return out

    def bic(self, n_grads, rss, k) -> float:
        return self.lib.orc_bic(n_grads, rss, k)

    # -- volume driver (src/fit.rs) ------------------------------------
    def fit_volume(self, data, mask, A, *, regularization="l1", alpha_mode="fixed", alpha=1.0,
                   alpha_ratio=1e-3, path_n_alphas=20, path_eps=None, slack=0.05, max_iter=1000,
                   tol=1e-6, non_negative=False, dn=None, dl=None, lambda_n=1e-8, lambda_l=1e-8,
                   n_threads=0) -> dict:
        data = np.ascontiguousarray(data, dtype=np.float32)
        nx, ny, nz, N = data.shape
        A = _f64(A)
        K = A.shape[1]
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        reg = {"l1": 1, "l2": 2}[regularization]
        strat = STRATEGIES[alpha_mode]
        if path_eps is None:
            path_eps = 1e-4 if alpha_mode == "l2-anchored" else 1e-3
        shp = (nx, ny, nz)
        out = {
            "coefficients": np.zeros(shp + (K,), np.float32),
            "r2": np.zeros(shp, np.float32),
            "residual_l2": np.zeros(shp, np.float32),
            "iterations": np.zeros(shp, np.uint32),
            "regularization_kind": np.zeros(shp, np.uint8),
            "alpha": np.zeros(shp, np.float32),
            "bic": np.zeros(shp, np.float32),
            "rss_l2": np.zeros(shp, np.float32),
        }
        dist = np.zeros(3)
        dn = None if dn is None else _f64(dn)
        dl = None if dl is None else _f64(dl)
        rc = self.lib.orc_fit_volume(
            _ptr(data, _pf), _ptr(m, _pu8), nx, ny, nz, N, _ptr(A, _pd), K, reg, strat, alpha, alpha_ratio,
            path_n_alphas, path_eps, slack, max_iter, tol, int(non_negative), _ptr(dn, _pd), _ptr(dl, _pd),
            lambda_n, lambda_l, int(n_threads), _ptr(out["coefficients"], _pf), _ptr(out["r2"], _pf),
            _ptr(out["residual_l2"], _pf), _ptr(out["iterations"], _pu32), _ptr(out["regularization_kind"], _pu8),
            _ptr(out["alpha"], _pf), _ptr(out["bic"], _pf), _ptr(out["rss_l2"], _pf), _ptr(dist, _pd))
        if rc != 0:
            raise ValueError("normal-equations matrix is not positive definite")
        out["alpha_distribution"] = None if (reg == 2 or strat == 0) else tuple(float(v) for v in dist)
        return out

    # -- graddev (src/graddev.rs) --------------------------------------
    def effective_table(self, T, R, bvals, bvecs_vox, bvecs_nominal, b0_threshold=50.0):
        T, R = _f64(T), _f64(R)
        b = _f64(bvals)
        gv = _f64(bvecs_vox).reshape(-1, 3)
        gn = _f64(bvecs_nominal).reshape(-1, 3)
        bo = np.zeros_like(b)
        do = np.zeros_like(gn)
        self.lib.orc_effective_table(_ptr(T, _pd), _ptr(R, _pd), _ptr(b, _pd), _ptr(gv, _pd), _ptr(gn, _pd),
                                     len(b), b0_threshold, _ptr(bo, _pd), _ptr(do, _pd))
        return bo, do

    def fit_graddev(self, data, mask, field, bvals, bvecs_vox, *, bvecs_nominal=None, frame_rotation=None,
                    big_delta, small_delta, b0_threshold=50.0, radial_order=6, zeta=700.0, alpha_mode="fixed",
                    alpha=1e-3, alpha_ratio=1e-3, max_iter=500, tol=1e-6, non_negative=False,
                    n_threads=0) -> dict:
        """Per-voxel-design L1 fit (src/graddev.rs + src/dti.rs:393-407
        pattern): field (X,Y,Z,3,3) or (X,Y,Z,9) row-major T."""
        data = np.ascontiguousarray(data, dtype=np.float32)
        nx, ny, nz, N = data.shape
        fld = np.ascontiguousarray(field, dtype=np.float32).reshape(nx, ny, nz, 9)
        m = None if mask is None else np.ascontiguousarray(mask, dtype=np.uint8)
        R = _f64(np.eye(3) if frame_rotation is None else frame_rotation).reshape(9)
        gv = _f64(bvecs_vox).reshape(-1, 3)
        gn = _f64(gv @ R.reshape(3, 3).T if bvecs_nominal is None else bvecs_nominal).reshape(-1, 3)
        b = _f64(bvals)
        K = self.n_coeffs(radial_order)
        shp = (nx, ny, nz)
        out = {
            "coefficients": np.zeros(shp + (K,), np.float32),
            "r2": np.zeros(shp, np.float32),
            "residual_l2": np.zeros(shp, np.float32),
            "iterations": np.zeros(shp, np.uint32),
            "regularization_kind": np.zeros(shp, np.uint8),
            "alpha": np.zeros(shp, np.float32),
            "bic": np.zeros(shp, np.float32),
            "lipschitz": np.zeros(shp, np.float64),
        }
        rc = self.lib.orc_fit_graddev(
            _ptr(data, _pf), _ptr(m, _pu8), nx, ny, nz, N, _ptr(b, _pd), _ptr(gv, _pd), _ptr(gn, _pd),
            _ptr(R, _pd), big_delta, small_delta, b0_threshold, radial_order, zeta, _ptr(fld, _pf),
            STRATEGIES[alpha_mode], alpha, alpha_ratio, max_iter, tol, int(non_negative), int(n_threads),
            _ptr(out["coefficients"], _pf), _ptr(out["r2"], _pf), _ptr(out["residual_l2"], _pf),
            _ptr(out["iterations"], _pu32), _ptr(out["regularization_kind"], _pu8), _ptr(out["alpha"], _pf),
            _ptr(out["bic"], _pf), _ptr(out["lipschitz"], _pd))
        if rc != 0:
            raise ValueError("graddev oracle supports alpha_mode fixed and alpha-ratio only")
        return out

    # -- differentiation (no reference counterpart) ----------------------
    def lasso_adjoint(self, A, coef, grad):
        """(grad_signal (V, N), grad_alpha (V,)) of the active-set implicit
        derivative, FP64."""
        A = _f64(A)
        N, K = A.shape
        c = _f64(np.atleast_2d(coef))
        g = _f64(np.atleast_2d(grad))
        V = c.shape[0]
        gs = np.zeros((V, N))
        ga = np.zeros(V)
        self.lib.orc_lasso_adjoint(_ptr(A, _pd), N, K, _ptr(c, _pd), _ptr(g, _pd), V, _ptr(gs, _pd), _ptr(ga, _pd))
        return gs, ga

    def lasso_polish(self, A, s, alpha, coef):
        """Exact minimiser on coef's active set and signs; (c, status), status
        0 when KKT-verified."""
        A, s, c = _f64(A), _f64(s), _f64(coef)
        N, K = A.shape
        out = np.zeros(K)
        st = self.lib.orc_lasso_polish(_ptr(A, _pd), N, K, _ptr(s, _pd), float(alpha), _ptr(c, _pd), _ptr(out, _pd))
        return out, st

    # -- post-fit (src/odf.rs, src/scalars/microstructure.rs, src/synth.rs) --
    def hyp2f1(self, a, b, c, z) -> float:
        return self.lib.orc_hyp2f1(a, b, c, z)

    def binomial_float(self, n, k) -> float:
        return self.lib.orc_binomial_float(n, k)

    def odf_radial_factor(self, n, ell, zeta) -> float:
        return self.lib.orc_odf_radial_factor(n, ell, zeta)

    def brainsuite_to_tournier(self, ell, k):
        row, sign = C.c_int(0), C.c_double(0)
        self.lib.orc_brainsuite_to_tournier(ell, k, C.byref(row), C.byref(sign))
        return row.value, sign.value

    def tournier_matrix(self, radial_order, zeta, lmax) -> np.ndarray:
        K = self.n_coeffs(radial_order)
        M = np.zeros(((lmax + 1) * (lmax + 2) // 2, K))
        self.lib.orc_tournier_matrix(radial_order, zeta, lmax, _ptr(M, _pd))
        return M

    def alpha_factor(self, n, ell, zeta) -> float:
        return self.lib.orc_alpha_factor(n, ell, zeta)

    def microstructure_raw(self, radial_order, zeta, coefs, direction=None) -> np.ndarray:
        """rtop, rtap, rtpp, msd, qiv, ng (FP64, mm units) of one coefficient vector."""
        c = _f64(coefs)
        d = None if direction is None else _f64(direction)
        out = np.zeros(6)
        self.lib.orc_microstructure_raw(radial_order, zeta, _ptr(c, _pd), _ptr(d, _pd), _ptr(out, _pd))
        return out

    def compute_microstructure(self, radial_order, zeta, coefs, dirs=None, units="um", outlier_factor=10.0):
        """Per masked voxel (rows of coefs), dict of the six f32 scalars."""
        c = np.ascontiguousarray(coefs, dtype=np.float32)
        V = c.shape[0]
        d = None if dirs is None else np.ascontiguousarray(dirs, dtype=np.float32)
        out = np.zeros((6, V), np.float32)
        self.lib.orc_compute_microstructure(radial_order, zeta, _ptr(c, _pf), _ptr(d, _pf), V,
                                            {"um": 0, "mm": 1}[units],
                                            -1.0 if outlier_factor is None else float(outlier_factor),
                                            _ptr(out, _pf))
        return dict(zip(("rtop", "rtap", "rtpp", "msd", "qiv", "ng"), out))

    def synthesize(self, A, coefs) -> np.ndarray:
        A = _f64(A)
        c = np.ascontiguousarray(coefs, dtype=np.float32)
        out = np.zeros((c.shape[0], A.shape[0]), np.float32)
        self.lib.orc_synthesize(_ptr(A, _pd), A.shape[0], A.shape[1], _ptr(c, _pf), c.shape[0], _ptr(out, _pf))
        return out

    def project(self, M, coefs) -> np.ndarray:
        M = _f64(M)
        c = np.ascontiguousarray(coefs, dtype=np.float32)
        out = np.zeros((c.shape[0], M.shape[0]), np.float32)
        self.lib.orc_project(_ptr(M, _pd), M.shape[0], M.shape[1], _ptr(c, _pf), c.shape[0], _ptr(out, _pf))
        return out


_SINGLETON: Oracle | None = None


def load() -> Oracle:
    global _SINGLETON
    if _SINGLETON is None:
        _SINGLETON = Oracle()
    return _SINGLETON


if os.environ.get("ORACLE_AUTOBUILD") == "1":  # pragma: no cover
    build()
