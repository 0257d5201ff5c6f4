"""Loaders and seeded problem builders shared by the tests."""

from __future__ import annotations

import json
import struct
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden"
LASSO_DIR = GOLDEN / "lasso_fixtures"


def lasso_fixtures():
    """The reference's sklearn LASSO fixtures (tests/optimization_sklearn.rs:32-127)."""
    names = json.loads((LASSO_DIR / "manifest.json").read_text())
    out = []
    for name in names:
        raw = (LASSO_DIR / name).read_bytes()
        magic, version, m, n, pos, alpha, amax, obj = struct.unpack("<IIIIB7xddd", raw[:48])
        assert magic == 0x434C4153 and version == 1
        body = np.frombuffer(raw[48:], dtype="<f8")
        assert body.size == m * n + m + n
        A = body[: m * n].reshape(m, n).copy()
        y = body[m * n: m * n + m].copy()
        beta = body[m * n + m:].copy()
        out.append(dict(name=name[:-4], m=m, n=n, positive=bool(pos), alpha=alpha, alpha_max=amax,
                        objective=obj, A=A, y=y, beta=beta))
    return out


def lasso_objective(A, y, b, alpha):
    """tests/common/mod.rs:111-115"""
    r = A @ b - y
    return 0.5 / A.shape[0] * float(r @ r) + alpha * float(np.abs(b).sum())


def kkt_lasso(A, y, b, alpha, zero_tol=1e-10):
    """tests/common/mod.rs:127-153 — (max active residual, max inactive excess)."""
    g = A.T @ (A @ b - y) / A.shape[0]
    act = np.abs(b) > zero_tol
    a = np.abs(g[act] + alpha * np.sign(b[act])).max(initial=0.0)
    ina = (np.abs(g[~act]) - alpha).max(initial=0.0)
    return float(a), float(max(ina, 0.0))


def kkt_nonneg_lasso(A, y, b, alpha, zero_tol=1e-10):
    """tests/common/mod.rs:161-189."""
    g = A.T @ (A @ b - y) / A.shape[0]
    act = b > zero_tol
    a = np.abs(g[act] + alpha).max(initial=0.0)
    ina = (-g[~act] - alpha).max(initial=0.0)
    return float(a), float(max(ina, 0.0))


def make_design(m, n, seed):
    """Gaussian design with iid N(0, 1/m) entries (tests/common/mod.rs:31-35; numpy RNG)."""
    return np.random.default_rng(seed).standard_normal((m, n)) / np.sqrt(m)


def make_ill_conditioned(m, n, kappa, seed):
    """SVD-rescaled design with singular values geometric in [1/kappa, 1] (tests/common/mod.rs:39-58)."""
    U, _, Vt = np.linalg.svd(make_design(m, n, seed), full_matrices=False)
    r = min(m, n)
    s = np.exp(np.arange(r) * (-np.log(kappa) / max(r - 1, 1)))
    return (U[:, :r] * s) @ Vt[:r]


def make_planted_beta(n, k, signed, seed):
    """tests/common/mod.rs:85-100."""
    rng = np.random.default_rng(seed)
    idx = rng.permutation(n)[:k]
    beta = np.zeros(n)
    mag = 0.5 + 1.5 * rng.random(k)
    sign = np.where(rng.random(k) < 0.5, -1.0, 1.0) if signed else np.ones(k)
    beta[idx] = sign * mag
    return beta


def make_noisy_signal(A, beta, sigma, seed):
    return A @ beta + sigma * np.random.default_rng(seed).standard_normal(A.shape[0])


def shell_gtab(per_shell=30, shells=(1000.0, 2000.0, 3000.0)):
    """1 b0 + golden-spiral shells (tests/optimization_kkt.rs:221-235)."""
    bvals = [0.0]
    bvecs = [[0.0, 0.0, 0.0]]
    phi = np.pi * (1.0 + np.sqrt(5.0))
    for b in shells:
        for i in range(per_shell):
            z = 1.0 - (2.0 * i + 1.0) / per_shell
            r = np.sqrt(max(1.0 - z * z, 0.0))
            th = phi * i
            bvecs.append([r * np.cos(th), r * np.sin(th), z])
            bvals.append(b)
    return np.array(bvals), np.array(bvecs)
