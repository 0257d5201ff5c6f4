"""Seeded synthetic SHORE-fit workloads (gradient tables and phantom volumes).

These are the inputs the benchmark and the parity tests run on. They replace
external DWI series (none ship with the reference) with a documented,
deterministic generator:

* gradient table — the non-zero points of the Cartesian q-grid
  ``{-5..5}^3`` with ``|q| <= 5`` (514 points), a seeded draw of ``n_dirs`` of
  them without replacement, plus one b0 at row 0. ``b = (bmax/25)|q|^2``,
  ``bvec = q/|q|``; ``Delta = 1/(4 pi^2)``, ``delta = 0`` so ``tau = 1/(4 pi^2)``.
* signal — per voxel 1-3 sticks (lambda_par 1.7e-3, lambda_perp 0.3e-3 mm^2/s,
  uniform random directions, Dirichlet(1) weights) plus an isotropic 3.0e-3
  compartment of fraction U(0, 0.3); ``S0 = 1000``; Rician noise of the given
  sigma (``sqrt((S + sigma n1)^2 + (sigma n2)^2)``).
* masks — full (16x16x8) or an ellipsoid in a 96x96x60 grid (250,176 voxels).

All draws use ``numpy.random.default_rng(seed)``.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

S0 = 1000.0
LAMBDA_PAR = 1.7e-3
LAMBDA_PERP = 0.3e-3
D_ISO = 3.0e-3
TAU_CARTESIAN = 1.0 / (4.0 * np.pi ** 2)


@dataclass
class GradientTable:
    bvals: np.ndarray      # (N,) float64, s/mm^2
    bvecs: np.ndarray      # (N, 3) float64, unit (zero for b0)
    big_delta: float
    small_delta: float
    b0_threshold: float = 50.0


def cartesian_qgrid() -> np.ndarray:
    """Non-zero integer points of {-5..5}^3 with |q| <= 5, in lexicographic order (514 rows)."""
    r = np.arange(-5, 6)
    g = np.stack(np.meshgrid(r, r, r, indexing="ij"), axis=-1).reshape(-1, 3)
    n2 = (g ** 2).sum(1)
    keep = (n2 > 0) & (n2 <= 25)
    return g[keep].astype(np.float64)


def cartesian_table(n_dirs: int, seed: int, bmax: float = 5000.0) -> GradientTable:
    grid = cartesian_qgrid()
    rng = np.random.default_rng(seed)
    pick = rng.choice(len(grid), size=n_dirs, replace=False)
    q = grid[pick]
    qn = np.linalg.norm(q, axis=1)
    bvals = np.concatenate([[0.0], bmax / 25.0 * qn ** 2])
    bvecs = np.vstack([np.zeros((1, 3)), q / qn[:, None]])
    return GradientTable(bvals, bvecs, TAU_CARTESIAN, 0.0)


def ellipsoid_mask(shape=(96, 96, 60), centre=(47.5, 47.5, 29.5), semi=(45.7, 45.7, 28.6)) -> np.ndarray:
    x, y, z = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    r = ((x - centre[0]) / semi[0]) ** 2 + ((y - centre[1]) / semi[1]) ** 2 + ((z - centre[2]) / semi[2]) ** 2
    return r <= 1.0


def _unit_vectors(rng, n):
    v = rng.standard_normal((n, 3))
    return v / np.linalg.norm(v, axis=1, keepdims=True)


def multi_tensor_signal(gtab: GradientTable, n_vox: int, sigma, seed: int, chunk: int = 16384) -> np.ndarray:
    """(n_vox, N) float32 Rician-noised multi-stick + isotropic signals.

    ``sigma`` is a scalar or a per-voxel array.
    """
    rng = np.random.default_rng(seed)
    b = gtab.bvals
    g = gtab.bvecs
    n_sticks = rng.integers(1, 4, size=n_vox)
    dirs = _unit_vectors(rng, 3 * n_vox).reshape(n_vox, 3, 3)
    w = rng.dirichlet(np.ones(3), size=n_vox)
    # Dirichlet(1) over the sticks that exist: renormalise the first n_sticks.
    active = np.arange(3)[None, :] < n_sticks[:, None]
    w = np.where(active, w, 0.0)
    w /= w.sum(1, keepdims=True)
    f_iso = rng.uniform(0.0, 0.3, size=n_vox)
    sig = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (n_vox,))
    n1 = rng.standard_normal((n_vox, len(b)))
    n2 = rng.standard_normal((n_vox, len(b)))
    out = np.empty((n_vox, len(b)), dtype=np.float32)
    iso = np.exp(-D_ISO * b)
    for s in range(0, n_vox, chunk):
        e = slice(s, min(n_vox, s + chunk))
        cos = np.einsum("nj,vkj->vkn", g, dirs[e])           # (v, 3, N)
        att = np.exp(-b[None, None, :] * (LAMBDA_PERP + (LAMBDA_PAR - LAMBDA_PERP) * cos ** 2))
        stick = np.einsum("vk,vkn->vn", w[e], att)
        clean = S0 * ((1.0 - f_iso[e, None]) * stick + f_iso[e, None] * iso[None, :])
        sv = sig[e, None]
        out[e] = np.sqrt((clean + sv * n1[e]) ** 2 + (sv * n2[e]) ** 2)
    return out


@dataclass
class Workload:
    name: str
    gtab: GradientTable
    data: np.ndarray       # (X, Y, Z, N) float32
    mask: np.ndarray       # (X, Y, Z) bool
    radial_order: int
    zeta: float = 700.0

    @property
    def n_voxels(self) -> int:
        return int(self.mask.sum())


def _volume(shape, mask, signals):
    data = np.zeros(tuple(shape) + (signals.shape[1],), dtype=np.float32)
    data[mask] = signals
    return data


def config1(shape=(16, 16, 8)) -> Workload:
    """Order-6 CPU reference case: 101 measurements, 2048 voxels, sigma = S0/30."""
    gt = cartesian_table(100, seed=1)
    mask = np.ones(shape, dtype=bool)
    sig = multi_tensor_signal(gt, int(mask.sum()), S0 / 30.0, seed=2)
    return Workload("shore6-cartesian101-16x16x8", gt, _volume(shape, mask, sig), mask, 6)


def config2(shape=(96, 96, 60), radial_order: int = 8) -> Workload:
    """Whole-brain shape: 258 measurements, ellipsoid mask (250,176 voxels), sigma = S0/20."""
    gt = cartesian_table(257, seed=3)
    mask = ellipsoid_mask(shape) if tuple(shape) == (96, 96, 60) else np.ones(shape, dtype=bool)
    sig = multi_tensor_signal(gt, int(mask.sum()), S0 / 20.0, seed=2)
    return Workload(f"shore{radial_order}-cartesian258-{'x'.join(map(str, shape))}", gt,
                    _volume(shape, mask, sig), mask, radial_order)


def config2_signals(n_vox: int, seed: int = 2, sigma: float = S0 / 20.0):
    """Voxel-major (n_vox, 258) signals with config-2's table, without a volume."""
    gt = cartesian_table(257, seed=3)
    return gt, multi_tensor_signal(gt, n_vox, sigma, seed=seed)


def mixed_snr_signals(gt: GradientTable, n_vox: int, seed: int = 5, snr=(5.0, 40.0)):
    """Per-voxel SNR drawn uniformly (config 4's uneven-convergence signal)."""
    rng = np.random.default_rng(seed)
    snr_v = rng.uniform(snr[0], snr[1], size=n_vox)
    return multi_tensor_signal(gt, n_vox, S0 / snr_v, seed=seed + 1000)


def graddev_tensors(n_vox: int, seed: int = 4, sd: float = 0.02) -> np.ndarray:
    """(n_vox, 3, 3) deviation tensors T = I + E, E symmetric with iid N(0, sd^2) entries."""
    rng = np.random.default_rng(seed)
    e6 = rng.normal(0.0, sd, size=(n_vox, 6))
    E = np.zeros((n_vox, 3, 3))
    iu = np.triu_indices(3)
    E[:, iu[0], iu[1]] = e6
    E = E + np.transpose(E, (0, 2, 1)) - np.eye(3)[None] * E
    return np.eye(3)[None] + E


def config4(shape=(96, 96, 60), radial_order: int = 8):
    """Uneven per-voxel-design workload: config 2's table and ellipsoid mask,
    per-voxel SNR ~ U(5, 40) (seed 5) and deviation tensors T = I + E (seed
    4) on the masked voxels (identity outside). The signals follow the
    nominal table. Returns (workload, field (X, Y, Z, 3, 3) float32)."""
    gt = cartesian_table(257, seed=3)
    mask = ellipsoid_mask(shape) if tuple(shape) == (96, 96, 60) else np.ones(shape, dtype=bool)
    n = int(mask.sum())
    sig = mixed_snr_signals(gt, n, seed=5)
    field = np.broadcast_to(np.eye(3, dtype=np.float32), tuple(shape) + (3, 3)).copy()
    field[mask] = graddev_tensors(n, seed=4).astype(np.float32)
    wl = Workload(f"shore{radial_order}-graddev-cartesian258-{'x'.join(map(str, shape))}", gt,
                  _volume(shape, mask, sig), mask, radial_order)
    return wl, field


def config5(n_vox: int = 65536, frac: float = 0.1, seed: int = 6):
    """Differentiation workload: config 2's table and signal generator
    (sigma = S0/20) for n_vox voxels; frac of the diffusion-weighted
    q-points held out (seed 6). Returns (table, signals (n_vox, N) float32,
    train row indices, held-out row indices); the b0 row is always trained."""
    gt, sig = config2_signals(n_vox)
    rng = np.random.default_rng(seed)
    dw = np.flatnonzero(gt.bvals > gt.b0_threshold)
    held = np.sort(rng.choice(dw, size=int(round(frac * len(dw))), replace=False))
    train = np.setdiff1d(np.arange(len(gt.bvals)), held)
    return gt, sig, train, held
