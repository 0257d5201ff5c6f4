"""Progress reporting of whole-volume fits: the on_voxel hook behind cs-fit's
heartbeat (src/fit.rs:69-73, src/bin/cs-fit.rs:285-312, src/progress.rs:17-80).

The solve kernels publish per-wave finished-voxel counts into host memory and
the calling thread polls them. The callback sequence starts at 0, ends at the
masked-voxel count, never decreases, and reporting changes no output bit.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import cs_dmri_amd as sf  # noqa: E402
from cs_dmri_amd import io as sio  # noqa: E402
from cs_dmri_amd import workloads as W  # noqa: E402


@pytest.fixture(scope="module")
def block():
    gt, sig = W.config2_signals(32 * 32 * 64, seed=8)
    data = np.ascontiguousarray(sig.reshape(32, 32, 64, -1).astype(np.float32))
    mask = np.ones(data.shape[:3], bool)
    mask[:, :, :2] = False  # masked-out voxels are not counted
    return gt, data, mask


def _fit(gt, data, mask, **kw):
    return sf.shore_fit(data, gt.bvals, gt.bvecs, mask, big_delta=gt.big_delta, small_delta=gt.small_delta,
                        radial_order=8, diagnostics=True, **kw)


@pytest.mark.parametrize("precision", ["fp32", "fp16"])
def test_progress_counts_during_solve(block, precision):
    gt, data, mask = block
    total = int(mask.sum())
    # tol = 0 runs every voxel the full max_iter, a solve of roughly 0.1 s
    kw = dict(alpha_mode="fixed", alpha=1e-3, max_iter=1000, tol=0.0, precision=precision)
    calls = []
    out = _fit(gt, data, mask, progress=lambda d, t: calls.append((d, t)), progress_interval=0.001, **kw)
    ref = _fit(gt, data, mask, **kw)
    print(f"{precision}: {len(calls)} progress calls, first {calls[:2]}, last {calls[-2:]}", flush=True)
    assert calls[0] == (0, total) and calls[-1] == (total, total)
    done = [d for d, _ in calls]
    assert all(t == total for _, t in calls)
    assert all(b >= a for a, b in zip(done, done[1:]))
    mid = done[1:-1]
    assert len(mid) >= 3, "the solve ended before the first poll"
    assert any(0 < d < total for d in mid), "no partial count published while the kernel ran"
    for key in ("coefficients", "residual_l2", "iterations"):
        np.testing.assert_array_equal(out[key], ref[key])


@pytest.mark.parametrize("mode", ["l2-anchored", "path-bic"])
def test_progress_path_modes(block, mode):
    gt, data, mask = block
    total = int(mask.sum())
    calls = []
    out = _fit(gt, data[:, :, :16], mask[:, :, :16], alpha_mode=mode, max_iter=300,
               progress=lambda d, t: calls.append(d), progress_interval=0.001)
    n = int(mask[:, :, :16].sum())
    assert n < total
    assert calls[0] == 0 and calls[-1] == n
    assert all(b >= a and b <= n for a, b in zip(calls, calls[1:]))
    ref = _fit(gt, data[:, :, :16], mask[:, :, :16], alpha_mode=mode, max_iter=300)
    np.testing.assert_array_equal(out["coefficients"], ref["coefficients"])


def test_progress_tikhonov_and_unregistered(block):
    gt, data, mask = block
    calls = []
    _fit(gt, data[:4], mask[:4], regularization="l2", progress=lambda d, t: calls.append((d, t)))
    n = int(mask[:4].sum())
    assert calls == [(0, n), (n, n)]
    # the hook is removed after the call: a later fit reports nothing
    _fit(gt, data[:4], mask[:4], regularization="l2")
    assert calls == [(0, n), (n, n)]


def test_heartbeat_lines_from_a_fit(block, capsys):
    gt, data, mask = block
    n = int(mask[:8].sum())
    hb = sio.Heartbeat("cs-fit", n, 0.001)
    _fit(gt, data[:8], mask[:8], alpha_mode="fixed", alpha=1e-3, max_iter=1000, tol=0.0,
         progress=hb.update, progress_interval=0.001)
    hb.finish()
    lines = capsys.readouterr().err.strip().splitlines()
    assert lines[-1].startswith(f"[cs-fit] done {n}/{n} (100.0%) elapsed=")
    assert all(ln.startswith("[cs-fit] ") for ln in lines)
