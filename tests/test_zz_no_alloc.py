"""shorefit_lasso_forward_noalloc against shorefit_lasso_forward.

The allocation-free entry point runs the same kernels with its per-call
buffers in the library's device scratch array, so its coefficients and
iteration counts must equal the allocating entry point's bit for bit. Two
calls in a row share that array; the second is smaller (and warm-started),
so stale contents of the first would show. A call larger than the scratch
array is refused before any device work (its buffers are sized for the call,
so nothing could be read out of bounds either way).
"""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import torch  # noqa: E402

from cs_dmri_amd import _lib as L  # noqa: E402
from cs_dmri_amd import design_matrix  # noqa: E402
from cs_dmri_amd import workloads as W  # noqa: E402
from cs_dmri_amd.autograd import LassoDesign  # noqa: E402


def _noalloc(d, S, alpha, warm, max_iter, tol):
    V = S.shape[0]
    coef = torch.full((V, d.K), float("nan"), dtype=torch.float32, device=S.device)
    its = torch.full((V,), -1, dtype=torch.int32, device=S.device)
    a_ptr, a_scalar = (alpha.data_ptr(), 0.0) if isinstance(alpha, torch.Tensor) else (None, float(alpha))
    rc = L.lib().shorefit_lasso_forward_noalloc(
        d._h, S.data_ptr(), V, a_ptr, a_scalar, None if warm is None else warm.data_ptr(), max_iter, tol, 0,
        coef.data_ptr(), its.data_ptr(), torch.cuda.current_stream().cuda_stream)
    return rc, coef, its


def test_noalloc_matches_lasso_forward():
    gt, sig = W.config2_signals(3000, seed=13)
    A = design_matrix(gt.bvals, gt.bvecs, radial_order=8, big_delta=gt.big_delta, small_delta=gt.small_delta)
    d = LassoDesign(A)
    dev = torch.device("cuda:0")
    S = torch.from_numpy(sig).to(dev)
    alpha = torch.from_numpy(np.random.default_rng(3).uniform(5e-4, 2e-3, size=len(sig)).astype(np.float32)).to(dev)
    # first call: 3000 signals, per-signal alpha, no warm start
    ref1, rits1 = d.forward(S, alpha, max_iter=400, tol=1e-6)
    rc, c1, i1 = _noalloc(d, S, alpha, None, 400, 1e-6)
    assert rc == L.OK, L.lib().shorefit_last_error()
    torch.cuda.synchronize()
    assert torch.equal(c1, ref1) and torch.equal(i1, rits1)
    assert int(i1.min()) >= 1 and int(i1.max()) > 2 * int(i1.min())  # uneven convergence: slots refill
    # second call, smaller: 700 other signals, scalar alpha, warm start
    S2 = S[1500:2200].contiguous()
    warm = ref1[:700].contiguous()
    ref2, rits2 = d.forward(S2, 1e-3, warm=warm, max_iter=250, tol=1e-6)
    rc, c2, i2 = _noalloc(d, S2, 1e-3, warm, 250, 1e-6)
    assert rc == L.OK, L.lib().shorefit_last_error()
    torch.cuda.synchronize()
    assert torch.equal(c2, ref2) and torch.equal(i2, rits2)
    # an empty call is a no-op
    rc, _, _ = _noalloc(d, S[:0], 1e-3, None, 100, 1e-6)
    assert rc == L.OK


def test_noalloc_refuses_calls_beyond_its_scratch():
    gt, sig = W.config2_signals(64, seed=14)
    A = design_matrix(gt.bvals, gt.bvecs, radial_order=8, big_delta=gt.big_delta, small_delta=gt.small_delta)
    d = LassoDesign(A)
    V = 120_000  # 4 (160 + 1) bytes each: 77 MB of scratch, above the 64 MiB the library holds
    S = torch.zeros((V, d.N), dtype=torch.float32, device="cuda:0")
    rc, _, _ = _noalloc(d, S, 1e-3, None, 100, 1e-6)
    assert rc == L.E_INVALID
    msg = L.lib().shorefit_last_error().decode()
    assert "scratch" in msg and "at most" in msg, msg
