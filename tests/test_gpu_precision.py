"""Config-3 precision report on the GPU: device fits at each arithmetic
against the FP64 oracle fit of the same block (coefficients, fitted signal,
RTOP / MSD / QIV, principal ODF peak on the 362-point sphere).

Bounds (medians over voxels): FP32 coefficients <= 1e-5, FP16 (k = 10)
<= 1e-4, BF16 (k = 10) <= 1e-3; the scalars and the fitted signal follow the
coefficients; the principal peak does not move for FP32 / FP16 in at least
95% of voxels (a peak can hop between two equal lobes)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from cs_dmri_amd import workloads as W  # noqa: E402
from cs_dmri_amd.precision import report  # noqa: E402


def test_precision_report_block(orc):
    wl = W.config1()
    g = wl.gtab
    data = np.ascontiguousarray(wl.data[:8, :8, :4])
    mask = np.ones(data.shape[:3], bool)
    A = orc.design_matrix(6, 700.0, g.bvals, g.bvecs, g.big_delta, g.small_delta)
    ref = orc.fit_volume(data, mask, A, alpha_mode="fixed", alpha=1e-3, max_iter=500, tol=1e-6)["coefficients"]
    rep = report(data, mask, g, ref, settings=(("fp32", 10), ("fp16", 10), ("bf16", 10)))
    by = {e["precision"]: e for e in rep["entries"]}
    print({k: (v["coefficients"]["median"], v["rtop"]["median"], v["principal_peak_angle_deg"]["median"])
           for k, v in by.items()})
    for prec, bound in (("fp32", 1e-5), ("fp16", 1e-4), ("bf16", 1e-3)):
        e = by[prec]
        assert e["coefficients"]["median"] <= bound
        assert e["fitted_signal"]["median"] <= bound
        for s in ("rtop", "msd", "qiv"):
            assert e[s]["median"] <= 10 * bound
    for prec in ("fp32", "fp16"):
        ang = by[prec]["principal_peak_angle_deg"]
        assert ang["median"] == 0.0 and by[prec]["same_peak_count_frac"] >= 0.95
