"""Config-3 precision report on the GPU: device fits at each arithmetic
against the FP64 oracle fit of the same block (coefficients, fitted signal,
RTOP / MSD / QIV, principal ODF peak on the 362-point sphere).

Settings: FP32, and FP16 / BF16 at refresh 5, 10 and 20. Bounds (medians
over voxels): coefficients and fitted signal <= 1e-5 for every setting, the
RTOP / MSD / QIV scalars <= 1e-4; p99 of the coefficient error <= 1e-3; the
principal peak does not move in at least 95% of voxels (a peak can hop
between two equal lobes)."""

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
    settings = (("fp32", 10),) + tuple((p, k) for p in ("fp16", "bf16") for k in (5, 10, 20))
    rep = report(data, mask, g, ref, settings=settings)
    assert len(rep["entries"]) == len(settings)
    for e in rep["entries"]:
        tag = f"{e['precision']} k={e.get('refresh_every')}"
        print(tag, e["coefficients"]["median"], e["coefficients"]["p99"], e["rtop"]["median"],
              e["principal_peak_angle_deg"]["median"])
        assert e["coefficients"]["median"] <= 1e-5, tag
        assert e["coefficients"]["p99"] <= 1e-3, tag
        assert e["fitted_signal"]["median"] <= 1e-5, tag
        for s in ("rtop", "msd", "qiv"):
            assert e[s]["median"] <= 1e-4, (tag, s)
        assert e["principal_peak_angle_deg"]["median"] == 0.0 and e["same_peak_count_frac"] >= 0.95, tag
