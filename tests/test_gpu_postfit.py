"""GPU parity of the post-fit maps against the FP64 oracle.

* Tournier SH projection (src/odf.rs:79-102, python/src/models.rs:203-248):
  the library's matrix to 1e-13 relative, the projected volume to 1e-6
  relative (f32 outputs of FP64 sums), zeros outside the mask.
* Synthesis (src/synth.rs:26-84): f32 outputs of FP64 sums to 1e-6 of the
  volume's largest value; all-zero voxels give zeros.
* Microstructure (src/scalars/microstructure.rs, src/io/microstructure.rs):
  the six maps to 1e-5 relative where finite, identical NaN pattern, with
  and without fit-failure rejection, in both unit systems.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

import cs_dmri_amd as sf  # noqa: E402
from cs_dmri_amd import workloads as W  # noqa: E402

NAMES = ("rtop", "rtap", "rtpp", "msd", "qiv", "ng")


@pytest.fixture(scope="module")
def fitted():
    wl = W.config1()
    g = wl.gtab
    data = np.ascontiguousarray(wl.data[:8, :8, :6])
    mask = np.ones(data.shape[:3], bool)
    mask[0, :2, 0] = False
    out = sf.shore_fit(data, g.bvals, g.bvecs, mask, big_delta=g.big_delta, small_delta=g.small_delta,
                       radial_order=6, alpha_mode="fixed", alpha=1e-3, max_iter=300)
    return g, out["coefficients"], mask


def test_tournier_matrix_and_projection(orc, fitted):
    g, C, mask = fitted
    for order, lmax in [(6, 6), (6, 4), (8, 8)]:
        M = sf.tournier_matrix(order, 700.0, lmax)
        Mr = orc.tournier_matrix(order, 700.0, lmax)
        np.testing.assert_allclose(M, Mr, rtol=1e-13, atol=1e-16)
    sh = sf.shore_odf_sh(C, 6, 700.0, 6, mask)
    ref = orc.project(orc.tournier_matrix(6, 700.0, 6), C[mask])
    np.testing.assert_allclose(sh[mask], ref, rtol=1e-6, atol=1e-7 * np.abs(ref).max())
    assert not sh[~mask].any() and sh.shape == C.shape[:3] + (28,)
    full = sf.shore_odf_sh(C, 6, 700.0, 2)  # no mask: every voxel
    np.testing.assert_allclose(full[mask], sh[mask][:, :6], rtol=1e-6, atol=1e-7 * np.abs(ref).max())
    with pytest.raises(ValueError, match="lmax must be even"):
        sf.shore_odf_sh(C, 6, 700.0, 3)


def test_predict_matches_oracle(orc, fitted):
    g, C, mask = fitted
    tgt = W.cartesian_table(64, seed=41)
    pred = sf.shore_predict(C, 6, 700.0, tgt.bvals, tgt.bvecs, big_delta=tgt.big_delta, small_delta=tgt.small_delta)
    A = orc.design_matrix(6, 700.0, tgt.bvals, tgt.bvecs, tgt.big_delta, tgt.small_delta)
    ref = orc.synthesize(A, C.reshape(-1, C.shape[-1])).reshape(C.shape[:3] + (len(tgt.bvals),))
    np.testing.assert_allclose(pred, ref, rtol=1e-6, atol=1e-6 * np.abs(ref).max())
    assert not pred[~mask].any()
    # the fit's own table reproduces the fitted signal (R^2 of the fit)
    own = sf.shore_predict(C, 6, 700.0, g.bvals, g.bvecs, big_delta=g.big_delta, small_delta=g.small_delta)
    A0 = orc.design_matrix(6, 700.0, g.bvals, g.bvecs, g.big_delta, g.small_delta)
    np.testing.assert_allclose(own[mask], (C[mask].astype(np.float64) @ A0.T), rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("units,factor", [("um", 10.0), ("mm", None), ("mm", 3.0)])
def test_microstructure_matches_oracle(orc, fitted, units, factor):
    g, C, mask = fitted
    rng = np.random.default_rng(7)
    dirs = rng.standard_normal(C.shape[:3] + (3,)).astype(np.float32)
    dirs[1, 1, 1] = 0.0
    dirs[2, 3, 4, 0] = np.nan
    out = sf.shore_microstructure(C, 6, 700.0, mask, dirs, units=units, outlier_factor=factor)
    ref = orc.compute_microstructure(6, 700.0, C[mask], dirs[mask], units=units, outlier_factor=factor)
    for k in NAMES:
        a, b = out[k][mask], ref[k]
        np.testing.assert_array_equal(np.isnan(a), np.isnan(b), err_msg=k)
        fin = np.isfinite(b)
        np.testing.assert_allclose(a[fin], b[fin], rtol=1e-5, atol=0, err_msg=k)
        assert not out[k][~mask].any()
    assert np.isnan(out["rtap"][1, 1, 1]) and np.isnan(out["rtpp"][2, 3, 4])
    nod = sf.shore_microstructure(C, 6, 700.0, mask, None, units=units, outlier_factor=factor)
    assert np.isnan(nod["rtap"][mask]).all() and np.isnan(nod["rtpp"][mask]).all()
    np.testing.assert_array_equal(nod["rtop"], out["rtop"])


def test_fit_object_methods(fitted):
    g, C, mask = fitted
    wl = W.config1()
    model = sf.ShoreModel(sf.GradientTable(wl.gtab.bvals, wl.gtab.bvecs, big_delta=wl.gtab.big_delta,
                                           small_delta=wl.gtab.small_delta),
                          radial_order=6, alpha_mode="fixed", alpha=1e-3, max_iter=300)
    fit = model.fit(wl.data[:4, :4, :2], np.ones((4, 4, 2), bool), diagnostics=True)
    sh = fit.odf_sh()
    assert sh.shape == (4, 4, 2, 28)
    pred = fit.predict()
    assert pred.shape == wl.data[:4, :4, :2].shape and (pred >= 0).all()
    ms = fit.microstructure(units="mm", outlier_factor=None)
    assert set(ms) == set(NAMES) and np.isfinite(ms["rtop"][fit.mask]).all()


class _Hip:
    """Device buffers through the HIP runtime (the library's on_device path)."""

    def __init__(self):
        import ctypes as C
        self.C = C
        self.rt = C.CDLL("libamdhip64.so")

    def upload(self, a):
        C = self.C
        p = C.c_void_p()
        assert self.rt.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 1))) == 0
        assert self.rt.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0  # host to device
        return p

    def download(self, p, a):
        C = self.C
        assert self.rt.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, C.c_size_t(a.nbytes), 2) == 0  # device to host
        return a

    def free(self, *ps):
        for p in ps:
            self.rt.hipFree(p)


def test_device_resident_inputs(fitted):
    from cs_dmri_amd import _lib as L
    g, C, mask = fitted
    hip = _Hip()
    c = np.ascontiguousarray(C)
    m = mask.astype(np.uint8)
    out = np.zeros(C.shape[:3] + (28,), np.float32)
    dc, dm, do = hip.upload(c), hip.upload(m), hip.upload(out)
    nx, ny, nz = C.shape[:3]
    try:
        L.check(L.lib().shorefit_odf_sh(dc.value, dm.value, nx, ny, nz, 6, 700.0, 6, do.value, 1, None))
        hip.download(do, out)
    finally:
        hip.free(dc, dm, do)
    np.testing.assert_array_equal(out, sf.shore_odf_sh(C, 6, 700.0, 6, mask))
