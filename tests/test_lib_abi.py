"""CPU checks of the C-ABI library: it loads, exports every function declared
in include/shorefit.h, its ctypes struct mirrors match the header's layout,
and the host-side validation of shore_fit (src/fit.rs:555-730,
python/src/models.rs:112-147) refuses bad input before any device work."""

import ctypes as C
import re
import subprocess
from pathlib import Path

import numpy as np
import pytest

from cs_dmri_amd import _lib as L
from cs_dmri_amd import workloads as W

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "shorefit.h"


def header_functions():
    text = HEADER.read_text()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char\*)\s+(shorefit_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_function():
    lib = C.CDLL(str(L.LIB_PATH))
    names = header_functions()
    assert len(names) >= 8
    for name in names:
        assert hasattr(lib, name), name
    assert set(names) == set(L.EXPORTS)


def test_struct_layout_matches_header(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "shorefit.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\", sizeof(ShorefitSpec), sizeof(ShorefitGradients),"
        " sizeof(ShorefitVolume), sizeof(ShorefitOutput), offsetof(ShorefitOutput, warnings),"
        " offsetof(ShorefitOutput, fit_ms), offsetof(ShorefitSpec, refresh_every)); return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(L.ShorefitSpec), C.sizeof(L.ShorefitGradients), C.sizeof(L.ShorefitVolume),
            C.sizeof(L.ShorefitOutput), L.ShorefitOutput.warnings.offset, L.ShorefitOutput.fit_ms.offset,
            L.ShorefitSpec.refresh_every.offset]
    assert got == want


def test_spec_defaults_match_reference():
    s = L.ShorefitSpec()
    L.lib().shorefit_spec_default(C.byref(s))
    # src/fit.rs:647-669 and python/src/models.rs:71-75
    assert (s.radial_order, s.zeta, s.regularization, s.alpha_mode) == (6, 700.0, 1, 3)
    assert (s.alpha, s.alpha_ratio, s.path_n_alphas, s.slack) == (1.0, 1e-3, 20, 0.05)
    assert (s.max_iter, s.tol, s.non_negative, s.lambda_n, s.lambda_l) == (1000, 1e-6, 0, 1e-8, 1e-8)
    assert L.lib().shorefit_version().decode().startswith("shorefit")


def test_n_coeffs():
    from cs_dmri_amd import n_coeffs
    assert [n_coeffs(o) for o in (2, 4, 6, 8, 10)] == [8, 29, 72, 145, 256]
    with pytest.raises(ValueError):
        n_coeffs(7)


def test_graddev_variant_table():
    """Every radial order 2-10 has variants (host-side table, no device
    work): the padded coefficient count covers K in multiples of 16, the Gram
    rows per chunk are a multiple of 4 of at most 64, and an index past the
    order's variants leaves the outputs alone."""
    from cs_dmri_amd import graddev_variants, n_coeffs
    for order in (2, 4, 6, 8, 10):
        vs = graddev_variants(order)
        assert vs, order
        K = n_coeffs(order)
        for v in vs:
            assert v["kp"] >= K and v["kp"] % 16 == 0 and v["kp"] - K < 16, (order, v)
            assert v["rows"] % 4 == 0 and 4 <= v["rows"] <= 64, (order, v)
            assert v["gram_waves"] >= 1 and v["fista_waves"] >= 1
        assert len({v["kp"] for v in vs}) == 1, order
        out = C.c_int32(-7)
        n = L.lib().shorefit_graddev_variants(order, len(vs), C.byref(out), None, None, None)
        assert n == len(vs) and out.value == -7
    assert L.lib().shorefit_graddev_variants(12, 0, None, None, None, None) == 0


def _series(shells):
    bvals = [0.0, 0.0]
    bvecs = [[0, 0, 1.0]] * 2
    for b in shells:
        for i in range(40):
            z = 1 - 2 * (i + 0.5) / 40
            r = np.sqrt(1 - z * z)
            bvals.append(b)
            bvecs.append([r * np.cos(2.399963 * i), r * np.sin(2.399963 * i), z])
    bvals = np.array(bvals)
    data = np.broadcast_to(np.exp(-bvals * 0.0008).astype(np.float32), (2, 2, 1, len(bvals))).copy()
    return data, bvals, np.array(bvecs), np.ones((2, 2, 1), bool)


def test_refusals_are_value_errors():
    # These are raised by the host side before any device call.
    from cs_dmri_amd import shore_fit
    data, bvals, bvecs, mask = _series([1000.0])
    with pytest.raises(ValueError, match="at least 2 b-value shells") as e:
        shore_fit(data, bvals, bvecs, mask, regularization="l2", big_delta=0.04, small_delta=0.01)
    assert "b = 1000" in str(e.value) and "allow_single_shell" in str(e.value)
    d0, b0, v0, m0 = _series([])
    with pytest.raises(ValueError, match="needs diffusion-weighted data"):
        shore_fit(d0, b0, v0, m0, regularization="l2", allow_single_shell=True)
    data, bvals, bvecs, mask = _series([1000.0, 2000.0])
    bad = [dict(regularization="l3"), dict(alpha_mode="nope"), dict(radial_order=5), dict(radial_order=0),
           dict(alpha_mode="alpha-ratio", alpha_ratio=1.5), dict(alpha_mode="path-bic", path_n_alphas=1),
           dict(alpha_mode="path-bic", path_eps=2.0), dict(alpha_mode="l2-anchored", slack=-1.0),
           dict(bvec_frame="sideways"), dict(precision="fp8")]
    for kw in bad:
        with pytest.raises(ValueError):
            shore_fit(data, bvals, bvecs, mask, **kw)
    with pytest.raises(ValueError, match="volumes but the gradient table"):
        shore_fit(data[..., :-1], bvals, bvecs, mask)
    with pytest.raises(ValueError, match="mask shape"):
        shore_fit(data, bvals, bvecs, np.ones((2, 2, 2), bool))
    with pytest.raises(ValueError, match="bvecs must have shape"):
        shore_fit(data, bvals, bvecs[:, :2], mask)


def test_progress_registration_checks():
    # no device call: registration only
    lib = L.lib()
    fn = L.PROGRESS_FN(lambda d, t, u: None)
    assert lib.shorefit_set_progress(fn, None, 0.0) == L.E_INVALID
    assert "interval" in lib.shorefit_last_error().decode()
    assert lib.shorefit_set_progress(fn, None, 0.5) == 0
    assert lib.shorefit_set_progress(L.PROGRESS_FN(), None, 0.0) == 0  # NULL switches it off
    from cs_dmri_amd import shore_fit
    data, bvals, bvecs, mask = _series([1000.0, 2000.0])
    with pytest.raises(ValueError, match="callable"):
        shore_fit(data, bvals, bvecs, mask, progress=3)
    with pytest.raises(ValueError, match="progress_interval"):
        shore_fit(data, bvals, bvecs, mask, progress=print, progress_interval=0)


def test_nonneg_is_reported_unsupported():
    from cs_dmri_amd import shore_fit
    data, bvals, bvecs, mask = _series([1000.0, 2000.0])
    with pytest.raises(L.ShorefitError) as e:
        shore_fit(data, bvals, bvecs, mask, regularization="nonneg")
    assert e.value.code == L.E_UNSUPPORTED


def test_workload_shapes():
    assert len(W.cartesian_qgrid()) == 514
    assert int(W.ellipsoid_mask().sum()) == 250_176
    gt = W.cartesian_table(100, 1)
    assert gt.bvals.shape == (101,) and gt.bvals[0] == 0 and gt.bvals.max() <= 5000.0
    np.testing.assert_allclose(np.linalg.norm(gt.bvecs[1:], axis=1), 1.0)
    T = W.graddev_tensors(5)
    np.testing.assert_allclose(T, np.transpose(T, (0, 2, 1)))
