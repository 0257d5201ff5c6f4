"""On-disk formats (host side): gradient-file parsing, NIfTI-1 headers, the
affine rules, masks, the atomic writers and path helpers.

Pins: the reference's own unit tests (src/io/dwi.rs:422-516,
src/qspace.rs:284-317, src/io/mod.rs:229-256) restated; the NIfTI-1 header
layout checked against a header assembled field by field with ``struct`` at
the offsets of the NIfTI-1 standard (nifti1.h), independently of the
module's own field table; the quaternion rotation against closed forms.
The device layout conversion is covered in test_gpu_io.py.
"""

import json
import struct
from pathlib import Path

import numpy as np
import pytest

from cs_dmri_amd import io as sio
from cs_dmri_amd import nifti as nf


# ---- gradient files (src/io/dwi.rs:422-448) ----------------------------------
def test_parse_bvals_simple():
    assert sio.parse_bvals(" 0 1000 2000\n 3000 ").tolist() == [0.0, 1000.0, 2000.0, 3000.0]


def test_parse_bvecs_row_and_column_layouts():
    v = sio.parse_bvecs("1 0 0\n0 1 0\n0 0 1\n")
    assert v.shape == (3, 3) and v[0].tolist() == [1, 0, 0] and v[2].tolist() == [0, 0, 1]
    v = sio.parse_bvecs("1 0 0\n0 1 0\n0 0 1\n0 0 -1\n")
    assert v.shape == (4, 3) and v[3].tolist() == [0, 0, -1]
    rows = sio.parse_bvecs("1 0 0 0.6\n0 1 0 0.8\n0 0 1 0\n")
    assert rows.shape == (4, 3) and rows[3].tolist() == [0.6, 0.8, 0.0]
    with pytest.raises(ValueError, match="inconsistent"):
        sio.parse_bvecs("1 0\n0 1 0\n0 0 1\n")
    with pytest.raises(ValueError, match="could not interpret"):
        sio.parse_bvecs("1 0\n0 1\n")


def test_parse_mrtrix_grad_rows():
    b, d = sio.parse_mrtrix_grad("# comment\n0 0 0 0\n0.6 -0.8 0 1000\n\n0,0,1,2000\n")
    assert b.tolist() == [0.0, 1000.0, 2000.0]
    assert d.tolist() == [[0, 0, 0], [0.6, -0.8, 0], [0, 0, 1]]
    with pytest.raises(ValueError):
        sio.parse_mrtrix_grad("1 2 3\n")


# ---- affine rotation (src/qspace.rs:284-317) -----------------------------------
def test_affine_rotation_strips_voxel_sizes():
    aff = np.array([[-1.7, 0, 0, 79.9], [0, -1.7, 0, 80.75], [0, 0, 1.7, -80.0], [0, 0, 0, 1]])
    r = sio.affine_rotation(aff)
    np.testing.assert_allclose(np.linalg.norm(r, axis=0), 1.0, atol=1e-12)
    assert r[0, 0] == pytest.approx(-1.0, abs=1e-12) and r[1, 1] == pytest.approx(-1.0, abs=1e-12)
    assert r[2, 2] == pytest.approx(1.0, abs=1e-12)


def test_rotate_bvecs_lps_to_ras_flips_x_and_y():
    r = sio.affine_rotation(np.diag([-1.7, -1.7, 1.7, 1.0]))
    out = sio.rotate_bvecs(np.eye(3), r)
    assert out.tolist() == [[-1, 0, 0], [0, -1, 0], [0, 0, 1]]
    rng = np.random.default_rng(0)
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    v = rng.normal(size=(20, 3))
    np.testing.assert_allclose(sio.rotate_bvecs(v, q), v @ q.T, rtol=0, atol=1e-14)


# ---- masks (src/io/dwi.rs:347-365, 507-516) ------------------------------------
def test_default_mask_from_b0_uses_the_b0_volumes():
    data = np.zeros((3, 1, 1, 3), np.float32)
    data[0, 0, 0, 0] = 100.0  # bright only at b = 0
    data[1, 0, 0, 1] = 100.0  # bright only at b = 1000
    mask = sio.default_mask_from_b0(data, [0.0, 1000.0, 2000.0])
    assert mask.ravel().tolist() == [True, False, False]
    # no b0 frames: the mean over all volumes
    mask = sio.default_mask_from_b0(data, [1000.0, 1000.0, 2000.0])
    assert mask.ravel().tolist() == [True, True, False]


def test_load_mask_threshold_and_shape(tmp_path):
    m = np.zeros((4, 3, 2), np.float32)
    m[1, 2, 1] = 0.5
    m[3, 0, 0] = -1.0
    p = tmp_path / "mask.nii.gz"
    nf.save(p, nf.new_header(m.shape), m.transpose(2, 1, 0))
    got = sio.load_mask(p, (4, 3, 2))
    assert got.dtype == bool and got.sum() == 1 and got[1, 2, 1]
    with pytest.raises(ValueError, match="does not match"):
        sio.load_mask(p, (4, 3, 3))


# ---- NIfTI-1 header against the standard's offsets -----------------------------
def _spec_header(endian, dims, datatype, bitpix, pixdim, vox_offset, slope, inter, qform, sform, quat, qoff, srows):
    """348 bytes assembled field by field at the nifti1.h offsets."""
    b = bytearray(348)
    e = endian
    struct.pack_into(e + "i", b, 0, 348)
    struct.pack_into(e + "8h", b, 40, *dims)
    struct.pack_into(e + "h", b, 70, datatype)
    struct.pack_into(e + "h", b, 72, bitpix)
    struct.pack_into(e + "8f", b, 76, *pixdim)
    struct.pack_into(e + "f", b, 108, vox_offset)
    struct.pack_into(e + "f", b, 112, slope)
    struct.pack_into(e + "f", b, 116, inter)
    struct.pack_into("B", b, 123, 10)
    struct.pack_into(e + "h", b, 252, qform)
    struct.pack_into(e + "h", b, 254, sform)
    struct.pack_into(e + "3f", b, 256, *quat)
    struct.pack_into(e + "3f", b, 268, *qoff)
    struct.pack_into(e + "4f", b, 280, *srows[0])
    struct.pack_into(e + "4f", b, 296, *srows[1])
    struct.pack_into(e + "4f", b, 312, *srows[2])
    b[344:348] = b"n+1\x00"
    return bytes(b)


@pytest.mark.parametrize("endian", ["<", ">"])
def test_header_fields_at_standard_offsets(tmp_path, endian):
    srows = [(2.0, 0, 0, -90.0), (0, 2.0, 0, -126.0), (0, 0, 2.5, -72.0)]
    raw = _spec_header(endian, (4, 3, 4, 5, 6, 1, 1, 1), 4, 16, (1, 2, 2, 2.5, 1, 0, 0, 0), 352.0, 0.5, 10.0,
                       1, 2, (0.0, 0.0, 0.0), (1.0, 2.0, 3.0), srows)
    h = nf.parse_header(raw)
    assert h["dim"].tolist()[:5] == [4, 3, 4, 5, 6]
    assert int(h["datatype"]) == 4 and int(h["bitpix"]) == 16
    assert float(h["vox_offset"]) == 352.0
    assert float(h["scl_slope"]) == 0.5 and float(h["scl_inter"]) == 10.0
    assert int(h["xyzt_units"]) == 10
    np.testing.assert_array_equal(nf.affine(h)[:3], np.array(srows))
    # the data block follows at vox_offset, in the file's byte order
    vals = np.arange(3 * 4 * 5 * 6, dtype=np.int16) - 100
    p = tmp_path / "h.nii"
    p.write_bytes(raw + b"\x00" * 4 + vals.astype(np.dtype("i2").newbyteorder(endian)).tobytes())
    img = nf.load(p)
    assert img.shape == (3, 4, 5, 6) and img.data.shape == (6, 5, 4, 3)
    np.testing.assert_array_equal(img.data.ravel(), vals)
    assert img.scl == (0.5, 10.0)


def test_written_header_matches_the_standard_layout(tmp_path):
    aff = np.array([[-1.5, 0, 0, 90.0], [0, 1.5, 0, -120.0], [0, 0, 2.0, -60.0], [0, 0, 0, 1]])
    ref = nf.new_header((7, 8, 9, 4), aff, datatype=4, bitpix=16)
    h = nf.derived_header(ref, (7, 8, 9, 11))
    data = np.arange(7 * 8 * 9 * 11, dtype=np.float32).reshape(11, 9, 8, 7)
    p = tmp_path / "c.nii"
    nf.save(p, h, data)
    b = p.read_bytes()
    assert len(b) == 352 + data.nbytes
    assert struct.unpack_from("<i", b, 0)[0] == 348
    assert struct.unpack_from("<8h", b, 40) == (4, 7, 8, 9, 11, 1, 1, 1)
    assert struct.unpack_from("<hh", b, 70) == (16, 32)
    assert struct.unpack_from("<fff", b, 108) == (352.0, 0.0, 0.0)
    assert struct.unpack_from("<h", b, 254)[0] == 2
    np.testing.assert_allclose(struct.unpack_from("<4f", b, 280), aff[0], rtol=1e-6)
    assert b[344:348] == b"n+1\x00" and b[348:352] == b"\x00\x00\x00\x00"
    np.testing.assert_array_equal(np.frombuffer(b[352:], "<f4"), data.ravel())


@pytest.mark.parametrize("name", ["r.nii", "r.nii.gz"])
def test_save_load_round_trip(tmp_path, name):
    rng = np.random.default_rng(1)
    data = rng.normal(size=(6, 5, 4, 3)).astype(np.float32)  # file order (t, z, y, x)
    h = nf.new_header((3, 4, 5, 6), np.diag([1.25, 1.25, 2.0, 1.0]))
    nf.save(tmp_path / name, h, data)
    img = nf.load(tmp_path / name)
    assert img.shape == (3, 4, 5, 6) and img.datatype == 16
    np.testing.assert_array_equal(img.data, data)
    np.testing.assert_allclose(img.affine(), np.diag([1.25, 1.25, 2.0, 1.0]))


def test_quaternion_affine_closed_forms():
    h = nf.new_header((2, 2, 2))
    h["sform_code"] = 0
    h["qform_code"] = 1
    h["pixdim"][:4] = [1.0, 2.0, 3.0, 4.0]
    h["qoffset_x"], h["qoffset_y"], h["qoffset_z"] = 5.0, 6.0, 7.0
    a = nf.affine(h)  # identity quaternion: diag(pixdim) + offsets
    np.testing.assert_allclose(a, [[2, 0, 0, 5], [0, 3, 0, 6], [0, 0, 4, 7], [0, 0, 0, 1]], atol=1e-12)
    h["quatern_d"] = 1.0  # 180 degrees about z: a = 0 branch
    np.testing.assert_allclose(nf.affine(h)[:3, :3], np.diag([-2.0, -3.0, 4.0]), atol=1e-6)
    h["quatern_d"] = np.sin(np.pi / 4)  # 90 degrees about z
    np.testing.assert_allclose(nf.affine(h)[:3, :3], [[0, -3, 0], [2, 0, 0], [0, 0, 4]], atol=1e-6)
    h["quatern_d"] = 0.0
    h["pixdim"][0] = -1.0  # qfac flips the third column
    np.testing.assert_allclose(nf.affine(h)[:3, :3], np.diag([2.0, 3.0, -4.0]), atol=1e-12)
    h["qform_code"] = 0  # no qform, no sform: voxel sizes only
    np.testing.assert_allclose(nf.affine(h), np.diag([2.0, 3.0, 4.0, 1.0]), atol=1e-12)


def test_rejects_non_nifti(tmp_path):
    p = tmp_path / "x.nii"
    p.write_bytes(b"\x00" * 400)
    with pytest.raises(ValueError, match="sizeof_hdr"):
        nf.load(p)


# ---- paths and atomic writes (src/io/coeffs.rs:174-186, src/io/output.rs:113-123, src/io/mod.rs:229-256) --
def test_path_helpers():
    assert str(sio.sidecar_path_for("a/foo.nii.gz")) == "a/foo.json"
    assert str(sio.sidecar_path_for("foo.nii")) == "foo.json"
    assert str(sio.sidecar_path_for("foo")) == "foo.json"
    assert str(sio.sibling_path("a/foo.nii.gz", "_r2.nii.gz")) == "a/foo_r2.nii.gz"
    assert str(sio.sibling_path("foo", "_r2.nii.gz")) == "foo_r2.nii.gz"


def test_atomic_write_pair_lands_both_files_or_neither(tmp_path):
    a, b = tmp_path / "a.bin", tmp_path / "b.json"
    sio.atomic_write_pair(a, b, False, lambda t: t.write_bytes(b"a-data"), lambda t: t.write_bytes(b"b-data"))
    assert a.read_bytes() == b"a-data" and b.read_bytes() == b"b-data"

    def boom(t):
        raise RuntimeError("boom")
    with pytest.raises(RuntimeError, match="boom"):
        sio.atomic_write_pair(a, b, True, lambda t: t.write_bytes(b"new-a"), boom)
    assert a.read_bytes() == b"a-data" and b.read_bytes() == b"b-data"
    assert sorted(p.name for p in tmp_path.iterdir()) == ["a.bin", "b.json"]
    with pytest.raises(FileExistsError, match="refusing to overwrite"):
        sio.atomic_write_pair(a, b, False, lambda t: t.write_bytes(b"x"), lambda t: t.write_bytes(b"y"))


def test_sparsity_map():
    c = np.zeros((2, 1, 1, 4), np.float32)
    c[0, 0, 0, :3] = 1.0
    c[1, 0, 0, 0] = 2.0
    m = np.array([True, False]).reshape(2, 1, 1)
    assert sio.sparsity_map(c, m).ravel().tolist() == [0.75, 0.0]


def test_sidecar_json_shape():
    """Keys and tags as the reference's serde serialization writes them
    (src/io/coeffs.rs:26-98, src/solver/alpha.rs:193-205)."""
    from cs_dmri_amd.models import _sidecar
    from cs_dmri_amd import _lib as L

    out = L.ShorefitOutput()
    out.n_shells = 2
    out.dwi_shells[0], out.dwi_shells[1] = 1000.0, 2000.0
    sc = _sidecar(out, None, "l1", "fixed", 1e-3, 1e-3, 20, None, 0.05, 500, 1e-6, False, 1e-8, 1e-8,
                  "world-ras", 72, 6, 700.0, None)
    assert sc["basis"] == {"name": "shore", "radial_order": 6, "zeta": 700.0}
    assert sc["solver"]["kind"] == "fista"
    assert sc["solver"]["alpha_strategy"] == {"kind": "fixed", "alpha": 1e-3}
    assert sc["solver"]["chosen_alpha"] == {"scope": "global", "alpha": 1e-3}
    assert sc["bvec_frame"] == "world-ras" and sc["dwi_shells"] == [1000.0, 2000.0]
    assert "provenance" not in sc
    json.dumps(sc)


def test_ensure_nifti_extension(capsys):
    # src/io/output.rs:461-486
    assert sio.ensure_nifti_extension("/tmp/foo.nii.gz") == Path("/tmp/foo.nii.gz")
    assert sio.ensure_nifti_extension("/tmp/foo.nii") == Path("/tmp/foo.nii")
    assert sio.ensure_nifti_extension("/tmp/foo") == Path("/tmp/foo.nii.gz")
    assert sio.ensure_nifti_extension("foo.tar.gz") == Path("foo.tar.gz.nii.gz")
    assert "appending .nii.gz" in capsys.readouterr().err


def test_precheck_writable(tmp_path):
    # src/io/output.rs:488-524: a missing parent is created and the probe
    # removed; an existing file is refused without overwrite and untouched
    target = tmp_path / "nested" / "sub" / "out.nii.gz"
    sio.precheck_writable(target, False)
    assert target.parent.is_dir() and not any("precheck" in p.name for p in target.parent.iterdir())
    target.write_bytes(b"existing")
    with pytest.raises(FileExistsError, match="refusing to overwrite"):
        sio.precheck_writable(target, False)
    sio.precheck_writable(target, True)
    assert target.read_bytes() == b"existing"


def test_precheck_outputs_set_follows_the_fit(tmp_path):
    out = tmp_path / "coef.nii.gz"
    names = lambda **kw: sorted(p.name for p in sio.fit_output_paths(out, **kw))  # noqa: E731
    assert names() == ["coef.json", "coef.nii.gz"]
    assert names(diagnostics=True, regularization="l2") == sorted(
        ["coef.nii.gz", "coef.json", "coef_r2.nii.gz", "coef_rmse.nii.gz", "coef_sparsity.nii.gz"])
    assert "coef_rss_l2.nii.gz" in names(diagnostics=True, alpha_mode="l2-anchored")
    assert "coef_rss_l2.nii.gz" not in names(diagnostics=True, alpha_mode="fixed")
    assert {"coef_alpha.nii.gz", "coef_bic.nii.gz"} <= set(names(diagnostics=True, alpha_mode="fixed"))


def test_fit_files_checks_outputs_before_reading(tmp_path):
    # an existing diagnostic map stops the run before the inputs are read
    # (the inputs here do not even exist), so no partial output set is left
    out = tmp_path / "coef.nii.gz"
    (tmp_path / "coef_bic.nii.gz").write_bytes(b"old")
    with pytest.raises(FileExistsError, match="coef_bic"):
        sio.fit_files(tmp_path / "missing.nii.gz", tmp_path / "b.bval", tmp_path / "b.bvec", out,
                      alpha_mode="fixed", diagnostics=True)
    assert sorted(p.name for p in tmp_path.iterdir()) == ["coef_bic.nii.gz"]
    with pytest.raises(ValueError, match="provenance"):
        sio.ProvenanceRecorder("cs-fit", "verbose")


# ---- progress heartbeat (src/progress.rs:98-143, tests :147-185) -------------
def test_format_hms():
    assert sio.format_hms(0) == "00:00:00"
    assert sio.format_hms(59.9) == "00:00:59"
    assert sio.format_hms(3600 + 2 * 60 + 3) == "01:02:03"
    assert sio.format_hms(100 * 3600) == "100:00:00"


def test_heartbeat_quiet_prints_nothing(capsys):
    hb = sio.Heartbeat("t", 100, 1e-9, quiet=True)
    for i in range(101):
        hb.update(i, 100)
    hb.finish()
    assert hb.done == 100
    assert capsys.readouterr().err == ""


def test_heartbeat_lines_and_eta(capsys, monkeypatch):
    clock = iter([0.0, 10.0, 10.5, 30.0, 40.0])
    monkeypatch.setattr(sio.time, "perf_counter", lambda: next(clock))
    hb = sio.Heartbeat("cs-fit", 1000, 5.0)  # t0 = 0
    hb.update(250, 1000)   # t = 10: 250 in 10 s, 750 left -> eta 30 s
    hb.update(300, 1000)   # t = 10.5: inside the interval, silent
    hb.update(200, 1000)   # t = 30: counts never go backwards
    hb.finish()            # t = 40
    err = capsys.readouterr().err.splitlines()
    assert err == ["[cs-fit] 250/1000 (25.0%) elapsed=00:00:10 eta=00:00:30",
                   "[cs-fit] 300/1000 (30.0%) elapsed=00:00:30 eta=00:01:10",
                   "[cs-fit] done 300/1000 (30.0%) elapsed=00:00:40"]


def test_heartbeat_empty_mask_and_bad_interval(capsys):
    hb = sio.Heartbeat("cs-fit", 0, 1e-9)
    hb.update(0, 0)
    hb.finish()
    assert capsys.readouterr().err.splitlines() == ["[cs-fit] done 0/0 (100.0%) elapsed=00:00:00"]
    with pytest.raises(ValueError):
        sio.Heartbeat("cs-fit", 1, 0.0)
