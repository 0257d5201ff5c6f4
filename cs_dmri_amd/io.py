"""On-disk formats around the fit: DWI series, masks, gradient tables,
coefficient files and their JSON sidecars, and the ``cs-fit`` data path.

Mirrors the reference's loaders and writers:

* ``parse_bvals`` / ``parse_bvecs`` / ``parse_mrtrix_grad`` —
  src/io/dwi.rs:356-420 (FSL rows or columns; MRtrix ``x y z b`` rows);
* ``read_dwi_volume`` — src/io/dwi.rs:219-313: any stored type and scaling
  to f32, voxel-major (X, Y, Z, N). The file's bytes go to the device as they
  are; the type conversion, the scaling and the layout change run there
  (``shorefit_volume_to_voxel_major``);
* ``load_mask`` / ``default_mask_from_b0`` — src/io/dwi.rs:319-365;
* ``load_dwi`` / ``load_dwi_mrtrix_grad`` — src/io/dwi.rs:107-217, with the
  image-axis -> world-RAS rotation of the bvecs (src/qspace.rs:46-70);
* ``write_coefficients`` / ``read_coefficients`` — CoefficientsFile
  (src/io/coeffs.rs:100-199): the (X, Y, Z, K) float32 NIfTI on the DWI's
  grid plus the sidecar JSON, written as an atomic pair
  (src/io/mod.rs:60-110, 259-282);
* ``write_map`` — write_3d_f32 (src/io/output.rs:113-141, 364-383);
* ``fit_files`` — the data path of the ``cs-fit`` binary
  (src/bin/cs-fit.rs:204-434): load, fit, write coefficients, sidecar and
  (with ``diagnostics``) the ``_r2``, ``_rmse``, ``_alpha``, ``_bic``,
  ``_rss_l2`` and ``_sparsity`` maps.
"""

from __future__ import annotations

import json
import os
import socket
import sys
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib as L
from . import nifti as nf

B0_THRESHOLD = 50.0


# ------------------------------------------------------------ gradient files --
def parse_bvals(text: str) -> np.ndarray:
    """Whitespace-separated b-values (src/io/dwi.rs:356-363)."""
    try:
        return np.array([float(t) for t in text.split()], dtype=np.float64)
    except ValueError as e:
        raise ValueError(f"bval parse: {e}") from None


def parse_bvecs(text: str) -> np.ndarray:
    """FSL bvecs: three rows of N values, or N rows of three
    (src/io/dwi.rs:392-420). Returns (N, 3)."""
    try:
        lines = [[float(t) for t in ln.split()] for ln in text.splitlines()]
    except ValueError as e:
        raise ValueError(f"bvec parse: {e}") from None
    lines = [ln for ln in lines if ln]
    if len(lines) == 3:
        n = len(lines[0])
        if len(lines[1]) != n or len(lines[2]) != n:
            raise ValueError("bvec rows have inconsistent lengths")
        return np.array(lines, dtype=np.float64).T.copy()
    if lines and all(len(ln) == 3 for ln in lines):
        return np.array(lines, dtype=np.float64)
    raise ValueError(f"could not interpret bvec file (lines: {len(lines)}, widths: {[len(ln) for ln in lines]})")


def parse_mrtrix_grad(text: str) -> tuple[np.ndarray, np.ndarray]:
    """MRtrix ``.b``: one ``x y z b`` row per gradient (world RAS), ``#``
    comments and blank lines skipped, commas allowed (src/io/dwi.rs:366-390)."""
    bvals, dirs = [], []
    for line in text.splitlines():
        line = line.split("#")[0].strip()
        if not line:
            continue
        toks = [t for t in line.replace(",", " ").split() if t]
        try:
            vals = [float(t) for t in toks]
        except ValueError as e:
            raise ValueError(f".b token: {e}") from None
        if len(vals) != 4:
            raise ValueError(f".b rows must have 4 values (x y z b), got {len(vals)} in '{line}'")
        dirs.append(vals[:3])
        bvals.append(vals[3])
    return np.array(bvals, dtype=np.float64), np.array(dirs, dtype=np.float64).reshape(-1, 3)


def affine_rotation(aff: np.ndarray) -> np.ndarray:
    """Column-normalised 3x3 of the affine: voxel sizes stripped, a pure
    rotation / reflection from image axes to world (src/qspace.rs:46-60)."""
    r = np.zeros((3, 3))
    for j in range(3):
        col = np.asarray(aff, np.float64)[:3, j]
        norm = np.sqrt(col[0] * col[0] + col[1] * col[1] + col[2] * col[2])
        r[:, j] = col * (1.0 / norm if norm > 0.0 else 0.0)
    return r


def rotate_bvecs(bvecs: np.ndarray, rot: np.ndarray) -> np.ndarray:
    """Each bvec times the 3x3 (src/qspace.rs:62-70), in the same order of
    operations."""
    v = np.asarray(bvecs, np.float64)
    out = np.empty_like(v)
    for i in range(3):
        out[:, i] = rot[i, 0] * v[:, 0] + rot[i, 1] * v[:, 1] + rot[i, 2] * v[:, 2]
    return out


# ----------------------------------------------------------------- volumes --
def _torch_device(device):
    import torch
    return torch.device(device)


def to_voxel_major(img: nf.Nifti, device=None):
    """(X, Y, Z, N) float32 of a 4D image, converted on the device from the
    file's stored type, scaling and order. ``device`` None returns a numpy
    array, a torch device (or string) a tensor resident there."""
    shp = img.shape
    if len(shp) == 3:
        shp = shp + (1,)
    if len(shp) != 4:
        raise ValueError(f"expected a 4D series, got {len(shp)}D ({shp})")
    nx, ny, nz, nt = shp
    slope, inter = img.scl
    raw = np.ascontiguousarray(img.data)
    if img.datatype not in (2, 4, 8, 16, 64, 256, 512, 768):
        raw = raw.astype(np.float64)  # 64-bit integers: exact below 2^53, then converted on the device
        code = 64
    else:
        code = img.datatype
    if device is None:
        out = np.empty((nx, ny, nz, nt), np.float32)
        L.check(L.lib().shorefit_volume_to_voxel_major(L.ptr(raw), code, nx, ny, nz, nt, slope, inter, L.ptr(out),
                                                       0, None))
        return out
    import torch
    dev = _torch_device(device)
    d_raw = torch.from_numpy(np.require(raw.reshape(-1), requirements=["C", "W"]).view(np.uint8)).to(dev)
    out = torch.empty((nx, ny, nz, nt), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    L.check(L.lib().shorefit_volume_to_voxel_major(d_raw.data_ptr(), code, nx, ny, nz, nt, slope, inter,
                                                   out.data_ptr(), 1, stream))
    return out


def to_file_order(vol) -> np.ndarray:
    """File-order float32 data of a voxel-major (X, Y, Z, T) volume (numpy
    array or device tensor), laid out on the device; returned on the host."""
    shp = tuple(vol.shape)
    if len(shp) == 3:
        shp = shp + (1,)
    nx, ny, nz, nt = shp
    out = np.empty((nt, nz, ny, nx), np.float32)
    if isinstance(vol, np.ndarray):
        v = np.ascontiguousarray(vol, dtype=np.float32)
        L.check(L.lib().shorefit_volume_to_file_order(L.ptr(v), nx, ny, nz, nt, L.ptr(out), 0, None))
        return out
    import torch
    v = vol.to(torch.float32).contiguous()
    d_out = torch.empty(out.shape, dtype=torch.float32, device=v.device)
    stream = torch.cuda.current_stream(v.device).cuda_stream
    L.check(L.lib().shorefit_volume_to_file_order(v.data_ptr(), nx, ny, nz, nt, d_out.data_ptr(), 1, stream))
    return d_out.cpu().numpy()


def read_dwi_volume(path, device=None):
    """4D series as (X, Y, Z, N) float32 (src/io/dwi.rs:219-260)."""
    img = nf.load(path)
    if len(img.shape) != 4:
        raise ValueError(f"expected 4D DWI, got {len(img.shape)}D ({list(img.shape)})")
    return to_voxel_major(img, device)


def load_mask(path, expected_shape) -> np.ndarray:
    """3D mask, true where the (scaled) value is > 0 (src/io/dwi.rs:319-345)."""
    img = nf.load(path)
    if len(img.shape) != 3:
        raise ValueError(f"expected 3D mask, got {len(img.shape)}D ({list(img.shape)})")
    if tuple(img.shape) != tuple(expected_shape):
        raise ValueError(f"mask shape {list(img.shape)} does not match DWI spatial shape {list(expected_shape)}")
    v = np.asarray(img.data).astype(np.float32).transpose(2, 1, 0)
    slope, inter = img.scl
    if slope != 0.0:
        v = v * np.float32(slope) + np.float32(inter)
    return np.ascontiguousarray(v > 0.0)


def default_mask_from_b0(data, bvals, b0_threshold: float = B0_THRESHOLD) -> np.ndarray:
    """Voxels whose mean b <= b0_threshold signal (all volumes if there is no
    b0) exceeds 1% of the largest such mean; the mean is an f32 sum in frame
    order over the frame count (src/io/dwi.rs:347-365)."""
    d = data if isinstance(data, np.ndarray) else data.cpu().numpy()
    b0 = [t for t, b in enumerate(bvals) if b <= b0_threshold] or list(range(d.shape[3]))
    m = np.zeros(d.shape[:3], np.float32)
    for t in b0:
        m += d[..., t]
    m /= np.float32(len(b0))
    # a fold of f32::max from 0 (NaN means ignored), times 0.01 in f32
    thresh = np.fmax.reduce(m.ravel(), initial=np.float32(0.0)) * np.float32(0.01)
    return m > thresh


@dataclass
class DwiData:
    """What ``load_dwi`` returns (DwiData, src/io/dwi.rs:22-60): ``data``
    (X, Y, Z, N) float32 (numpy, or a tensor on the requested device), the
    mask, the gradient table in the fitting frame, the voxel-frame bvecs and
    the frame rotation."""
    data: object
    mask: np.ndarray
    bvals: np.ndarray
    bvecs: np.ndarray
    bvecs_vox: np.ndarray
    frame_rotation: np.ndarray
    bvec_frame: str
    big_delta: float | None
    small_delta: float | None
    affine: np.ndarray
    source: Path
    header: np.ndarray = field(repr=False, default=None)


def _load_dwi_vox(dwi_path, bvals, bvecs_vox, mask_path, big_delta, small_delta, bvec_frame, device) -> DwiData:
    if bvec_frame not in ("image-axis", "world-ras"):
        raise ValueError(f"bvec_frame must be 'image-axis' or 'world-ras', got {bvec_frame!r}")
    img = nf.load(dwi_path)
    if len(img.shape) != 4:
        raise ValueError(f"expected 4D DWI, got {len(img.shape)}D ({list(img.shape)})")
    if img.shape[3] != len(bvals):
        raise ValueError(f"DWI 4th dimension ({img.shape[3]}) does not match bval count ({len(bvals)})")
    data = to_voxel_major(img, device)
    mask = load_mask(mask_path, img.shape[:3]) if mask_path is not None else default_mask_from_b0(data, bvals)
    aff = img.affine()
    rot = affine_rotation(aff) if bvec_frame == "world-ras" else np.eye(3)
    bvecs = rotate_bvecs(bvecs_vox, rot) if bvec_frame == "world-ras" else np.array(bvecs_vox, np.float64)
    return DwiData(data, mask, np.asarray(bvals, np.float64), bvecs, np.array(bvecs_vox, np.float64), rot, bvec_frame,
                   big_delta, small_delta, aff, Path(dwi_path), img.header)


def load_dwi(dwi, bval, bvec, mask=None, *, big_delta=None, small_delta=None, bvec_frame="world-ras",
             device=None) -> DwiData:
    """DWI series, FSL gradient files and an optional mask
    (src/io/dwi.rs:107-127, 162-217). Without a mask the b0 fallback mask is
    used; ``world-ras`` rotates the bvecs by the affine's column-normalised
    3x3 (cs-fit's default)."""
    bvals = parse_bvals(Path(bval).read_text())
    bvecs = parse_bvecs(Path(bvec).read_text())
    if len(bvals) != len(bvecs):
        raise ValueError(f"bval count ({len(bvals)}) != bvec count ({len(bvecs)})")
    return _load_dwi_vox(dwi, bvals, bvecs, mask, big_delta, small_delta, bvec_frame, device)


def load_dwi_mrtrix_grad(dwi, grad, mask=None, *, big_delta=None, small_delta=None, bvec_frame="world-ras",
                         device=None) -> DwiData:
    """DWI with an MRtrix ``.b`` table: world directions taken back to voxel
    axes through the inverse of the affine's column-normalised 3x3, then as
    ``load_dwi`` (src/io/dwi.rs:129-160)."""
    bvals, world = parse_mrtrix_grad(Path(grad).read_text())
    p = affine_rotation(nf.affine(nf.read_header(dwi)))
    try:
        p_inv = np.linalg.inv(p)
    except np.linalg.LinAlgError:
        raise ValueError(f"affine of {dwi} is singular") from None
    vox = world @ p_inv.T
    return _load_dwi_vox(dwi, bvals, vox, mask, big_delta, small_delta, bvec_frame, device)


# -------------------------------------------------------- coefficient files --
def sidecar_path_for(path) -> Path:
    """foo.nii.gz / foo.nii / foo -> foo.json (src/io/coeffs.rs:174-186)."""
    s = str(path)
    for suf in (".nii.gz", ".nii"):
        if s.endswith(suf):
            return Path(s[: -len(suf)] + ".json")
    return Path(s + ".json")


def sibling_path(path, suffix: str) -> Path:
    """foo.nii.gz + "_r2.nii.gz" -> foo_r2.nii.gz (src/io/output.rs:113-123)."""
    s = str(path)
    for suf in (".nii.gz", ".nii"):
        if s.endswith(suf):
            return Path(s[: -len(suf)] + suffix)
    return Path(s + suffix)


def _temp_sibling(final: Path) -> Path:
    """Hidden temp name beside the final file, keeping its (compound)
    extension so the writer picks the same format (src/io/mod.rs:259-282)."""
    name = final.name or "out"
    i = name.find(".")
    stem, ext = (name[:i], name[i:]) if i > 0 else (name, "")
    return final.parent / f".{stem}.tmp.{os.getpid()}.{time.time_ns() % 1_000_000_000}{ext}"


def atomic_write_pair(final_a, final_b, overwrite: bool, write_a, write_b) -> None:
    """Both files land via temp + rename, or neither; refuses to clobber
    unless ``overwrite`` (src/io/mod.rs:60-110)."""
    final_a, final_b = Path(final_a), Path(final_b)
    for p in (final_a, final_b):
        if p.exists() and not overwrite:
            raise FileExistsError(f"refusing to overwrite existing {p} — pass overwrite=True to replace")
        if str(p.parent):
            p.parent.mkdir(parents=True, exist_ok=True)
    tmp_a, tmp_b = _temp_sibling(final_a), _temp_sibling(final_b)

    def cleanup():
        for t in (tmp_a, tmp_b):
            try:
                t.unlink()
            except FileNotFoundError:
                pass
    try:
        write_a(tmp_a)
        write_b(tmp_b)
    except BaseException:
        cleanup()
        raise
    try:
        os.replace(tmp_a, final_a)
    except BaseException:
        cleanup()
        raise
    try:
        os.replace(tmp_b, final_b)
    except BaseException:
        try:
            tmp_b.unlink()
        except FileNotFoundError:
            pass
        raise


def atomic_write(final, overwrite: bool, write) -> None:
    """One file via temp + rename (src/io/output.rs: atomic_write)."""
    final = Path(final)
    if final.exists() and not overwrite:
        raise FileExistsError(f"refusing to overwrite existing {final} — pass overwrite=True to replace")
    if str(final.parent):
        final.parent.mkdir(parents=True, exist_ok=True)
    tmp = _temp_sibling(final)
    try:
        write(tmp)
        os.replace(tmp, final)
    except BaseException:
        try:
            tmp.unlink()
        except FileNotFoundError:
            pass
        raise


def _grid_header(reference) -> np.ndarray:
    """The header whose grid a new file takes: a NIfTI path, or a header
    array (``nifti.read_header`` / ``nifti.new_header``)."""
    if isinstance(reference, (np.ndarray, np.void)) and reference.dtype.names:
        return reference
    return nf.read_header(reference)


def write_coefficients(path, coeffs, reference_nifti, sidecar: dict, overwrite: bool = False) -> None:
    """Coefficient NIfTI (X, Y, Z, K) float32 with the reference DWI's
    spatial header (a path or a header array), and its JSON sidecar, as an
    atomic pair (src/io/coeffs.rs:112-160). ``coeffs`` may be a device
    tensor."""
    shp = tuple(coeffs.shape)
    if len(shp) != 4:
        raise ValueError(f"coefficients must be 4D (X, Y, Z, K), got {shp}")
    if "n_coefficients" in sidecar and int(sidecar["n_coefficients"]) != shp[3]:
        raise ValueError(f"sidecar n_coefficients ({sidecar['n_coefficients']}) != coefficient count ({shp[3]})")
    header = nf.derived_header(_grid_header(reference_nifti), shp, 16, 32)
    data = to_file_order(coeffs)
    text = json.dumps(sidecar, indent=2)
    atomic_write_pair(path, sidecar_path_for(path), overwrite,
                      lambda tmp: nf.save(tmp, header, data),
                      lambda tmp: Path(tmp).write_text(text))


def read_coefficients(path, device=None):
    """(coefficients (X, Y, Z, K) float32, sidecar dict) of a coefficient
    file (src/io/coeffs.rs:162-172)."""
    img = nf.load(path)
    if len(img.shape) != 4:
        raise ValueError(f"coeffs not 4D: {list(img.shape)}")
    meta = json.loads(sidecar_path_for(path).read_text())
    meta.setdefault("bvec_frame", "image-axis")  # sidecars that predate the field (src/io/coeffs.rs:19-24)
    return to_voxel_major(img, device), meta


def write_map(path, values, reference_nifti, overwrite: bool = False) -> None:
    """3D float32 map on the reference's grid (a path or a header array;
    src/io/output.rs:125-141)."""
    v = values if isinstance(values, np.ndarray) else values.cpu().numpy()
    v = np.asarray(v, np.float32)
    if v.ndim != 3:
        raise ValueError(f"map must be 3D, got {v.shape}")
    header = nf.derived_header(_grid_header(reference_nifti), v.shape, 16, 32)
    atomic_write(path, overwrite, lambda tmp: nf.save(tmp, header, np.ascontiguousarray(v.transpose(2, 1, 0))))


def sparsity_map(coeffs: np.ndarray, mask: np.ndarray) -> np.ndarray:
    """Fraction of non-zero coefficients per masked voxel (src/fit.rs:810-827)."""
    nz = (coeffs != 0.0).sum(axis=-1).astype(np.float32)
    return np.where(mask, nz / np.float32(coeffs.shape[-1]), np.float32(0.0)).astype(np.float32)


# ------------------------------------------------------------ the cs-fit path --
def ensure_nifti_extension(path, label: str = "--output") -> Path:
    """``path`` itself when it ends in .nii or .nii.gz, else ``path`` + .nii.gz
    with a warning on stderr (src/io/output.rs:28-50): the writer picks the
    format from the extension."""
    s = str(path)
    if s.endswith(".nii.gz") or s.endswith(".nii"):
        return Path(path)
    new = Path(s + ".nii.gz")
    print(f"[cs_dmri] warning: {label} {s} has no .nii.gz/.nii extension; appending .nii.gz \u2192 {new}",
          file=sys.stderr)
    return new


def precheck_writable(path, overwrite: bool) -> None:
    """Refuses, before any work, an output that exists (without ``overwrite``)
    or whose directory cannot be created or written: the parent directory is
    created if missing and probed with a short-lived file
    (src/io/output.rs:52-103)."""
    path = Path(path)
    if path.exists() and not overwrite:
        raise FileExistsError(f"refusing to overwrite existing {path} \u2014 pass overwrite=True to replace")
    parent = path.parent if str(path.parent) not in ("", ".") else Path(".")
    try:
        parent.mkdir(parents=True, exist_ok=True)
    except OSError as e:
        raise PermissionError(f"cannot create parent directory {parent} for output {path}: {e}") from None
    probe = parent / f".cs_dmri-precheck.{os.getpid()}.{time.time_ns() % 1_000_000_000}"
    try:
        with open(probe, "wb") as f:
            f.write(b"x")
    except OSError as e:
        raise PermissionError(f"output directory {parent} is not writable (probe failed): {e}") from None
    finally:
        try:
            probe.unlink()
        except FileNotFoundError:
            pass


def fit_output_paths(output, *, diagnostics=False, regularization="l1", alpha_mode="l2-anchored") -> list[Path]:
    """Every file ``fit_files`` writes for these settings: the coefficient
    file, its sidecar and, with diagnostics, the maps the fit produces
    (``_r2``, ``_rmse``, ``_sparsity``; ``_alpha`` and ``_bic`` for L1;
    ``_rss_l2`` for the L2-anchored strategy)."""
    output = Path(output)
    paths = [output, sidecar_path_for(output)]
    if diagnostics:
        suffixes = ["_r2", "_rmse", "_sparsity"]
        if regularization == "l1":
            suffixes += ["_alpha", "_bic"]
            if alpha_mode == "l2-anchored":
                suffixes.append("_rss_l2")
        paths += [sibling_path(output, f"{sfx}.nii.gz") for sfx in suffixes]
    return paths


def precheck_outputs(output, overwrite: bool, **fit_kwargs) -> Path:
    """cs-fit's pre-flight (src/bin/cs-fit.rs:507-552): the coefficient path
    gets a NIfTI extension, then every output it will write is checked, so a
    path mistake surfaces before the data are read. Returns the output path."""
    output = ensure_nifti_extension(output)
    for p in fit_output_paths(output, diagnostics=bool(fit_kwargs.get("diagnostics")),
                              regularization=fit_kwargs.get("regularization", "l1"),
                              alpha_mode=fit_kwargs.get("alpha_mode", "l2-anchored")):
        precheck_writable(p, overwrite)
    return output


PROVENANCE_MODES = ("minimal", "full", "none")


class ProvenanceRecorder:
    """The sidecar's provenance record (src/io/provenance.rs:34-105): tool,
    library version, git SHA and build time of the library, host threads and
    wall time of the run; ``full`` adds the command line, host name and start
    time, which may identify a subject and are left out by default."""

    def __init__(self, tool: str = "cs-fit", mode: str = "minimal"):
        if mode not in PROVENANCE_MODES:
            raise ValueError(f"provenance must be one of {PROVENANCE_MODES}, got {mode!r}")
        self.tool, self.mode = tool, mode
        self._t0 = time.perf_counter()
        self._started = time.strftime("%Y-%m-%dT%H:%M:%SZ", time.gmtime())

    def finish(self, threads_used: int = 1) -> dict | None:
        if self.mode == "none":
            return None
        info = json.loads(L.lib().shorefit_build_info().decode())
        rec = {"tool": self.tool, "version": info["version"], "git_sha": info["git_sha"],
               "build_timestamp": info["build_timestamp"], "threads_used": int(threads_used),
               "duration_secs": time.perf_counter() - self._t0}
        if self.mode == "full":
            rec["argv"] = list(sys.argv)
            rec["hostname"] = os.environ.get("HOSTNAME") or socket.gethostname()
            rec["started_at_utc"] = self._started
        return rec


def format_hms(seconds: float) -> str:
    """Whole seconds as HH:MM:SS (src/progress.rs:137-143)."""
    t = int(max(seconds, 0.0))
    return f"{t // 3600:02d}:{(t % 3600) // 60:02d}:{t % 60:02d}"


class Heartbeat:
    """Plain-line progress on stderr for a voxel loop (src/progress.rs:17-143).

    ``update(done, total)`` is the progress callback of ``shore_fit``; it
    prints ``[label] done/total (pct%) elapsed=HH:MM:SS eta=HH:MM:SS`` at most
    once per ``interval`` seconds, and ``finish()`` prints the ``done`` line.
    ``quiet`` prints nothing. No carriage returns, so job logs stay readable.
    """

    def __init__(self, label: str, total: int, interval: float, quiet: bool = False, stream=None):
        if not interval > 0:
            raise ValueError(f"interval must be positive, got {interval}")
        self.label = label
        self.total = int(total)
        self.interval = float(interval)
        self.quiet = bool(quiet)
        self.done = 0
        self.stream = stream
        self._t0 = time.perf_counter()
        self._next = self._t0 + self.interval

    def _pct(self, done):
        return 100.0 if self.total == 0 else done * 100.0 / self.total

    def _eta(self, done, elapsed):
        if done == 0 or done >= self.total:
            return "--:--:--"
        return format_hms((self.total - done) * elapsed / done)

    def _print(self, line):
        print(line, file=self.stream if self.stream is not None else sys.stderr, flush=True)

    def update(self, done: int, total: int | None = None) -> None:
        self.done = max(self.done, int(done))
        now = time.perf_counter()
        if self.quiet or self.total == 0 or now < self._next:
            return
        self._next = now + self.interval
        el = now - self._t0
        self._print(f"[{self.label}] {self.done}/{self.total} ({self._pct(self.done):.1f}%) "
                    f"elapsed={format_hms(el)} eta={self._eta(self.done, el)}")

    def finish(self) -> None:
        if not self.quiet:
            el = time.perf_counter() - self._t0
            self._print(f"[{self.label}] done {self.done}/{self.total} ({self._pct(self.done):.1f}%) "
                        f"elapsed={format_hms(el)}")


def fit_files(dwi, bval, bvec, output, *, mask=None, overwrite=False, bvec_frame="world-ras", big_delta=None,
              small_delta=None, device=None, provenance="minimal", quiet=False, progress_interval_secs=30,
              **fit_kwargs) -> dict:
    """The ``cs-fit`` data path (src/bin/cs-fit.rs:204-434): check every
    output path first (``precheck_outputs``), load the series (``load_dwi``),
    fit with ``shore_fit`` (keywords as there; ``precision`` and
    ``refresh_every`` select the device arithmetic), write the coefficient
    NIfTI and sidecar (with the ``provenance`` record: "minimal", "full" or
    "none"), reporting progress on stderr every ``progress_interval_secs``
    (at least 1 s; ``quiet`` silences it, src/bin/cs-fit.rs:175-183,
    :285-312), and with ``diagnostics=True`` the ``_r2``, ``_rmse``, ``_alpha``,
    ``_bic``, ``_rss_l2`` and ``_sparsity`` maps beside it. Returns the
    sidecar dict. With ``device`` the series is converted and fitted without
    a host copy of the voxel-major volume."""
    from .models import shore_fit

    rec = ProvenanceRecorder("cs-fit", provenance)
    output = precheck_outputs(output, overwrite, **fit_kwargs)
    d = load_dwi(dwi, bval, bvec, mask, big_delta=big_delta, small_delta=small_delta, bvec_frame=bvec_frame,
                 device=device)
    n_mask = int((d.mask != 0).sum())
    hb = Heartbeat("cs-fit", n_mask, max(float(progress_interval_secs), 1.0), quiet)
    out = shore_fit(d.data, d.bvals, d.bvecs, d.mask, big_delta=big_delta, small_delta=small_delta,
                    bvec_frame=bvec_frame, progress=None if quiet else hb.update,
                    progress_interval=hb.interval, **fit_kwargs)
    hb.finish()
    sidecar = dict(out["sidecar"])
    prov = rec.finish(threads_used=1)
    if prov is not None:
        sidecar["provenance"] = prov
    coeffs = out["coefficients"]
    write_coefficients(output, coeffs, dwi, sidecar, overwrite=overwrite)
    if fit_kwargs.get("diagnostics"):
        write_map(sibling_path(output, "_r2.nii.gz"), out["r2"], dwi, overwrite)
        write_map(sibling_path(output, "_rmse.nii.gz"), out["rmse"], dwi, overwrite)
        for key in ("alpha", "bic", "rss_l2"):
            if key in out:
                write_map(sibling_path(output, f"_{key}.nii.gz"), out[key], dwi, overwrite)
        c = coeffs if isinstance(coeffs, np.ndarray) else coeffs.cpu().numpy()
        write_map(sibling_path(output, "_sparsity.nii.gz"), sparsity_map(c, d.mask), dwi, overwrite)
    return sidecar
