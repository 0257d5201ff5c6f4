"""NIfTI-1 single files (``.nii``, ``.nii.gz``): header, data and affine.

The reference reads and writes these through the ``nifti`` crate (DWI input,
src/io/dwi.rs:107-345; coefficient and map output, src/io/coeffs.rs:112-172,
src/io/output.rs:113-141, 364-383). This module does the byte-level part in
numpy: the 348-byte header (little- or big-endian), the data block at
``vox_offset`` in file order (volume after volume, x fastest), and the
voxel-to-world affine by the NIfTI-1 rules (sform when ``sform_code > 0``,
else the qform quaternion, else the voxel sizes). Turning file order into the
voxel-major layout the fits use (and back) is done on the device by
``shorefit_volume_to_voxel_major`` / ``shorefit_volume_to_file_order``.
Two-file ``.hdr``/``.img`` pairs and header extensions are not written; on
read, extensions are skipped through ``vox_offset``.
"""

from __future__ import annotations

import gzip
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

HEADER_SIZE = 348
VOX_OFFSET = 352  # header + the 4-byte extension flag

# NIfTI-1 header fields: (name, numpy type, byte offset).
_FIELDS = [
    ("sizeof_hdr", "i4", 0), ("data_type", "S10", 4), ("db_name", "S18", 14), ("extents", "i4", 32),
    ("session_error", "i2", 36), ("regular", "S1", 38), ("dim_info", "u1", 39), ("dim", ("i2", 8), 40),
    ("intent_p1", "f4", 56), ("intent_p2", "f4", 60), ("intent_p3", "f4", 64), ("intent_code", "i2", 68),
    ("datatype", "i2", 70), ("bitpix", "i2", 72), ("slice_start", "i2", 74), ("pixdim", ("f4", 8), 76),
    ("vox_offset", "f4", 108), ("scl_slope", "f4", 112), ("scl_inter", "f4", 116), ("slice_end", "i2", 120),
    ("slice_code", "u1", 122), ("xyzt_units", "u1", 123), ("cal_max", "f4", 124), ("cal_min", "f4", 128),
    ("slice_duration", "f4", 132), ("toffset", "f4", 136), ("glmax", "i4", 140), ("glmin", "i4", 144),
    ("descrip", "S80", 148), ("aux_file", "S24", 228), ("qform_code", "i2", 252), ("sform_code", "i2", 254),
    ("quatern_b", "f4", 256), ("quatern_c", "f4", 260), ("quatern_d", "f4", 264), ("qoffset_x", "f4", 268),
    ("qoffset_y", "f4", 272), ("qoffset_z", "f4", 276), ("srow_x", ("f4", 4), 280), ("srow_y", ("f4", 4), 296),
    ("srow_z", ("f4", 4), 312), ("intent_name", "S16", 328), ("magic", "S4", 344),
]


def header_dtype(endian: str = "<") -> np.dtype:
    def t(x):
        if isinstance(x, tuple):
            return (np.dtype(x[0]).newbyteorder(endian), x[1])
        return np.dtype(x).newbyteorder(endian) if x[0] in "iuf" else np.dtype(x)
    return np.dtype({"names": [f[0] for f in _FIELDS], "formats": [t(f[1]) for f in _FIELDS],
                     "offsets": [f[2] for f in _FIELDS], "itemsize": HEADER_SIZE})


# NIfTI-1 datatype code -> numpy element type
DATATYPES = {2: "u1", 4: "i2", 8: "i4", 16: "f4", 64: "f8", 256: "i1", 512: "u2", 768: "u4", 1024: "i8",
             1280: "u8"}


@dataclass
class Nifti:
    """A parsed file: ``header`` (a numpy record of the 348-byte header, in
    native byte order) and ``data`` (the elements in file order, shape
    ``dim[dim[0]] ... dim[1]``, i.e. reversed, native byte order)."""
    header: np.ndarray
    data: np.ndarray

    @property
    def shape(self) -> tuple:
        d = self.header["dim"]
        return tuple(int(v) for v in d[1:1 + int(d[0])])

    @property
    def datatype(self) -> int:
        return int(self.header["datatype"])

    @property
    def scl(self) -> tuple[float, float]:
        """(slope, inter); slope 0 (or non-finite) means unscaled."""
        s, i = float(self.header["scl_slope"]), float(self.header["scl_inter"])
        if not np.isfinite(s):
            s = 0.0
        return s, (i if np.isfinite(i) else 0.0)

    def affine(self) -> np.ndarray:
        return affine(self.header)


def _open(path: Path, mode: str):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


def parse_header(raw: bytes) -> np.ndarray:
    """The 348-byte header in native byte order; raises ValueError if it is
    not a single-file NIfTI-1 header."""
    if len(raw) < HEADER_SIZE:
        raise ValueError(f"NIfTI header needs {HEADER_SIZE} bytes, got {len(raw)}")
    for endian in "<>":
        arr = np.frombuffer(raw[:HEADER_SIZE], dtype=header_dtype(endian), count=1)
        if int(arr[0]["sizeof_hdr"]) == HEADER_SIZE:
            break
    else:
        raise ValueError("not a NIfTI-1 file (sizeof_hdr is not 348 in either byte order)")
    if bytes(arr[0]["magic"])[:3] != b"n+1":
        raise ValueError(f"unsupported NIfTI magic {bytes(arr[0]['magic'])!r} (only single-file n+1 is read)")
    return arr.astype(header_dtype("="))[0]


def read_header(path) -> np.ndarray:
    with _open(Path(path), "rb") as f:
        return parse_header(f.read(HEADER_SIZE))


def load(path, mmap: bool = True) -> Nifti:
    """Header and data of a ``.nii`` / ``.nii.gz`` file. Uncompressed
    little-endian data is memory-mapped (read-only) unless ``mmap`` is False."""
    path = Path(path)
    with _open(path, "rb") as f:
        head = f.read(HEADER_SIZE)
        le = np.frombuffer(head[:4], "<i4")[0] == HEADER_SIZE
        h = parse_header(head)
        code = int(h["datatype"])
        if code not in DATATYPES:
            raise ValueError(f"unsupported NIfTI datatype {code}")
        nd = int(h["dim"][0])
        if not 1 <= nd <= 7:
            raise ValueError(f"invalid NIfTI dim[0] = {nd}")
        shape = tuple(int(v) for v in h["dim"][1:1 + nd])
        if any(s < 0 for s in shape):
            raise ValueError(f"negative NIfTI dimension in {shape}")
        off = int(h["vox_offset"])
        if off < HEADER_SIZE:
            raise ValueError(f"{path}: vox_offset {off} lies inside the header")
        et = np.dtype(DATATYPES[code]).newbyteorder("<" if le else ">")
        count = int(np.prod(shape, dtype=np.int64))
        if mmap and not str(path).endswith(".gz"):
            data = np.memmap(path, dtype=et, mode="r", offset=off, shape=(count,))
        else:
            f.read(off - HEADER_SIZE)  # header extensions, if any
            buf = f.read(count * et.itemsize)
            if len(buf) != count * et.itemsize:
                raise ValueError(f"{path}: data block holds {len(buf)} bytes, the header declares "
                                 f"{count * et.itemsize}")
            data = np.frombuffer(buf, dtype=et)
    if not le and et.itemsize > 1:
        data = data.byteswap().view(et.newbyteorder("="))
    return Nifti(h, np.asarray(data).reshape(shape[::-1]))


def quaternion_matrix(b: float, c: float, d: float) -> np.ndarray:
    """Rotation of the qform quaternion (b, c, d), a = sqrt(1 - b² - c² - d²);
    when that is below 1e-7 the quaternion is renormalised and a = 0 (the
    NIfTI-1 reference implementation, nifti_quatern_to_mat44)."""
    a = 1.0 - (b * b + c * c + d * d)
    if a < 1e-7:
        s = 1.0 / np.sqrt(b * b + c * c + d * d)
        b, c, d, a = b * s, c * s, d * s, 0.0
    else:
        a = np.sqrt(a)
    return np.array([
        [a * a + b * b - c * c - d * d, 2 * b * c - 2 * a * d, 2 * b * d + 2 * a * c],
        [2 * b * c + 2 * a * d, a * a + c * c - b * b - d * d, 2 * c * d - 2 * a * b],
        [2 * b * d - 2 * a * c, 2 * c * d + 2 * a * b, a * a + d * d - c * c - b * b]])


def affine(h: np.ndarray) -> np.ndarray:
    """Voxel -> world 4x4 (float64): the sform rows when sform_code > 0, else
    the qform (quaternion, qfac = pixdim[0] sign, voxel sizes, offsets) when
    qform_code > 0, else diag(pixdim[1:4])."""
    aff = np.eye(4)
    pix = h["pixdim"].astype(np.float64)
    if int(h["sform_code"]) > 0:
        aff[0] = h["srow_x"]
        aff[1] = h["srow_y"]
        aff[2] = h["srow_z"]
    elif int(h["qform_code"]) > 0:
        R = quaternion_matrix(float(h["quatern_b"]), float(h["quatern_c"]), float(h["quatern_d"]))
        qfac = -1.0 if pix[0] < 0 else 1.0
        aff[:3, :3] = R * np.array([pix[1], pix[2], qfac * pix[3]])
        aff[:3, 3] = [float(h["qoffset_x"]), float(h["qoffset_y"]), float(h["qoffset_z"])]
    else:
        aff[:3, :3] = np.diag(pix[1:4])
    return aff


def derived_header(reference: np.ndarray, shape, datatype: int = 16, bitpix: int = 32) -> np.ndarray:
    """A header for new data on the reference's grid: the reference's fields
    (voxel sizes, qform / sform, units, description) with the given shape and
    type, no scaling, data at byte 352 (src/io/coeffs.rs:127-141,
    src/io/output.rs:364-383)."""
    h = np.array(reference, dtype=header_dtype("=")).copy()[()]
    dim = np.ones(8, np.int16)
    dim[0] = len(shape)
    dim[1:1 + len(shape)] = shape
    h["dim"] = dim
    h["datatype"] = datatype
    h["bitpix"] = bitpix
    h["scl_slope"] = 0.0
    h["scl_inter"] = 0.0
    h["vox_offset"] = float(VOX_OFFSET)
    h["sizeof_hdr"] = HEADER_SIZE
    h["magic"] = b"n+1\x00"
    return h


def new_header(shape, affine4: np.ndarray | None = None, datatype: int = 16, bitpix: int = 32) -> np.ndarray:
    """A header with only the shape, type and (sform-only) affine set."""
    h = np.zeros((), dtype=header_dtype("="))[()]
    h["sizeof_hdr"] = HEADER_SIZE
    h["magic"] = b"n+1\x00"
    h["pixdim"] = np.ones(8, np.float32)
    if affine4 is not None:
        a = np.asarray(affine4, np.float64)
        h["srow_x"], h["srow_y"], h["srow_z"] = a[0], a[1], a[2]
        h["sform_code"] = 2
        h["pixdim"][1:4] = np.sqrt((a[:3, :3] ** 2).sum(axis=0))
    h["xyzt_units"] = 10  # mm, s
    return derived_header(h, shape, datatype, bitpix)


def save(path, header: np.ndarray, data_file_order: np.ndarray) -> None:
    """Write header + 4 zero extension bytes + the data (already in file
    order, the header's type) to ``path`` (gzip-compressed for ``.gz``)."""
    h = np.array(header, dtype=header_dtype("<"))
    dt = np.dtype(DATATYPES[int(h["datatype"])]).newbyteorder("<")
    d = np.ascontiguousarray(data_file_order, dtype=dt)
    n = int(np.prod(h["dim"][1:1 + int(h["dim"][0])], dtype=np.int64))
    if d.size != n:
        raise ValueError(f"data holds {d.size} elements, the header declares {n}")
    with _open(Path(path), "wb") as f:
        f.write(h.tobytes())
        f.write(b"\x00" * (VOX_OFFSET - HEADER_SIZE))
        f.write(d.tobytes(order="C"))
        f.flush()
        if hasattr(f, "fileno") and not str(path).endswith(".gz"):
            os.fsync(f.fileno())
