"""Minimal NIfTI-1 reader/writer (single-file ``.nii`` / ``.nii.gz``).

It replaces the few ``nibabel`` calls of the reference pipeline:

* ``nib.load(f).get_fdata()`` / ``.affine`` (BINDER/cli/run_BINDER.py:127-130,
  BINDER/tests/cpu/test_3D_MI_registration.py:17-24);
* ``nib.save(nib.Nifti1Image(data, affine), f)`` (BINDER/utils.py:222-240);
* ``nib.affines.apply_affine`` (BINDER/FastRegistration.py:375-393).

The format is the published NIfTI-1 header: 348 bytes, the data after
``vox_offset``, and voxels stored with the first index fastest.  Scaling by
``scl_slope`` / ``scl_inter`` is applied on read as ``get_fdata`` does.
The affine is taken from the sform if ``sform_code > 0``, else from the
qform quaternion if ``qform_code > 0``, else from the pixel sizes.
"""
import gzip
import os
import struct

import numpy as np

_DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64,
           256: np.int8, 512: np.uint16, 768: np.uint32, 1024: np.int64, 1280: np.uint64}
_CODES = {np.dtype(v): k for k, v in _DTYPES.items()}


class NiftiImage:
    """Data (float64 via ``get_fdata``, like nibabel) plus the 4x4 affine."""

    def __init__(self, data, affine, header=None):
        self.dataobj = np.asarray(data)
        self.affine = np.asarray(affine, dtype=np.float64)
        self.header = header or {}

    @property
    def shape(self):
        return self.dataobj.shape

    def get_fdata(self):
        return np.asarray(self.dataobj, dtype=np.float64)


def _open(path, mode):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


def _quatern_affine(b, c, d, qfac, pixdim, offset):
    """qform of the NIfTI-1 standard: rotation from the unit quaternion
    (a, b, c, d) with a = sqrt(1 - b^2 - c^2 - d^2), voxel sizes pixdim[1:4]
    and the z size negated when qfac < 0."""
    a2 = 1.0 - (b * b + c * c + d * d)
    if a2 > 1e-7:
        a = np.sqrt(a2)
    else:  # a 180-degree rotation: renormalise (b, c, d), a = 0
        n = np.sqrt(b * b + c * c + d * d)
        b, c, d, a = b / n, c / n, d / n, 0.0
    R = np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                  [2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)],
                  [2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - c * c - b * b]])
    zooms = np.array(pixdim[1:4], dtype=np.float64)
    zooms[zooms == 0] = 1.0
    if qfac < 0:
        zooms[2] = -zooms[2]
    A = np.eye(4)
    A[:3, :3] = R * zooms[None, :]
    A[:3, 3] = offset
    return A


def load(path):
    """Read a NIfTI-1 file."""
    with _open(path, "rb") as f:
        raw = f.read()
    if len(raw) < 348:
        raise ValueError(f"{path}: too short for a NIfTI-1 header")
    endian = "<"
    if struct.unpack("<i", raw[:4])[0] != 348:
        endian = ">"
        if struct.unpack(">i", raw[:4])[0] != 348:
            raise ValueError(f"{path}: not a NIfTI-1 file (sizeof_hdr != 348)")
    u = lambda fmt, off: struct.unpack_from(endian + fmt, raw, off)  # noqa: E731
    dim = u("8h", 40)
    datatype = u("h", 70)[0]
    pixdim = u("8f", 76)
    vox_offset = int(u("f", 108)[0])
    slope, inter = u("2f", 112)
    qform_code, sform_code = u("2h", 252)
    qb, qc, qd, qx, qy, qz = u("6f", 256)
    srow = np.array(u("12f", 280), dtype=np.float64).reshape(3, 4)
    magic = raw[344:348]
    if magic not in (b"n+1\x00", b"ni1\x00"):
        raise ValueError(f"{path}: bad NIfTI magic {magic!r}")
    if datatype not in _DTYPES:
        raise ValueError(f"{path}: unsupported NIfTI datatype {datatype}")
    nd = dim[0]
    shape = tuple(int(x) for x in dim[1:1 + nd])
    # trailing singleton dimensions beyond 3 are dropped (as nibabel does for
    # 3D images stored with dim[0] = 4 and dim[4] = 1)
    while len(shape) > 3 and shape[-1] == 1:
        shape = shape[:-1]
    dt = np.dtype(_DTYPES[datatype]).newbyteorder(endian)
    count = int(np.prod(shape))
    data = np.frombuffer(raw, dtype=dt, count=count, offset=vox_offset)
    data = data.reshape(shape, order="F")
    if slope not in (0.0, 1.0) or inter != 0.0:
        if slope == 0.0:
            slope = 1.0
        data = data.astype(np.float64) * slope + inter
    if sform_code > 0:
        affine = np.eye(4)
        affine[:3, :] = srow
    elif qform_code > 0:
        affine = _quatern_affine(qb, qc, qd, pixdim[0], pixdim, [qx, qy, qz])
    else:
        affine = np.diag([pixdim[1] or 1.0, pixdim[2] or 1.0, pixdim[3] or 1.0, 1.0])
    hdr = {"datatype": datatype, "pixdim": pixdim, "qform_code": qform_code,
           "sform_code": sform_code}
    return NiftiImage(data, affine, hdr)


def save(img_or_data, path, affine=None):
    """Write a NIfTI-1 single file (gzip if the name ends with .gz).  The
    affine is stored as the sform (code 2, aligned) and the voxel sizes as
    pixdim; the data keep their dtype (bool is stored as uint8)."""
    if isinstance(img_or_data, NiftiImage):
        data, affine = img_or_data.dataobj, img_or_data.affine
    else:
        data = img_or_data
    data = np.asarray(data)
    if data.dtype == bool:
        data = data.astype(np.uint8)
    if data.dtype.byteorder == ">":
        data = data.astype(data.dtype.newbyteorder("<"))
    if np.dtype(data.dtype.type) not in _CODES:
        data = data.astype(np.float64)
    affine = np.eye(4) if affine is None else np.asarray(affine, dtype=np.float64)
    code = _CODES[np.dtype(data.dtype.type)]
    hdr = bytearray(352)
    struct.pack_into("<i", hdr, 0, 348)
    dims = [data.ndim] + list(data.shape) + [1] * (7 - data.ndim)
    struct.pack_into("<8h", hdr, 40, *dims)
    struct.pack_into("<2h", hdr, 70, code, data.dtype.itemsize * 8)
    zooms = np.sqrt((affine[:3, :3] ** 2).sum(axis=0))
    pix = [1.0] + list(zooms) + [1.0] * 4
    struct.pack_into("<8f", hdr, 76, *pix)
    struct.pack_into("<f", hdr, 108, 352.0)
    struct.pack_into("<2f", hdr, 112, 1.0, 0.0)
    struct.pack_into("<b", hdr, 123, 10)  # xyzt_units: mm, s
    struct.pack_into("<2h", hdr, 252, 0, 2)
    struct.pack_into("<12f", hdr, 280, *affine[:3, :].ravel())
    hdr[344:348] = b"n+1\x00"
    payload = np.ascontiguousarray(data.T).astype(data.dtype.newbyteorder("<"), copy=False)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with _open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(payload.tobytes(order="C"))


def apply_affine(aff, pts):
    """Apply a 4x4 affine to points in the last dimension (nibabel.affines)."""
    aff = np.asarray(aff, dtype=np.float64)
    pts = np.asarray(pts, dtype=np.float64)
    return pts @ aff[:3, :3].T + aff[:3, 3]
