"""NIfTI-1 images, the control-point grid header and the affine text file against niftilib and
the reference's writers (host code; no GPU).

Files with chosen header fields are written byte by byte here, then read by the library
(niftyreg_amd.nifti) and by the reference build (nifti_image_read + reg_checkAndCorrectDimension):
dimensions, spacings, both matrices, codes and voxel values must agree exactly.  The grid header
is compared with reg_createControlPointGrid's, and files written by either side are read by the
other.
"""
import ctypes
import gzip
import os

import numpy as np
import pytest

import oracles as O
from niftyreg_amd import nifti

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build absent")

HDR = np.dtype([
    ("sizeof_hdr", "i4"), ("data_type", "S10"), ("db_name", "S18"), ("extents", "i4"), ("session_error", "i2"),
    ("regular", "S1"), ("dim_info", "u1"), ("dim", "i2", 8), ("intent_p1", "f4"), ("intent_p2", "f4"),
    ("intent_p3", "f4"), ("intent_code", "i2"), ("datatype", "i2"), ("bitpix", "i2"), ("slice_start", "i2"),
    ("pixdim", "f4", 8), ("vox_offset", "f4"), ("scl_slope", "f4"), ("scl_inter", "f4"), ("slice_end", "i2"),
    ("slice_code", "u1"), ("xyzt_units", "u1"), ("cal_max", "f4"), ("cal_min", "f4"), ("slice_duration", "f4"),
    ("toffset", "f4"), ("glmax", "i4"), ("glmin", "i4"), ("descrip", "S80"), ("aux_file", "S24"),
    ("qform_code", "i2"), ("sform_code", "i2"), ("quatern_b", "f4"), ("quatern_c", "f4"), ("quatern_d", "f4"),
    ("qoffset_x", "f4"), ("qoffset_y", "f4"), ("qoffset_z", "f4"), ("srow_x", "f4", 4), ("srow_y", "f4", 4),
    ("srow_z", "f4", 4), ("intent_name", "S16"), ("magic", "S4")])
assert HDR.itemsize == 348

DTYPES = {2: "u1", 4: "i2", 8: "i4", 16: "f4", 64: "f8", 256: "i1", 512: "u2"}


def _write_nifti(path, data, datatype=16, pixdim=(1, 1, 1, 1), qform=None, sform=None, qfac=1.0, slope=0.0,
                 inter=0.0, units=2, big_endian=False):
    """A single-file NIfTI-1 image with the given fields; data in (z, y, x) order."""
    bo = ">" if big_endian else "<"
    h = np.zeros((), HDR.newbyteorder(bo))
    h["sizeof_hdr"] = 348
    h["dim"][: data.ndim + 1] = [data.ndim] + list(data.shape[::-1])
    h["dim"][data.ndim + 1:] = 1
    h["datatype"] = datatype
    h["bitpix"] = 8 * np.dtype(DTYPES[datatype]).itemsize
    h["pixdim"][0] = qfac
    h["pixdim"][1:1 + len(pixdim)] = pixdim
    h["vox_offset"] = 352
    h["scl_slope"], h["scl_inter"] = slope, inter
    h["xyzt_units"] = units
    if qform is not None:
        h["qform_code"] = 1
        (h["quatern_b"], h["quatern_c"], h["quatern_d"]), (h["qoffset_x"], h["qoffset_y"], h["qoffset_z"]) = qform
    if sform is not None:
        h["sform_code"] = 2
        h["srow_x"], h["srow_y"], h["srow_z"] = sform[0], sform[1], sform[2]
    h["magic"] = b"n+1"
    raw = h.tobytes() + b"\0\0\0\0" + np.ascontiguousarray(data, np.dtype(DTYPES[datatype]).newbyteorder(bo)).tobytes()
    if str(path).endswith(".gz"):
        with gzip.open(path, "wb") as f:
            f.write(raw)
    else:
        with open(path, "wb") as f:
            f.write(raw)


def _oblique_sform():
    th = 0.25
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = R @ np.diag([1.1, 0.9, 1.3])
    return np.hstack([m, [[-40.5], [12.25], [3.0]]]).astype(np.float32)


def _cases(tmp):
    rng = np.random.default_rng(0)
    q = ((0.1, -0.2, 0.3), (-30.0, 20.5, -11.25))
    return {
        "float_qform_sform": dict(path=tmp / "a.nii", data=rng.normal(size=(13, 17, 19)).astype(np.float32),
                                  kw=dict(pixdim=(1.1, 0.9, 1.3), qform=q, sform=_oblique_sform())),
        "int16_qform_neg_qfac_gz": dict(path=tmp / "b.nii.gz", data=rng.integers(-300, 300, (11, 12, 14)).astype(np.int16),
                                        kw=dict(datatype=4, pixdim=(0.8, 0.8, 2.0), qform=q, qfac=-1.0, slope=2.0,
                                                inter=-5.0)),
        "uint8_no_orientation": dict(path=tmp / "c.nii", data=rng.integers(0, 255, (9, 10, 12)).astype(np.uint8),
                                     kw=dict(datatype=2, pixdim=(1.5, 1.25, 0.0))),
        "float64_big_endian_sform": dict(path=tmp / "d.nii", data=rng.normal(size=(8, 9, 10)),
                                         kw=dict(datatype=64, pixdim=(2.0, 2.0, 2.0), sform=_oblique_sform(),
                                                 big_endian=True)),
        "micron_units": dict(path=tmp / "e.nii", data=rng.normal(size=(7, 8, 9)).astype(np.float32),
                             kw=dict(pixdim=(500.0, 500.0, 750.0), qform=q, units=3)),
    }


def _assert_fields_equal(info, f, what):
    assert list(info.dim) == list(f.dim), what
    assert np.array_equal(np.array(info.pixdim[1:]), np.array(f.pixdim[1:])), what
    assert info.qform_code == f.qform_code and info.sform_code == f.sform_code, what
    assert np.array_equal(np.array(info.qto_xyz[:]), np.array(f.qto_xyz[:])), what
    if f.sform_code > 0:
        assert np.array_equal(np.array(info.sto_xyz[:]), np.array(f.sto_xyz[:])), what


@needs_ref
@pytest.mark.parametrize("name", ["float_qform_sform", "int16_qform_neg_qfac_gz", "uint8_no_orientation",
                                  "float64_big_endian_sform", "micron_units"])
def test_read_matches_niftilib(tmp_path, name):
    c = _cases(tmp_path)[name]
    _write_nifti(c["path"], c["data"], **c["kw"])
    geom, data, info = nifti.read_image(c["path"])
    f, want = O.ref_nifti_read(c["path"])
    _assert_fields_equal(info, f, name)
    assert info.scl_slope == f.scl_slope and info.scl_inter == f.scl_inter
    assert np.array_equal(data, want)
    assert geom.dims == tuple(f.dim[1:4])
    sel = np.array(f.sto_xyz[:] if f.sform_code > 0 else f.qto_xyz[:], np.float32).reshape(4, 4)
    assert np.array_equal(geom.xyz, sel)
    assert np.array_equal(geom.ijk, O.ref_mat44_inverse(sel))


@needs_ref
@pytest.mark.parametrize("name", ["float_qform_sform", "int16_qform_neg_qfac_gz", "uint8_no_orientation",
                                  "float64_big_endian_sform"])
@pytest.mark.parametrize("spacing", [(5.0, 5.0, 5.0), (4.4, 3.6, 6.5)])
def test_control_point_grid_header_matches_reference(tmp_path, name, spacing):
    c = _cases(tmp_path)[name]
    _write_nifti(c["path"], c["data"], **c["kw"])
    ref_grid_path = tmp_path / "ref_grid.nii"
    f = O.ref_control_point_grid(c["path"], spacing, ref_grid_path)
    g = nifti.control_point_grid_info(nifti.read_info(c["path"]), spacing)
    _assert_fields_equal(g, f, name)
    assert np.array_equal(np.array(g.quatern[:]), np.array(f.quatern[:]))
    assert np.array_equal(np.array(g.qoffset[:]), np.array(f.qoffset[:]))
    assert g.intent_code == f.intent_code == 1007
    assert g.intent_name == f.intent_name == b"NREG_TRANS"
    assert g.intent_p[0] == f.intent_p1 == 2.0  # CUB_SPLINE_GRID
    # the file our writer makes reads back, in niftilib, exactly as the file the reference writes
    # (a grid without qform_code keeps only niftilib's default qform, in both files)
    n = g.dim[1] * g.dim[2] * g.dim[3] * 3
    cpp = np.random.default_rng(1).normal(size=n).astype(np.float32)
    ours = tmp_path / "grid.nii.gz"
    nifti.write_image(ours, g, cpp)
    f2, back = O.ref_nifti_read(ours)
    f3, _ = O.ref_nifti_read(ref_grid_path)
    _assert_fields_equal(f2, f3, name)
    assert np.array_equal(np.array(f2.quatern[:]), np.array(f3.quatern[:]))
    assert np.array_equal(np.array(f2.qoffset[:]), np.array(f3.qoffset[:]))
    assert (f2.intent_code, f2.intent_name, f2.intent_p1) == (f3.intent_code, f3.intent_name, f3.intent_p1)
    assert np.array_equal(back, cpp)
    # and our reader sees the reference's file as niftilib does
    _, _, info3 = nifti.read_image(ref_grid_path)
    _assert_fields_equal(info3, f3, name)


@needs_ref
def test_affine_files_match_reference(tmp_path):
    m = np.array([[1.0234567, 0.1, -0.2, 12.3456789], [-0.05, 0.97, 0.15, -8.000001],
                  [0.25, -0.185, 1.0000001, 1e-8], [0, 0, 0, 1]], np.float32)
    nifti.write_affine(tmp_path / "ours.txt", m)
    O.ref_affine_write(tmp_path / "ref.txt", m)
    assert (tmp_path / "ours.txt").read_bytes() == (tmp_path / "ref.txt").read_bytes()
    assert np.array_equal(nifti.read_affine(tmp_path / "ref.txt"), O.ref_affine_read(tmp_path / "ours.txt"))
    assert np.array_equal(nifti.read_affine(tmp_path / "ours.txt"), O.ref_affine_read(tmp_path / "ref.txt"))


def test_round_trip_and_errors(tmp_path):
    data = np.random.default_rng(2).normal(size=(6, 7, 8)).astype(np.float32)
    _write_nifti(tmp_path / "x.nii", data, pixdim=(1.0, 2.0, 3.0), sform=_oblique_sform())
    geom, got, info = nifti.read_image(tmp_path / "x.nii")
    assert np.array_equal(got, data.ravel()) and geom.dims == (8, 7, 6) and geom.sform == 1
    nifti.write_image(tmp_path / "y.nii", info, got)
    _, again, info2 = nifti.read_image(tmp_path / "y.nii")
    assert np.array_equal(again, data.ravel())
    assert np.array_equal(np.array(info2.sto_xyz[:]), np.array(info.sto_xyz[:]))
    from niftyreg_amd import _lib
    with pytest.raises(_lib.NiftyRegError):
        nifti.read_info(tmp_path / "missing.nii")
    (tmp_path / "bad.nii").write_bytes(b"\0" * 400)
    with pytest.raises(_lib.NiftyRegError):
        nifti.read_info(tmp_path / "bad.nii")
    with pytest.raises(_lib.NiftyRegError):
        nifti.write_image(tmp_path / "z.hdr", info, got)


def test_oversized_header_is_rejected(tmp_path):
    """A header whose seven dimensions multiply past 64 bits (or past 2^40 bytes of data) is an
    error before any allocation or read, not a wrapped size."""
    import struct
    from niftyreg_amd import _lib
    data = np.zeros((4, 4, 4), np.float32)
    _write_nifti(tmp_path / "x.nii", data, pixdim=(1.0, 1.0, 1.0), sform=_oblique_sform())
    raw = bytearray((tmp_path / "x.nii").read_bytes())
    raw[40:56] = struct.pack("<8h", 7, 32767, 32767, 32767, 32767, 32767, 32767, 32767)
    raw[70:74] = struct.pack("<hh", 64, 64)  # float64
    (tmp_path / "big.nii").write_bytes(bytes(raw))
    out = np.zeros(8, np.float32)
    with pytest.raises(_lib.NiftyRegError):
        _lib.call("nrh_nifti_read_data", str(tmp_path / "big.nii").encode(), _lib.fptr(out), ctypes.c_size_t(2 ** 62))
    raw[40:56] = struct.pack("<8h", 3, 32767, 32767, 32767, 1, 1, 1, 1)  # 2^45 voxels of 8 bytes: too large
    (tmp_path / "big2.nii").write_bytes(bytes(raw))
    with pytest.raises(_lib.NiftyRegError):
        _lib.call("nrh_nifti_read_data", str(tmp_path / "big2.nii").encode(), _lib.fptr(out), ctypes.c_size_t(2 ** 62))
