"""NIfTI-1 images and affine text files at the edges of a registration.

reg_f3d / reg_aladin read their inputs with reg_io_ReadImageFile (reg-io/_reg_ReadWriteImage.cpp:
51-81, niftilib + reg_checkAndCorrectDimension) and write the control-point grid
(reg_createControlPointGrid's header, reg-lib/cpu/_reg_localTrans.cpp:21-128) and the affine text
matrix (reg_tool_WriteAffineFile, reg-io/_reg_ReadWriteMatrix.cpp:105-116).  These functions do
the same through the library's host entry points, so the arrays they produce feed the device
registrations and the files they write are read by the reference's tools.
"""
import ctypes

import numpy as np

from . import _lib
from .geometry import Geom


def read_info(path):
    info = _lib.nrh_nifti_info()
    _lib.call("nrh_nifti_read_info", str(path).encode(), ctypes.byref(info))
    return info


def geometry(info):
    """The Geom the registrations use (sform if sform_code > 0, else qform)."""
    g = _lib.nrh_geom()
    _lib.call("nrh_nifti_geometry", ctypes.byref(info), ctypes.byref(g))
    return Geom((g.nx, g.ny, g.nz), np.array(g.xyz[:], np.float32).reshape(4, 4),
                np.array(g.ijk[:], np.float32).reshape(4, 4), pixdim=(g.dx, g.dy, g.dz), sform=g.sform)


def read_image(path):
    """(Geom, float32 voxels in file order, header info)."""
    info = read_info(path)
    n = int(np.prod([info.dim[i] for i in range(1, 8)]))
    data = np.empty(n, np.float32)
    _lib.call("nrh_nifti_read_data", str(path).encode(), _lib.fptr(data), ctypes.c_size_t(n))
    return geometry(info), data, info


def write_image(path, info, data):
    data = np.ascontiguousarray(data, np.float32).reshape(-1)
    n = int(np.prod([info.dim[i] for i in range(1, info.dim[0] + 1)]))
    if data.size != n:
        raise ValueError(f"{data.size} values for an image of {n} voxels")
    _lib.call("nrh_nifti_write", str(path).encode(), ctypes.byref(info), _lib.fptr(data))


def control_point_grid_info(ref_info, spacing_mm):
    out = _lib.nrh_nifti_info()
    sp = (ctypes.c_float * 3)(*[float(v) for v in spacing_mm])
    _lib.call("nrh_control_point_grid_info", ctypes.byref(ref_info), sp, ctypes.byref(out))
    return out


def write_control_point_grid(path, ref_info, spacing_mm, cpp):
    """The control-point grid file reg_f3d writes (-cpp): planar positions in mm."""
    write_image(path, control_point_grid_info(ref_info, spacing_mm), cpp)


def read_affine(path):
    m = np.zeros(16, np.float32)
    _lib.call("nrh_affine_read", str(path).encode(), _lib.fptr(m))
    return m.reshape(4, 4)


def write_affine(path, m):
    m = np.ascontiguousarray(m, np.float32).reshape(16)
    _lib.call("nrh_affine_write", str(path).encode(), _lib.fptr(m))
