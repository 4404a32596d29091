"""Image geometry (the part of a nifti header the hot path reads).

A Geom carries the dims, pixdim and the voxel->mm / mm->voxel matrix pair that the reference
selects from a header (sform when sform_code > 0, else qform; reg-lib/Content.h:31-36).
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from . import _lib


def mat44_inverse(m):
    """nifti_mat44_inverse (reg-io/niftilib/nifti1_io.c:1680-1717) via the library's host helper."""
    m = np.ascontiguousarray(m, dtype=np.float32).reshape(16)
    out = np.zeros(16, np.float32)
    _lib.call("nrh_mat44_inverse", _lib.fptr(m), _lib.fptr(out))
    return out.reshape(4, 4)


def mat44_inverse_np(m):
    """The same adjugate inverse in numpy float64, stored in float32 (usable without the library)."""
    R = np.asarray(m, np.float64).reshape(4, 4)
    r11, r12, r13, v1 = R[0]
    r21, r22, r23, v2 = R[1]
    r31, r32, r33, v3 = R[2]
    deti = r11*r22*r33 - r11*r32*r23 - r21*r12*r33 + r21*r32*r13 + r31*r12*r23 - r31*r22*r13
    if deti != 0.0:
        deti = 1.0 / deti
    Q = np.zeros((4, 4), np.float32)
    Q[0, 0] = deti*(r22*r33 - r32*r23)
    Q[0, 1] = deti*(-r12*r33 + r32*r13)
    Q[0, 2] = deti*(r12*r23 - r22*r13)
    Q[0, 3] = deti*(-r12*r23*v3 + r12*v2*r33 + r22*r13*v3 - r22*v1*r33 - r32*r13*v2 + r32*v1*r23)
    Q[1, 0] = deti*(-r21*r33 + r31*r23)
    Q[1, 1] = deti*(r11*r33 - r31*r13)
    Q[1, 2] = deti*(-r11*r23 + r21*r13)
    Q[1, 3] = deti*(r11*r23*v3 - r11*v2*r33 - r21*r13*v3 + r21*v1*r33 + r31*r13*v2 - r31*v1*r23)
    Q[2, 0] = deti*(r21*r32 - r31*r22)
    Q[2, 1] = deti*(-r11*r32 + r31*r12)
    Q[2, 2] = deti*(r11*r22 - r21*r12)
    Q[2, 3] = deti*(-r11*r22*v3 + r11*r32*v2 + r21*r12*v3 - r21*r32*v1 - r31*r12*v2 + r31*r22*v1)
    Q[3, 3] = 0.0 if deti == 0.0 else 1.0
    return Q


@dataclass
class Geom:
    dims: tuple
    xyz: np.ndarray                  # 4x4 float32 voxel -> mm
    ijk: np.ndarray = None           # 4x4 float32 mm -> voxel
    pixdim: tuple = None
    sform: int = 1

    def __post_init__(self):
        self.dims = tuple(int(d) for d in self.dims)
        self.xyz = np.asarray(self.xyz, np.float32).reshape(4, 4)
        if self.ijk is None:
            self.ijk = mat44_inverse_np(self.xyz)
        self.ijk = np.asarray(self.ijk, np.float32).reshape(4, 4)
        if self.pixdim is None:
            # pixdim as nifti_mat44_to_quatern derives it: column norms, in double
            c = np.asarray(self.xyz, np.float64)[:3, :3]
            self.pixdim = tuple(float(np.float32(np.sqrt((c[:, i] ** 2).sum()))) for i in range(3))

    @property
    def nvox(self):
        return int(np.prod(self.dims))

    @property
    def shape_zyx(self):
        return (self.dims[2], self.dims[1], self.dims[0])

    def to_c(self):
        g = _lib.nrh_geom()
        g.nx, g.ny, g.nz = self.dims
        g.dx, g.dy, g.dz = self.pixdim
        g.xyz[:] = [float(v) for v in self.xyz.reshape(16)]
        g.ijk[:] = [float(v) for v in self.ijk.reshape(16)]
        g.sform = int(self.sform)
        return g

    @staticmethod
    def from_c(g):
        return Geom((g.nx, g.ny, g.nz), np.array(g.xyz[:], np.float32).reshape(4, 4),
                    np.array(g.ijk[:], np.float32).reshape(4, 4), (g.dx, g.dy, g.dz), g.sform)

    @staticmethod
    def identity(dims, spacing=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0)):
        m = np.eye(4, dtype=np.float32)
        m[0, 0], m[1, 1], m[2, 2] = spacing
        m[:3, 3] = origin
        return Geom(dims, m)


def create_control_point_grid(ref: Geom, spacing_mm):
    """reg_createControlPointGrid<float> (reg-lib/cpu/_reg_localTrans.cpp:21-128).

    Returns (grid Geom, identity control-point positions, planar float32 [3*G])."""
    sp = np.asarray(spacing_mm, np.float32)
    g = _lib.nrh_geom()
    rc = ref.to_c()
    _lib.call("nrh_create_control_point_grid", ctypes.byref(rc), _lib.fptr(sp), ctypes.byref(g), None)
    geom = Geom.from_c(g)
    data = np.zeros(3 * geom.nvox, np.float32)
    _lib.call("nrh_create_control_point_grid", ctypes.byref(rc), _lib.fptr(sp), ctypes.byref(g), _lib.fptr(data))
    return geom, data


def create_symmetric_control_point_grids(ref: Geom, flo: Geom, spacing_mm, affine=None):
    """reg_createSymmetricControlPointGrids<float> (reg-lib/cpu/_reg_localTrans.cpp:132-358).

    affine: optional reference -> floating 4x4.  Returns (grid Geom, forward identity positions,
    backward identity positions); the half affines come from symmetric_half_affines."""
    sp = np.asarray(spacing_mm, np.float32)
    a = None if affine is None else np.ascontiguousarray(affine, np.float32).reshape(16)
    ap = None if a is None else _lib.fptr(a)
    g = _lib.nrh_geom()
    rc, fc = ref.to_c(), flo.to_c()
    _lib.call("nrh_create_symmetric_control_point_grids", ctypes.byref(rc), ctypes.byref(fc), _lib.fptr(sp), ap,
              ctypes.byref(g), None, None)
    geom = Geom.from_c(g)
    fw = np.zeros(3 * geom.nvox, np.float32)
    bw = np.zeros(3 * geom.nvox, np.float32)
    _lib.call("nrh_create_symmetric_control_point_grids", ctypes.byref(rc), ctypes.byref(fc), _lib.fptr(sp), ap,
              ctypes.byref(g), _lib.fptr(fw), _lib.fptr(bw))
    return geom, fw, bw


def symmetric_half_affines(affine):
    """The half affines of reg_createSymmetricControlPointGrids (reg-lib/cpu/_reg_localTrans.cpp:150-166).

    Returns (half forward 4x4, half backward 4x4, whether the grids carry them as extensions)."""
    a = np.ascontiguousarray(affine, np.float32).reshape(16)
    hf = np.zeros(16, np.float32)
    hb = np.zeros(16, np.float32)
    ext = ctypes.c_int()
    _lib.call("nrh_symmetric_half_affines", _lib.fptr(a), _lib.fptr(hf), _lib.fptr(hb), ctypes.byref(ext))
    return hf.reshape(4, 4), hb.reshape(4, 4), bool(ext.value)
