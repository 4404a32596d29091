"""Aladin (block-matching affine registration) on the GPU through libnrhip's C ABI.

AladinContent mirrors the reference's AladinContent plus its kernels for one pyramid level
(reg-lib/AladinContent.h, reg-lib/cpu/Cpu*Kernel.h): the images, block parameters,
correspondences and the current matrix stay on the device.  Methods are named after the
reference kernels they replace; register() runs reg_aladin<float>::Run.
"""
import ctypes

import numpy as np

from . import _lib
from .geometry import Geom


def default_params(block_percent=50, inlier_lts=50, padding=float("nan")):
    return _lib.nrh_aladin_params(block_percent, inlier_lts, padding)


class AladinContent:
    def __init__(self, ref_geom: Geom, ref, flo_geom: Geom, flo, mask=None, params=None):
        self.ref_geom, self.flo_geom = ref_geom, flo_geom
        self.N = ref_geom.nvox
        self.params = params if params is not None else default_params()
        ref = np.ascontiguousarray(ref, np.float32).reshape(-1)
        flo = np.ascontiguousarray(flo, np.float32).reshape(-1)
        assert ref.size == self.N and flo.size == flo_geom.nvox
        m = None
        if mask is not None:
            m = np.ascontiguousarray(mask, np.int32).reshape(-1)
            assert m.size == self.N
        self._rg, self._fg = ref_geom.to_c(), flo_geom.to_c()
        h = ctypes.c_void_p()
        _lib.call("nrh_aladin_create", ctypes.byref(self._rg), _lib.fptr(ref), ctypes.byref(self._fg),
                  _lib.fptr(flo), _lib.iptr(m) if m is not None else None, ctypes.byref(self.params),
                  ctypes.byref(h))
        self.h = h
        counts = self.block_counts()
        self.total_blocks, self.active_blocks = counts[0], counts[1]

    def close(self):
        if self.h:
            _lib.call("nrh_aladin_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _op(self, name, *args):
        _lib.call(name, self.h, *args)

    def block_counts(self):
        c = np.zeros(3, np.int32)
        self._op("nrh_aladin_blocks", _lib.iptr(c), None)
        return [int(v) for v in c]

    def total_block(self):
        """_reg_blockMatchingParam::totalBlock: rank of each block or -1."""
        c = np.zeros(3, np.int32)
        tb = np.zeros(self.total_blocks, np.int32)
        self._op("nrh_aladin_blocks", _lib.iptr(c), _lib.iptr(tb))
        return tb

    # --- matrix (reg_aladin::affineTransformation)
    def set_matrix(self, m):
        m = np.ascontiguousarray(m, np.float32).reshape(16)
        self._op("nrh_aladin_set_matrix", _lib.fptr(m))

    def get_matrix(self):
        m = np.zeros(16, np.float32)
        self._op("nrh_aladin_get_matrix", _lib.fptr(m))
        return m.reshape(4, 4)

    # --- kernels (AffineDeformationFieldKernel + ResampleImageKernel, BlockMatchingKernel, LtsKernel)
    def GetWarpedImage(self):
        self._op("nrh_aladin_warp")

    def BlockMatching(self):
        self._op("nrh_aladin_block_matching")

    def Lts(self, affine=True):
        self._op("nrh_aladin_lts", int(affine))

    # --- transfers
    def warped(self):
        out = np.empty(self.N, np.float32)
        self._op("nrh_aladin_download_warped", _lib.fptr(out))
        return out

    def set_warped(self, w):
        w = np.ascontiguousarray(w, np.float32).reshape(-1)
        assert w.size == self.N
        self._op("nrh_aladin_upload_warped", _lib.fptr(w))

    def positions(self):
        """referencePosition, warpedPosition (3 floats per active block, rank order) and the
        defined block count."""
        rp = np.empty(3 * self.active_blocks, np.float32)
        wp = np.empty(3 * self.active_blocks, np.float32)
        d = ctypes.c_int()
        self._op("nrh_aladin_download_positions", _lib.fptr(rp), _lib.fptr(wp), ctypes.byref(d))
        return rp, wp, d.value


def lts(ref_pos, warp_pos, inlier_lts=50, affine=True, matrix=None):
    """optimize() on host arrays through the device LTS kernel; returns the updated matrix."""
    rp = np.ascontiguousarray(ref_pos, np.float32).reshape(-1)
    wp = np.ascontiguousarray(warp_pos, np.float32).reshape(-1)
    assert rp.size == wp.size and rp.size % 3 == 0
    m = np.eye(4, dtype=np.float32) if matrix is None else np.array(matrix, np.float32)
    m = np.ascontiguousarray(m.reshape(16))
    _lib.call("nrh_lts", _lib.fptr(rp), _lib.fptr(wp), rp.size // 3, inlier_lts, int(affine), _lib.fptr(m))
    return m.reshape(4, 4)


def symmetric_update(forward: "AladinContent", backward: "AladinContent"):
    """The matrix averaging of reg_aladin_sym::UpdateTransformationMatrix on the two handles'
    device matrices (reg-lib/_reg_aladin_sym.cpp:140-153)."""
    _lib.call("nrh_aladin_symmetric_update", forward.h, backward.h)


def register(ref_geom: Geom, ref, flo_geom: Geom, flo, levels=3, levels_to_perform=0, maxit=5, block_percent=50,
             inlier_lts=50, rigid=True, affine=True, align_centre=True, symmetric=False):
    """reg_aladin<float>::Run, or reg_aladin_sym<float>::Run with symmetric=True (the reg_aladin
    application's default); returns the 4x4 affine."""
    cfg = _lib.nrh_aladin_config(levels, levels_to_perform, maxit, block_percent, inlier_lts, int(rigid), int(affine),
                                 int(align_centre), int(symmetric))
    ref = np.ascontiguousarray(ref, np.float32).reshape(-1)
    flo = np.ascontiguousarray(flo, np.float32).reshape(-1)
    rg, fg = ref_geom.to_c(), flo_geom.to_c()
    m = np.zeros(16, np.float32)
    _lib.call("nrh_aladin_register", ctypes.byref(rg), _lib.fptr(ref), ctypes.byref(fg), _lib.fptr(flo),
              ctypes.byref(cfg), _lib.fptr(m))
    return m.reshape(4, 4)


def register_batch(pairs, levels=3, levels_to_perform=0, maxit=5, block_percent=50, inlier_lts=50, rigid=True,
                   affine=True, align_centre=True):
    """reg_aladin<float>::Run (non-symmetric) for many pairs at once (nrh_aladin_register_batch).

    pairs: list of (ref_geom, ref, flo_geom, flo).  Returns (matrices [n, 4, 4], status [n], stats)."""
    n = len(pairs)
    cfg = _lib.nrh_aladin_config(levels, levels_to_perform, maxit, block_percent, inlier_lts, int(rigid), int(affine),
                                 int(align_centre), 0)
    refs = (_lib.nrh_geom * n)(*[p[0].to_c() for p in pairs])
    flos = (_lib.nrh_geom * n)(*[p[2].to_c() for p in pairs])
    keep = [(np.ascontiguousarray(p[1], np.float32).reshape(-1), np.ascontiguousarray(p[3], np.float32).reshape(-1))
            for p in pairs]
    rdata = (ctypes.POINTER(ctypes.c_float) * n)(*[_lib.fptr(k[0]) for k in keep])
    fdata = (ctypes.POINTER(ctypes.c_float) * n)(*[_lib.fptr(k[1]) for k in keep])
    m = np.zeros(16 * n, np.float32)
    status = (ctypes.c_int * n)()
    stats = _lib.nrh_batch_stats()
    _lib.call("nrh_aladin_register_batch", n, refs, rdata, flos, fdata, ctypes.byref(cfg), _lib.fptr(m), status,
              ctypes.byref(stats))
    st = {"seconds": stats.seconds, "device_seconds": stats.device_seconds, "graph_replays": stats.graph_replays}
    return m.reshape(n, 4, 4), list(status), st
