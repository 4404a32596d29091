"""ctypes binding of libnrhip.so (include/niftyreg_hip.h).

The library is loaded from this package directory only; a missing library raises
ImportError-style errors on first use so that no caller can silently run without the HIP path.
"""
import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnrhip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "niftyreg_hip.h")

NRH_OK = 0
NRH_ERR_ARG, NRH_ERR_HIP, NRH_ERR_STATE, NRH_ERR_NOMEM, NRH_ERR_UNSUPPORTED, NRH_ERR_NUMERIC = 1, 2, 3, 4, 5, 6
ERR_NAMES = {1: "ARG", 2: "HIP", 3: "STATE", 4: "NOMEM", 5: "UNSUPPORTED", 6: "NUMERIC"}


class NiftyRegError(RuntimeError):
    """Raised for a non-zero status of the C ABI (the reference raises std::runtime_error)."""

    def __init__(self, code, message):
        super().__init__(f"[{ERR_NAMES.get(code, code)}] {message}")
        self.code = code


class nrh_geom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
                ("xyz", ctypes.c_float * 16), ("ijk", ctypes.c_float * 16), ("sform", ctypes.c_int)]


class nrh_f3d_params(ctypes.Structure):
    _fields_ = [("ref_bins", ctypes.c_int), ("flo_bins", ctypes.c_int), ("padding", ctypes.c_float),
                ("similarity_weight", ctypes.c_float), ("be_weight", ctypes.c_float),
                ("le_weight", ctypes.c_float), ("image_storage", ctypes.c_int), ("measure", ctypes.c_int),
                ("lncc_sigma", ctypes.c_float), ("field_order", ctypes.c_int)]


class nrh_f3d_config(ctypes.Structure):
    _fields_ = [("bins", ctypes.c_int), ("be_weight", ctypes.c_float), ("le_weight", ctypes.c_float),
                ("spacing", ctypes.c_float), ("levels", ctypes.c_int), ("levels_to_perform", ctypes.c_int),
                ("max_iterations", ctypes.c_int), ("padding", ctypes.c_float), ("use_pyramid", ctypes.c_int),
                ("image_storage", ctypes.c_int), ("measure", ctypes.c_int), ("lncc_sigma", ctypes.c_float),
                ("affine", ctypes.POINTER(ctypes.c_float)), ("field_order", ctypes.c_int)]

MEASURE_NMI = 0
MEASURE_LNCC = 1
FIELD_ORDER_SCALAR = 0  # the reference built without USE_SSE
FIELD_ORDER_SSE = 1     # the reference's default x86 build (USE_SSE, SSE3)


class nrh_batch_stats(ctypes.Structure):
    _fields_ = [("seconds", ctypes.c_double), ("device_seconds", ctypes.c_double), ("graph_replays", ctypes.c_int),
                ("blocks_per_xcd", ctypes.c_int * 8), ("tile", ctypes.c_int * 3)]


class nrh_server_stats(ctypes.Structure):
    _fields_ = [("pending", ctypes.c_int), ("active", ctypes.c_int), ("completed", ctypes.c_int),
                ("failed", ctypes.c_int), ("peak_active", ctypes.c_int), ("peak_bytes", ctypes.c_size_t),
                ("budget_bytes", ctypes.c_size_t)]


class nrh_aladin_params(ctypes.Structure):
    _fields_ = [("block_percent", ctypes.c_int), ("inlier_lts", ctypes.c_int), ("padding", ctypes.c_float)]


class nrh_aladin_config(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("levels_to_perform", ctypes.c_int), ("max_iterations", ctypes.c_int),
                ("block_percent", ctypes.c_int), ("inlier_lts", ctypes.c_int), ("rigid", ctypes.c_int),
                ("affine", ctypes.c_int), ("align_centre", ctypes.c_int), ("symmetric", ctypes.c_int)]


class nrh_request(ctypes.Structure):
    _fields_ = [("ref", nrh_geom), ("flo", nrh_geom), ("ref_data", ctypes.POINTER(ctypes.c_float)),
                ("flo_data", ctypes.POINTER(ctypes.c_float)), ("aladin", nrh_aladin_config), ("f3d", nrh_f3d_config)]


class nrh_result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("error", ctypes.c_char * 256), ("affine", ctypes.c_float * 16),
                ("grid", nrh_geom), ("similarity", ctypes.c_double), ("iterations", ctypes.c_int),
                ("submitted", ctypes.c_double), ("started", ctypes.c_double), ("finished", ctypes.c_double),
                ("device_bytes", ctypes.c_size_t)]


class nrh_nifti_info(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int * 8), ("pixdim", ctypes.c_float * 8), ("datatype", ctypes.c_int),
                ("scl_slope", ctypes.c_float), ("scl_inter", ctypes.c_float), ("qform_code", ctypes.c_int),
                ("sform_code", ctypes.c_int), ("quatern", ctypes.c_float * 3), ("qoffset", ctypes.c_float * 3),
                ("qfac", ctypes.c_float), ("srow", ctypes.c_float * 12), ("qto_xyz", ctypes.c_float * 16),
                ("sto_xyz", ctypes.c_float * 16), ("intent_code", ctypes.c_int), ("intent_p", ctypes.c_float * 3),
                ("intent_name", ctypes.c_char * 16), ("cal_min", ctypes.c_float), ("cal_max", ctypes.c_float),
                ("xyzt_units", ctypes.c_int), ("descrip", ctypes.c_char * 80)]


_lib = None


def declared_symbols():
    """Function names declared in include/niftyreg_hip.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(nrh_[a-z0-9_]+)\s*\(", text)))


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
        _lib = ctypes.CDLL(LIB_PATH)
        _lib.nrh_last_error.restype = ctypes.c_char_p
    return _lib


def check(rc):
    if rc != NRH_OK:
        msg = load().nrh_last_error().decode(errors="replace")
        raise NiftyRegError(rc, msg)


def call(name, *args):
    fn = getattr(load(), name)
    check(fn(*args))


def fptr(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def dptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def iptr(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
