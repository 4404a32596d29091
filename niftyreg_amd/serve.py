"""Registration server (nrh_server_*): requests of reg_aladin's affine followed by reg_f3d from it,
admitted by a device-memory budget and run concurrently on one device."""
import ctypes

import numpy as np

from . import _lib
from .geometry import Geom


def request_config(levels=3, maxit=300, bins=64, be=0.001, le=0.0, measure=_lib.MEASURE_NMI, lncc_sigma=-5.0,
                   aladin_levels=3, aladin_maxit=5, spacing=-5.0, field_order=_lib.FIELD_ORDER_SSE):
    """(Aladin config, F3D config) of a request: affine-only non-symmetric Aladin, then F3D."""
    a = _lib.nrh_aladin_config(aladin_levels, 0, aladin_maxit, 50, 50, 0, 1, 1, 0)
    f = _lib.nrh_f3d_config()
    f.bins, f.be_weight, f.le_weight, f.spacing = bins, be, le, spacing
    f.levels, f.levels_to_perform, f.max_iterations = levels, 0, maxit
    f.padding, f.use_pyramid, f.image_storage = float("nan"), 1, 0
    f.measure, f.lncc_sigma = measure, lncc_sigma
    f.affine = None
    f.field_order = field_order
    return a, f


class Server:
    def __init__(self, budget_bytes, max_active=4, device=0):
        s = ctypes.c_void_p()
        _lib.call("nrh_server_create", int(device), ctypes.c_size_t(int(budget_bytes)), int(max_active), ctypes.byref(s))
        self.s = s
        self._keep = {}

    def submit(self, ref_geom: Geom, ref, flo_geom: Geom, flo, config):
        ref = np.ascontiguousarray(ref, np.float32).reshape(-1)
        flo = np.ascontiguousarray(flo, np.float32).reshape(-1)
        r = _lib.nrh_request()
        r.ref, r.flo = ref_geom.to_c(), flo_geom.to_c()
        r.ref_data, r.flo_data = _lib.fptr(ref), _lib.fptr(flo)
        r.aladin, r.f3d = config
        rid = ctypes.c_uint64()
        _lib.call("nrh_server_submit", self.s, ctypes.byref(r), ctypes.byref(rid))
        self._keep[rid.value] = ref_geom
        return rid.value

    def wait(self, rid):
        """The request's result as a dict (status, error, affine, grid, cpp, similarity, iterations,
        submitted, started, finished, device_bytes)."""
        g = self._keep.pop(rid)
        cap = 3 * int(np.prod([n + 8 for n in g.dims]))
        cpp = np.zeros(cap, np.float32)
        res = _lib.nrh_result()
        _lib.call("nrh_server_wait", self.s, ctypes.c_uint64(rid), ctypes.byref(res), _lib.fptr(cpp),
                  ctypes.c_size_t(cap))
        out = {"status": res.status, "error": res.error.decode(), "affine": np.array(res.affine[:]).reshape(4, 4),
               "similarity": res.similarity, "iterations": res.iterations, "submitted": res.submitted,
               "started": res.started, "finished": res.finished, "device_bytes": res.device_bytes}
        if res.status == _lib.NRH_OK:
            grid = Geom.from_c(res.grid)
            out["grid"] = grid
            out["cpp"] = cpp[:3 * grid.nvox].copy()
        return out

    def stats(self):
        st = _lib.nrh_server_stats()
        _lib.call("nrh_server_get_stats", self.s, ctypes.byref(st))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.s:
            _lib.call("nrh_server_destroy", self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
