"""nrh_lts_static, the allocation-free LTS fit, against nrh_lts on the same correspondences.

Two calls in a row (the second with fewer correspondences, so stale scratch from the first would
show), each compared bit for bit with nrh_lts's matrix; the status is 0; sizes beyond the static
scratch are rejected with an argument error.
"""
import ctypes

import numpy as np
import pytest

from niftyreg_amd import _lib

pytestmark = pytest.mark.gpu


def correspondences(n, seed):
    """Reference block positions and their warped positions under a known affine, with noise and
    30% outliers (NaN-free)."""
    rng = np.random.default_rng(seed)
    rp = rng.uniform(0, 128, (n, 3)).astype(np.float32)
    A = np.eye(4)
    A[:3, :3] += rng.uniform(-0.08, 0.08, (3, 3))
    A[:3, 3] = rng.uniform(-5, 5, 3)
    wp = (rp.astype(np.float64) @ A[:3, :3].T + A[:3, 3]).astype(np.float32)
    wp += rng.normal(0, 0.2, wp.shape).astype(np.float32)
    out = rng.uniform(size=n) < 0.3
    wp[out] += rng.uniform(-10, 10, (int(out.sum()), 3)).astype(np.float32)
    return rp.ravel(), wp.ravel()


class Device:
    """Device buffers through the HIP runtime the library itself loaded (no second runtime in the
    process)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def check(self, rc, what):
        assert rc == 0, f"{what} returned {rc}"

    def put(self, a):
        p = ctypes.c_void_p()
        self.check(self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)), "hipMalloc")
        self.ptrs.append(p)
        self.check(self.hip.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1),
                   "hipMemcpy")
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        self.check(self.hip.hipDeviceSynchronize(), "hipDeviceSynchronize")
        self.check(self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), p, ctypes.c_size_t(out.nbytes), 2),
                   "hipMemcpy")
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def test_lts_static_matches_lts(gpu):
    d = Device()
    try:
        for n, seed in ((20000, 1), (5000, 2)):
            rp, wp = correspondences(n, seed)
            drp, dwp = d.put(rp), d.put(wp)
            for affine in (1, 0):
                want = np.eye(4, dtype=np.float32).ravel().copy()
                _lib.call("nrh_lts", _lib.fptr(rp), _lib.fptr(wp), n, 50, affine, _lib.fptr(want))
                eye = np.eye(4, dtype=np.float32).ravel().copy()
                dm = d.put(eye)
                ds = d.put(np.full(1, -7, np.int32))
                _lib.call("nrh_lts_static", drp, dwp, n, 50, affine, dm, ds, None)
                assert int(d.get(ds, np.zeros(1, np.int32))[0]) == 0
                got = d.get(dm, eye)
                print(f"n {n} affine {affine}: static fit equals nrh_lts: {np.array_equal(got, want)}")
                assert np.array_equal(got, want), (n, affine, got, want)
        big = 262145
        with pytest.raises(_lib.NiftyRegError):
            _lib.call("nrh_lts_static", drp, dwp, big, 50, 1, dm, ds, None)
    finally:
        d.free()
