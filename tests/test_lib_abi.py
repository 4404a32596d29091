"""The C ABI library (CPU side): it loads, exports every declared entry point, and its host
helpers (no device use) agree with the reference."""
import ctypes
import subprocess

import numpy as np
import pytest

import oracles as O
from niftyreg_amd import _lib
from niftyreg_amd.geometry import Geom, create_control_point_grid, mat44_inverse, mat44_inverse_np


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    names = _lib.declared_symbols()
    assert len(names) > 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(names) <= exported


def test_product_library_does_not_link_the_oracle():
    out = subprocess.run(["ldd", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in out and "nrref" not in out
    syms = subprocess.run(["nm", "-D", _lib.LIB_PATH], capture_output=True, text=True).stdout
    assert "orc_" not in syms and "ref_f3d" not in syms


def test_mat44_inverse_matches_nifti_formula():
    rng = np.random.default_rng(0)
    for _ in range(20):
        m = np.eye(4, dtype=np.float32)
        m[:3, :4] = rng.normal(size=(3, 4)).astype(np.float32)
        a = mat44_inverse(m)
        b = mat44_inverse_np(m)
        assert np.array_equal(a, b)
        assert np.allclose(a.astype(np.float64) @ m.astype(np.float64), np.eye(4), atol=1e-4)


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("spacing", [(5.0, 5.0, 5.0), (4.0, 3.0, 6.5), (10.0, 10.0, 10.0)])
def test_control_point_grid_matches_reference(spacing):
    th = 0.2
    R = np.array([[np.cos(th), 0, np.sin(th)], [0, 1, 0], [-np.sin(th), 0, np.cos(th)]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.1, 0.9, 1.3])).astype(np.float32)
    m[:3, 3] = [-20.0, 14.0, 3.5]
    for refg in (Geom.identity((37, 29, 23)), Geom((37, 29, 23), m)):
        sp = [s * d for s, d in zip(spacing, refg.pixdim)]
        cg, data = create_control_point_grid(refg, sp)
        dims, sform, rdata, pd = O.ref_create_cpp(refg, sp)
        assert cg.dims == dims
        assert np.array_equal(cg.xyz, sform)
        assert np.array_equal(data, rdata)
        assert np.allclose(cg.pixdim, pd)


def test_error_codes_without_device():
    lib = _lib.load()
    g = Geom.identity((10, 10, 10)).to_c()
    out = _lib.nrh_geom()
    bad = np.array([-5.0, 5.0, 5.0], np.float32)
    rc = lib.nrh_create_control_point_grid(ctypes.byref(g), _lib.fptr(bad), ctypes.byref(out), None)
    assert rc == 1
    assert b"spacing" in lib.nrh_last_error()
    with pytest.raises(_lib.NiftyRegError):
        _lib.call("nrh_create_control_point_grid", ctypes.byref(g), _lib.fptr(bad), ctypes.byref(out), None)
    rc = lib.nrh_f3d_destroy(None)
    assert rc == 0


def _refine_lib(refg, cg, data):
    lib = _lib.load()
    rc, gc = refg.to_c(), cg.to_c()
    out = _lib.nrh_geom()
    _lib.call("nrh_refine_control_point_grid", ctypes.byref(rc), ctypes.byref(gc), _lib.fptr(data), ctypes.byref(out), None)
    res = np.zeros(3 * out.nx * out.ny * out.nz, np.float32)
    _lib.call("nrh_refine_control_point_grid", ctypes.byref(rc), ctypes.byref(gc), _lib.fptr(data), ctypes.byref(out),
              _lib.fptr(res))
    return Geom.from_c(out), res


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_grid_refinement_matches_reference():
    refg = Geom.identity((47, 41, 38))
    coarse = Geom.identity((24, 21, 19), spacing=(2.0, 2.0, 2.0))
    cg, data = create_control_point_grid(coarse, [20.0, 20.0, 20.0])
    data = (data + np.random.default_rng(0).uniform(-3, 3, data.size)).astype(np.float32)
    ng, res = _refine_lib(refg, cg, data)
    out = O.nr_geom()
    r = O.ref()
    O.rcheck(r.ref_refine_cpp(ctypes.byref(O.rg(refg)), ctypes.byref(O.rg(cg)), O.fp(data), ctypes.byref(out), None))
    want = np.zeros(3 * out.nx * out.ny * out.nz, np.float32)
    O.rcheck(r.ref_refine_cpp(ctypes.byref(O.rg(refg)), ctypes.byref(O.rg(cg)), O.fp(data), ctypes.byref(out), O.fp(want)))
    assert ng.dims == (out.nx, out.ny, out.nz)
    assert np.array_equal(ng.xyz, np.array(out.sform[:], np.float32).reshape(4, 4))
    assert np.allclose(res, want, rtol=1e-6, atol=1e-5)


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_downsample_matches_reference():
    from niftyreg_amd import phantoms
    g = Geom.identity((70, 66, 64), origin=(-3.0, 5.0, 1.0))
    img = phantoms.blob_phantom(g.dims, 4).ravel()
    lib = _lib.load()
    gc = g.to_c()
    out = _lib.nrh_geom()
    axes = np.array([1, 1, 1], np.int32)
    _lib.call("nrh_downsample_image", ctypes.byref(gc), _lib.fptr(img), _lib.iptr(axes), ctypes.byref(out), None)
    res = np.zeros(out.nx * out.ny * out.nz, np.float32)
    _lib.call("nrh_downsample_image", ctypes.byref(gc), _lib.fptr(img), _lib.iptr(axes), ctypes.byref(out), _lib.fptr(res))
    ro = O.nr_geom()
    r = O.ref()
    O.rcheck(r.ref_downsample(ctypes.byref(O.rg(g)), O.fp(img), ctypes.byref(ro), None))
    want = np.zeros(ro.nx * ro.ny * ro.nz, np.float32)
    O.rcheck(r.ref_downsample(ctypes.byref(O.rg(g)), O.fp(img), ctypes.byref(ro), O.fp(want)))
    assert (out.nx, out.ny, out.nz) == (ro.nx, ro.ny, ro.nz)
    assert np.allclose(np.array(out.xyz[:]).reshape(4, 4), np.array(ro.sform[:]).reshape(4, 4))
    # bit-exact with the reference's default (SSE3) build, whose line sums the pyramid follows
    assert np.array_equal(res, want)
