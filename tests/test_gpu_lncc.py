"""LNCC on the device against the reference's golden vectors, the oracle and the reference's
reg_f3d driver.

The device kernels follow the reference's scalar build operation for operation (line sums of
the masked Gaussian convolution in double, per-voxel statistics in float, derivative terms in
double), so the voxel gradient is bit-exact; the value's voxel sum is reduced in a different
(fixed) order than the reference's OpenMP reduction and agrees to 1e-12.
"""
import os

import numpy as np
import pytest

import oracles as O
from niftyreg_amd import _lib, phantoms
from niftyreg_amd.f3d import F3dContent, default_params
from niftyreg_amd.geometry import Geom, create_control_point_grid

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
VALUE_RTOL = 1e-12


def lncc_params(**kw):
    return default_params(measure=_lib.MEASURE_LNCC, lncc_sigma=-5.0, **kw)


def _content(g, ref, flo, cppg, cpp, mask=None, **kw):
    h = F3dContent(g, ref, g, flo, cppg, cpp, mask=mask, params=lncc_params(**kw))
    h.rescale_intensities((0, 1), (0, 1))
    return h


@pytest.mark.parametrize("name", ["blobs", "masked"])
def test_lncc_golden(gpu, name):
    d = np.load(os.path.join(HERE, "golden", f"lncc_{name}.npz"))
    g = Geom.identity(tuple(int(v) for v in d["dims"]))
    cppg, _ = create_control_point_grid(g, [5.0] * 3)
    mask = d["mask"] if d["mask"].size else None
    h = _content(g, d["ref"], d["flo"], cppg, d["cpp"], mask)
    h.GetDeformationField()
    h.ResampleImage()
    if mask is None:
        assert np.array_equal(h.warped(), d["warped"], equal_nan=True)
    v = h.GetLnccValue()
    assert abs(v - d["scalar_value"]) <= VALUE_RTOL * abs(d["scalar_value"])
    h.GetImageGradient()
    h.ZeroVoxelBasedMeasureGradient()
    h.GetVoxelBasedLnccGradient(1.0)
    vg = h.voxel_gradient()
    assert np.array_equal(vg, d["scalar_grad"]), np.abs(vg - d["scalar_grad"]).max()
    h.close()


@pytest.mark.parametrize("dims,sigma,masked", [((61, 47, 39), -5.0, False), ((40, 44, 36), -3.0, True),
                                               ((33, 30, 70), 4.5, False)])
def test_lncc_matches_oracle(gpu, dims, sigma, masked):
    g = Geom.identity(dims, spacing=(1.0, 1.2, 0.9))
    rng = np.random.default_rng(sum(dims))
    ref = phantoms.blob_phantom(dims, 1).ravel()
    flo = phantoms.blob_phantom(dims, 2).ravel()
    cppg, cpp = create_control_point_grid(g, [5.0 * p for p in g.pixdim])
    cpp = (cpp + rng.uniform(-2, 2, cpp.size)).astype(np.float32)
    mask = (rng.uniform(size=g.nvox) > 0.2).astype(np.int32) - 1 if masked else None
    h = F3dContent(g, ref, g, flo, cppg, cpp, mask=mask,
                   params=default_params(measure=_lib.MEASURE_LNCC, lncc_sigma=sigma))
    h.rescale_intensities((0, 1), (0, 1))
    h.GetDeformationField()
    h.ResampleImage()
    h.GetImageGradient()
    war, wg = h.warped(), h.warped_gradient()
    r, _ = h.get_images()
    v = h.GetLnccValue()
    want = O.orc_lncc(g, r, war, mask, sigma)
    assert abs(v - want) <= VALUE_RTOL * abs(want)
    h.ZeroVoxelBasedMeasureGradient()
    h.GetVoxelBasedLnccGradient(0.7)
    got = h.voxel_gradient()
    assert np.array_equal(got, O.orc_lncc_grad(g, r, war, wg, mask, sigma, 0.7))
    h.close()


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_lncc_objective_and_gradient_match_reference_driver(gpu):
    """nrh_f3d_objective / nrh_f3d_gradient with LNCC against reg_f3d<float> (SSE build, one
    level): the value to 1e-6, the normalised node gradient to 1e-3 (L2, relative)."""
    g = Geom.identity((48, 44, 40))
    ref, flo = phantoms.config0_pair(3, g.dims)
    r = O.RefF3d(g, ref, g, flo, be=0.001, le=0.0, spacing=-5.0, levels=1, measure=1, lncc_sigma=-5.0)
    cpp0 = r.get_cpp()
    cpp = (cpp0 + np.random.default_rng(1).uniform(-1.5, 1.5, cpp0.size)).astype(np.float32)
    r.set_cpp(cpp)
    rv, rterms = r.value()
    rgrad, rmax = r.gradient(normalise=True)
    simw, bew = r.weights()
    cg = O.geom_from_nr(r.cpp_geom)
    r.close()
    h = F3dContent(g, ref.ravel(), g, flo.ravel(), cg, cpp,
                   params=default_params(similarity_weight=simw, be_weight=bew, measure=_lib.MEASURE_LNCC))
    h.rescale_intensities((0, 1), (0, 1))
    v, terms = h.objective()
    assert abs(v - rv) <= 1e-6 * abs(rv)
    maxlen = h.gradient(normalise=True)
    grad = h.node_gradient()
    assert abs(maxlen - rmax) <= 1e-4 * rmax
    rel = np.linalg.norm(grad - rgrad) / np.linalg.norm(rgrad)
    assert rel <= 1e-3, rel
    h.close()
