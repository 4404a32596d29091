"""Symmetric diffeomorphic F3D (reg_f3d2) evaluations on the device against the reference.

The reference session is reg_f3d2<float> from the reference's own sources built without USE_SSE
(oracle/_ref/libnrref_nosse.so), prepared at one level like reg_base::Run: symmetric velocity
grids, symmetric NMI, bending and linear energy of both grids.  The device runs the same chain
(scaling-and-squaring exponentiation of both grids, both warps, both histograms, voxel
gradients exponentiated along the opposite grid, convolution and node projection, penalty
gradients, the shared normalisation) with each direction on its own stream.  Every kernel on
that chain follows the scalar reference's arithmetic; the objective's terms are FP64 sums whose
order differs from the reference's OpenMP loops (1e-12 relative).  Where the grid's nodes lie on
voxels at an integer spacing, the convolution + node extraction runs as the separable node
projection (the same weights summed in another order): the normalised gradients then agree to
1e-5 relative (L2) and the maximal length to 1e-6; otherwise (an oblique reference, a grid
offset by a fraction of a voxel) the reference-order convolution runs and both directions'
gradients are compared bit for bit.
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import _lib, phantoms
from niftyreg_amd.f3d import F3d2Pair, F3dContent, default_params
from niftyreg_amd.geometry import Geom

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not O.ref_available(), reason="reference build absent")]


def _oblique(dims):
    th = 0.25
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.1, 0.95, 1.2])).astype(np.float32)
    m[:3, 3] = [4.0, -3.0, 2.0]
    return Geom(dims, m)


def _affine():
    c, s = np.cos(0.06), np.sin(0.06)
    m = np.eye(4)
    m[:3, :3] = np.array([[c, -s, 0.02], [s, c, -0.03], [0.01, 0.04, 1.0]]) @ np.diag([1.02, 0.98, 1.01])
    m[:3, 3] = [2.0, -1.5, 1.0]
    return m.astype(np.float32)


def _session(which, be=0.001, le=0.01, seed=1, amplitude=1.5, lncc=False, affine=None):
    if which == "identity":
        rg = fg = Geom.identity((40, 36, 32))
    else:
        rg, fg = _oblique((36, 34, 30)), Geom.identity((34, 38, 28), (1.1, 1.0, 1.2), (-2.0, 1.0, 3.0))
    ref, flo = phantoms.config0_pair(3, rg.dims)
    if fg.dims != rg.dims:
        flo = phantoms.blob_phantom(fg.dims, 4)
    r = O.RefF3d2(rg, ref, fg, flo, be=be, le=le, spacing=-5.0, levels=1, measure=1 if lncc else 0,
                  affine=affine)
    fw0, bw0, _ = r.get_cpp()
    rng = np.random.default_rng(seed)
    fw = (fw0 + rng.uniform(-amplitude, amplitude, fw0.size)).astype(np.float32)
    bw = (bw0 + rng.uniform(-amplitude, amplitude, bw0.size)).astype(np.float32)
    r.set_cpp(fw, bw)
    simw, bew, lew = r.weights()
    cg = O.geom_from_nr(r.cpp_geom)
    m = _lib.MEASURE_LNCC if lncc else _lib.MEASURE_NMI
    pf = default_params(similarity_weight=simw, be_weight=bew, le_weight=lew, measure=m, lncc_sigma=-5.0)
    pb = default_params(similarity_weight=simw, be_weight=bew, le_weight=lew, measure=m, lncc_sigma=-5.0)
    hf = F3dContent(rg, ref.ravel(), fg, flo.ravel(), cg, fw, params=pf)
    hb = F3dContent(fg, flo.ravel(), rg, ref.ravel(), cg, bw, params=pb)
    if affine is not None:
        # the grids' extensions hold the half affines; the contents hold A and its inverse
        from niftyreg_amd.geometry import symmetric_half_affines
        halfF, halfB, ext = symmetric_half_affines(affine)
        hf.set_velocity_affines(halfF if ext else None, halfF if ext else None, affine)
        hb.set_velocity_affines(halfB if ext else None, halfB if ext else None, O.orc_nifti_inverse(affine))
    rng_ = (0, 1) if lncc else (2, 65)
    hf.rescale_intensities(rng_, rng_)
    hb.rescale_intensities(rng_, rng_)
    return r, F3d2Pair(hf, hb, steps=6)


def _check_gradients(pair, rgf, rgb, rmax, m, exact):
    gf, gb = pair.fw.node_gradient(), pair.bw.node_gradient()
    if exact:
        assert m == rmax
        assert np.array_equal(gf, rgf), np.count_nonzero(gf != rgf)
        assert np.array_equal(gb, rgb), np.count_nonzero(gb != rgb)
    else:
        assert abs(m - rmax) <= 1e-6 * rmax
        for g, rg in ((gf, rgf), (gb, rgb)):
            rel = np.linalg.norm(g.astype(np.float64) - rg) / np.linalg.norm(rg)
            assert rel <= 1e-5, rel


@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_symmetric_objective_and_gradient_match_reference(gpu, which):
    r, pair = _session(which)
    rv, rterms = r.value()
    v, terms = pair.objective()
    for k in range(3):
        assert abs(terms[k] - rterms[k]) <= 1e-12 * abs(rterms[k]), (k, terms, rterms)
    assert abs(v - rv) <= 1e-12 * abs(rv)
    rgf, rgb, rmax = r.gradient(normalise=True)
    m = pair.gradient(normalise=True)
    _check_gradients(pair, rgf, rgb, rmax, m, exact=(which == "oblique"))
    # the bending-energy gradient's displacement round trip reaches both device grids too
    f, b, steps = r.get_cpp()
    assert np.array_equal(pair.fw.get_control_points(), f) and np.array_equal(pair.bw.get_control_points(), b)
    assert pair.fw.velocity_steps == steps[0] and pair.bw.velocity_steps == steps[1]
    r.close()
    pair.close()


def test_large_velocities_update_the_step_number(gpu):
    r, pair = _session("identity", amplitude=30.0, seed=5)
    rv, _ = r.value()
    v, _ = pair.objective()
    f, b, steps = r.get_cpp()
    assert steps[0] > 6 and pair.fw.velocity_steps == steps[0] and pair.bw.velocity_steps == steps[1]
    assert abs(v - rv) <= 1e-12 * abs(rv)
    rgf, rgb, rmax = r.gradient(normalise=True)
    m = pair.gradient(normalise=True)
    _check_gradients(pair, rgf, rgb, rmax, m, exact=False)
    r.close()
    pair.close()


@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_symmetric_lncc_matches_reference(gpu, which):
    """reg_f3d2 with LNCC (Gaussian sigma 5 voxels): both directions' LNCC values and voxel
    gradients on the device (the backward one on the floating image's geometry)."""
    r, pair = _session(which, lncc=True)
    rv, rterms = r.value()
    v, terms = pair.objective()
    for k in range(3):
        assert abs(terms[k] - rterms[k]) <= 1e-12 * abs(rterms[k]), (k, terms, rterms)
    rgf, rgb, rmax = r.gradient(normalise=True)
    m = pair.gradient(normalise=True)
    _check_gradients(pair, rgf, rgb, rmax, m, exact=(which == "oblique"))
    r.close()
    pair.close()


@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_symmetric_with_affine_matches_reference(gpu, which):
    """reg_f3d2 with an input affine: half affines in both grids' extensions, A / A^-1 in the
    contents.  The node gradient maps through the extension (the reference-order convolution
    path): objective terms to 1e-12, gradients bit for bit."""
    r, pair = _session(which, affine=_affine())
    rv, rterms = r.value()
    v, terms = pair.objective()
    for k in range(3):
        assert abs(terms[k] - rterms[k]) <= 1e-12 * abs(rterms[k]), (k, terms, rterms)
    rgf, rgb, rmax = r.gradient(normalise=True)
    m = pair.gradient(normalise=True)
    _check_gradients(pair, rgf, rgb, rmax, m, exact=True)
    r.close()
    pair.close()
