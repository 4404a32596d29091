"""Per-operation HIP kernels against the oracle (bit-exact where the operation is a pure function
of its inputs with the reference's rounding; FP64 reductions within 1e-12 relative)."""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import phantoms
from niftyreg_amd.f3d import F3dContent, default_params
from niftyreg_amd.geometry import Geom, create_control_point_grid

pytestmark = pytest.mark.gpu

RB = FB = 68


def oblique_geom(dims, spacing=(1.2, 0.9, 1.5), theta=0.3, origin=(3.0, -7.0, 11.0)):
    c, s = np.cos(theta), np.sin(theta)
    R = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag(spacing)).astype(np.float32)
    m[:3, 3] = origin
    return Geom(dims, m)


def make_case(dims=(45, 40, 35), spacing_vox=5.0, seed=0, geom=None, masked=False, flo_dims=None):
    rng = np.random.default_rng(seed)
    refg = geom if geom is not None else Geom.identity(dims)
    flog = Geom.identity(flo_dims) if flo_dims else refg
    ref = phantoms.blob_phantom(refg.dims, seed).ravel()
    flo = phantoms.blob_phantom(flog.dims, seed + 7).ravel()
    sp = [spacing_vox * d for d in refg.pixdim]
    cppg, cpp = create_control_point_grid(refg, sp)
    cpp = (cpp + rng.uniform(-2, 2, cpp.size)).astype(np.float32)
    mask = None
    if masked:
        mask = (rng.uniform(size=refg.nvox) > 0.15).astype(np.int32) - 1
    return refg, flog, cppg, ref, flo, cpp, mask


CASES = {
    "lut5": dict(),
    "spacing4": dict(spacing_vox=4.0),
    "spacing_frac": dict(spacing_vox=4.6),
    "masked": dict(masked=True),
    "oblique": dict(geom=oblique_geom((30, 28, 26))),
    "flo_other_size": dict(flo_dims=(41, 37, 39)),
}


@pytest.fixture(params=sorted(CASES))
def case(request, gpu):
    refg, flog, cppg, ref, flo, cpp, mask = make_case(**CASES[request.param])
    ref_r = O.orc_rescale(ref, 2, RB - 3)
    flo_r = O.orc_rescale(flo, 2, FB - 3)
    h = F3dContent(refg, ref, flog, flo, cppg, cpp, mask=mask, params=default_params())
    h.rescale_intensities((2, RB - 3), (2, FB - 3))
    yield dict(name=request.param, refg=refg, flog=flog, cppg=cppg, ref=ref_r, flo=flo_r, cpp=cpp, mask=mask, h=h)
    h.close()


def test_rescale_exact(case):
    r, f = case["h"].get_images()
    assert np.array_equal(r, case["ref"])
    assert np.array_equal(f, case["flo"])


def test_field_exact(case):
    h = case["h"]
    h.GetDeformationField()
    got = h.deformation_field()
    want = O.orc_field(case["refg"], case["cppg"], case["cpp"], case["mask"])
    if case["mask"] is not None:
        act = np.tile(case["mask"] > -1, 3)
        got, want = got[act], want[act]
    assert np.array_equal(got, want)


def test_resample_and_gradient_exact(case):
    h = case["h"]
    h.GetDeformationField()
    field = h.deformation_field()
    h.ResampleImage()
    h.GetImageGradient()
    war = O.orc_resample(case["flog"], case["flo"], case["refg"], field, case["mask"])
    wg = O.orc_image_gradient(case["flog"], case["flo"], case["refg"], field, case["mask"])
    got = h.warped()
    if case["mask"] is not None:
        act = case["mask"] > -1
        assert np.array_equal(got[act], war[act], equal_nan=True)
    else:
        assert np.array_equal(got, war, equal_nan=True)
    assert np.isnan(war).any() or case["name"] == "masked"  # the warp reaches outside the floating image
    assert np.array_equal(h.warped_gradient(), wg)


def test_nmi_value(case):
    h = case["h"]
    h.GetDeformationField()
    h.ResampleImage()
    nmi, ent = h.GetSimilarityMeasureValue()
    war = h.warped()
    v, pro, logh, ent_o = O.orc_nmi(case["ref"], war, RB, FB, case["mask"])
    assert abs(nmi - v) <= 1e-12 * abs(v)
    assert np.allclose(ent, ent_o, rtol=1e-12, atol=0)
    p, l = h.histogram()
    assert np.allclose(p, pro, rtol=1e-12, atol=1e-18)
    assert np.allclose(l, logh, rtol=1e-12, atol=1e-15)


def test_nmi_voxel_gradient(case):
    h = case["h"]
    h.GetDeformationField()
    h.ResampleImage()
    h.GetImageGradient()
    h.GetSimilarityMeasureValue()
    h.ZeroVoxelBasedMeasureGradient()
    h.GetVoxelBasedSimilarityMeasureGradient(1.0)
    war, wg = h.warped(), h.warped_gradient()
    _, _, logh, ent = O.orc_nmi(case["ref"], war, RB, FB, case["mask"])
    want = O.orc_nmi_grad(case["ref"], war, wg, RB, FB, logh, ent, 1.0, case["mask"])
    got = h.voxel_gradient()
    scale = np.abs(want).max()
    assert scale > 0
    assert np.abs(got - want).max() <= 1e-6 * scale


def test_convolution_and_node_projection(case):
    h = case["h"]
    h.GetDeformationField()
    h.ResampleImage()
    h.GetImageGradient()
    h.GetSimilarityMeasureValue()
    h.ZeroVoxelBasedMeasureGradient()
    h.GetVoxelBasedSimilarityMeasureGradient(1.0)
    vg = h.voxel_gradient()
    h.ConvolveImage()
    conv = h.voxel_gradient()
    want = O.orc_convolve_voxel_gradient(case["refg"], case["cppg"], vg)
    assert np.array_equal(conv, want)
    h.VoxelCentricToNodeCentric(0.75)
    ng = h.node_gradient()
    want_n = O.orc_voxel_to_node(case["cppg"], case["refg"], conv, case["flog"], 0.75)
    assert np.array_equal(ng, want_n)


def test_bending_energy(case):
    h = case["h"]
    be = h.ApproxBendingEnergy()
    want = O.orc_be(case["cppg"], case["cpp"])
    assert abs(be - want) <= 1e-12 * abs(want)
    h.ZeroTransformationGradient()
    h.ApproxBendingEnergyGradient(0.01)
    assert np.array_equal(h.node_gradient(), O.orc_be_grad(case["cppg"], case["cpp"], 0.01))


def test_optimiser_helpers(case):
    h = case["h"]
    rng = np.random.default_rng(3)
    G = case["cppg"].nvox
    g1 = rng.normal(size=3 * G).astype(np.float32)
    h.set_node_gradient(g1)
    m = h.GetMaximalLength()
    assert m == float(O.orc().orc_max_length(O.fp(g1), G))
    h.NormaliseGradient(m)
    want = g1.copy()
    O.orc().orc_normalise_gradient(O.fp(want), G, O.ctypes.c_double(m))
    assert np.array_equal(h.node_gradient(), want)
    # conjugate gradient: first call then one update
    gg, hh = np.zeros_like(g1), np.zeros_like(g1)
    grad = want.copy()
    O.orc().orc_cg_update(O.fp(grad), O.fp(gg), O.fp(hh), 3 * G, 1)
    h.conjugate_gradient(True)
    assert np.array_equal(h.node_gradient(), grad)
    g2 = rng.normal(size=3 * G).astype(np.float32)
    h.set_node_gradient(g2)
    grad = g2.copy()
    O.orc().orc_cg_update(O.fp(grad), O.fp(gg), O.fp(hh), 3 * G, 0)
    h.conjugate_gradient(False)
    assert np.allclose(h.node_gradient(), grad, rtol=1e-6, atol=1e-7)
    # DOF update cur = best + scale * grad
    h.store_best()
    h.update_control_points(-0.37)
    best = case["cpp"]
    cur = np.empty_like(best)
    O.orc().orc_update_dof(O.fp(cur), O.fp(best), O.fp(h.node_gradient()), O.ctypes.c_float(-0.37), 3 * G)
    assert np.array_equal(h.get_control_points(), cur)
    h.restore_best()
    assert np.array_equal(h.get_control_points(), best)


def test_bending_energy_gradient_round_trips_the_points(gpu):
    """reg_spline_approxBendingEnergyGradient3D converts the grid to displacements and back in
    place, so points whose node coordinate differs in sign or magnitude from the point lose their
    last bits: the device applies the same (p - node) + node to the handle's points."""
    from niftyreg_amd.geometry import Geom, create_control_point_grid
    th = 0.2
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
                 @ np.diag([1.1, 0.9, 1.2])).astype(np.float32)
    m[:3, 3] = [4.0, -6.0, 2.5]
    g = Geom((34, 30, 28), m)
    cg, cpp = create_control_point_grid(g, [5.0 * p for p in g.pixdim])
    cpp = (cpp + np.random.default_rng(0).uniform(-2, 2, cpp.size)).astype(np.float32)
    img = np.random.default_rng(1).uniform(size=g.nvox).astype(np.float32)
    h = F3dContent(g, img, g, img, cg, cpp)
    h.ZeroTransformationGradient()
    h.ApproxBendingEnergyGradient(0.01)
    want = O.orc_round_trip(cg, cpp)
    assert np.count_nonzero(want != cpp) > 0  # the case exercises the rounding
    assert np.array_equal(h.get_control_points(), want)
    h.close()
