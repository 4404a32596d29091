"""Fused objective / gradient sweeps against the oracle's reference-order chain.

Tolerances (stated for this path): deformation 1e-4 voxel, NMI relative 1e-5, node gradient
relative 1e-3 in the L2 norm and 2e-3 of the largest component element-wise.  The reference's
gradient jumps wherever a warped position crosses a floating-voxel plane (the trilinear image
gradient changes cell there); the sweeps evaluate every voxel within rounding of such a plane
exactly as the reference build does (the exact pass), and so does the histogram sweep with every
voxel whose warped value lies within rounding of a histogram bin edge; the remaining differences
come from float rounding elsewhere (measured: 7e-7 to 4e-6 L2, at most 1.5e-5 element-wise).
At the full size the fused sweeps are compared with the per-operation HIP chain (bit-exact to
the oracle on the smaller cases) and checked for run-to-run determinism.
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import phantoms
from niftyreg_amd.f3d import F3dContent, default_params
from niftyreg_amd.geometry import Geom, create_control_point_grid

pytestmark = pytest.mark.gpu

RB = FB = 68
NMI_RTOL = 1e-5
GRAD_RTOL = 1e-3          # L2-relative
GRAD_MAX_RTOL = 2e-3      # max-abs, relative to the largest component
FULL_GRAD_RTOL = GRAD_RTOL
FIELD_TOL = 1e-4          # voxel


def l2rel(a, b):
    return float(np.linalg.norm((a - b).astype(np.float64)) / np.linalg.norm(b.astype(np.float64)))


def _oblique(dims):
    th = 0.2
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.1, 0.9, 1.2])).astype(np.float32)
    m[:3, 3] = [4.0, -6.0, 2.5]
    return Geom(dims, m)


def build(dims, spacing_vox=5.0, seed=0, amp=2.0, masked=False, be=0.001, flo_dims=None, storage=0, oblique=False):
    rng = np.random.default_rng(seed)
    refg = _oblique(dims) if oblique else Geom.identity(dims)
    flog = Geom.identity(flo_dims) if flo_dims else refg
    ref = phantoms.blob_phantom(refg.dims, seed).ravel()
    flo = phantoms.blob_phantom(flog.dims, seed + 5).ravel()
    cppg, cpp = create_control_point_grid(refg, [spacing_vox * p for p in refg.pixdim])
    cpp = (cpp + rng.uniform(-amp, amp, cpp.size)).astype(np.float32)
    mask = (rng.uniform(size=refg.nvox) > 0.1).astype(np.int32) - 1 if masked else None
    params = default_params(similarity_weight=1.0 - be, be_weight=be, image_storage=storage)
    h = F3dContent(refg, ref, flog, flo, cppg, cpp, mask=mask, params=params)
    h.rescale_intensities((2, RB - 3), (2, FB - 3))
    ref_r = O.orc_rescale(ref, 2, RB - 3)
    flo_r = O.orc_rescale(flo, 2, FB - 3)
    return h, dict(refg=refg, flog=flog, cppg=cppg, ref=ref_r, flo=flo_r, cpp=cpp, mask=mask, be=be)


FUSED_CASES = {
    "tile_aligned": dict(dims=(128, 64, 48)),
    "ragged": dict(dims=(101, 77, 53)),
    "spacing4": dict(dims=(90, 70, 50), spacing_vox=4.0),
    "masked": dict(dims=(80, 72, 40), masked=True),
    "flo_other_size": dict(dims=(70, 66, 45), flo_dims=(75, 60, 49)),
    "large_warp": dict(dims=(96, 80, 64), amp=6.0),
    "oblique": dict(dims=(66, 60, 44), oblique=True),
}


@pytest.fixture(params=sorted(FUSED_CASES))
def fcase(request, gpu):
    h, c = build(**FUSED_CASES[request.param])
    assert h.fused_available(), "geometry expected to take the fused path"
    yield h, c
    h.close()


def test_fused_objective_matches_oracle(fcase):
    h, c = fcase
    value, terms = h.objective()
    field = O.orc_field(c["refg"], c["cppg"], c["cpp"], c["mask"])
    war = O.orc_resample(c["flog"], c["flo"], c["refg"], field, c["mask"])
    nmi, _, _, _ = O.orc_nmi(c["ref"], war, RB, FB, c["mask"])
    be = O.orc_be(c["cppg"], c["cpp"])
    sim_w = 1.0 - c["be"]
    assert abs(terms[0] / np.float32(sim_w) - nmi) <= NMI_RTOL * nmi
    assert abs(terms[1] - np.float32(c["be"]) * be) <= 1e-9 * max(abs(be), 1e-30)
    assert abs(value - (terms[0] - terms[1])) < 1e-12


def test_fused_gradient_matches_oracle(fcase):
    h, c = fcase
    h.objective()
    h.gradient()
    got = h.node_gradient()
    want, _, _ = O.orc_node_gradient_chain(c["refg"], c["flog"], c["cppg"], c["ref"], c["flo"], c["cpp"], RB, FB,
                                           np.float32(1.0 - c["be"]), np.float32(c["be"]), mask=c["mask"])
    scale = np.abs(want).max()
    assert scale > 0
    err = np.abs(got - want).max()
    print(f"fused vs oracle: L2 {l2rel(got, want):.2e}, max-abs {err / scale:.2e}")
    assert l2rel(got, want) <= GRAD_RTOL
    assert err <= GRAD_MAX_RTOL * scale, f"max error {err} vs scale {scale}"


def check_field(field, cppg, refg, cpp, mask=None):
    """Fused field within FIELD_TOL voxel of the FP64 field, and within FIELD_TOL plus the
    reference order's own FP64 error of the reference-order (scalar) field."""
    exact = O.fp64_field(refg, cppg, cpp)
    ref_order = O.orc_field(refg, cppg, cpp, mask)
    if mask is not None:
        act = np.tile(mask > -1, 3)
        field, exact, ref_order = field[act], exact[act], ref_order[act]
    vox = min(refg.pixdim)
    err_exact = np.abs(field - exact).max()
    ref_err = np.abs(ref_order - exact).max()
    assert err_exact <= FIELD_TOL * vox, err_exact
    assert np.abs(field - ref_order).max() <= FIELD_TOL * vox + ref_err
    return err_exact, ref_err


def test_fused_field_matches_oracle(fcase):
    h, c = fcase
    check_field(h.fused_field(), c["cppg"], c["refg"], c["cpp"], c["mask"])


def test_gradient_without_objective_recomputes_histogram(fcase):
    h, c = fcase
    h.objective()
    h.gradient()
    g1 = h.node_gradient()
    # change the control points, then restore: the gradient must see the restored grid
    h.set_control_points(c["cpp"] + np.float32(0.5))
    h.set_control_points(c["cpp"])
    h.gradient()
    assert np.array_equal(h.node_gradient(), g1)


def test_normalised_gradient(fcase):
    h, c = fcase
    h.objective()
    h.gradient()
    g = h.node_gradient()
    G = c["cppg"].nvox
    m_want = float(O.orc().orc_max_length(O.fp(g), G))
    # the bending-energy gradient round-trips the points through displacements in place (as the
    # reference does); start the second evaluation from the same points
    h.set_control_points(c["cpp"])
    m = h.gradient(normalise=True)
    assert m == m_want
    want = g.copy()
    O.orc().orc_normalise_gradient(O.fp(want), G, O.ctypes.c_double(m))
    assert np.array_equal(h.node_gradient(), want)


def test_per_operation_fallback_geometry(gpu):
    # spacing 4.6 voxels: not an integer, the objective runs the per-operation chain
    h, c = build((50, 44, 40), spacing_vox=4.6)
    try:
        assert not h.fused_available()
        value, terms = h.objective()
        field = O.orc_field(c["refg"], c["cppg"], c["cpp"])
        war = O.orc_resample(c["flog"], c["flo"], c["refg"], field)
        nmi, _, _, _ = O.orc_nmi(c["ref"], war, RB, FB)
        assert abs(terms[0] / np.float32(1 - c["be"]) - nmi) <= 1e-12 * nmi
        h.gradient()
        want, _, _ = O.orc_node_gradient_chain(c["refg"], c["flog"], c["cppg"], c["ref"], c["flo"], c["cpp"], RB, FB,
                                               np.float32(1 - c["be"]), np.float32(c["be"]))
        assert np.allclose(h.node_gradient(), want, rtol=1e-5, atol=1e-6 * np.abs(want).max())
    finally:
        h.close()


@pytest.fixture(scope="module")
def full_size(gpu):
    """The 256 x 256 x 160 shell phantom pair at a random 5-voxel B-spline grid."""
    refg = Geom.identity((256, 256, 160))
    ref, flo = phantoms.config1_pair(seed=0)
    cppg, cpp = create_control_point_grid(refg, [5.0, 5.0, 5.0])
    rng = np.random.default_rng(11)
    cpp = (cpp + rng.uniform(-2, 2, cpp.size)).astype(np.float32)
    h = F3dContent(refg, ref, refg, flo, cppg, cpp, params=default_params(similarity_weight=0.999, be_weight=0.001))
    h.rescale_intensities((2, RB - 3), (2, FB - 3))
    yield h, cpp
    h.close()


def test_full_size_fused_vs_per_operation(full_size):
    h, cpp = full_size
    assert h.fused_available()
    value, terms = h.objective()
    h.gradient()
    g_fused = h.node_gradient()
    # per-operation chain on the same handle
    h.GetDeformationField()
    h.ResampleImage()
    nmi, _ = h.GetSimilarityMeasureValue()
    assert abs(terms[0] / np.float32(0.999) - nmi) <= NMI_RTOL * nmi
    h.GetImageGradient()
    h.ZeroVoxelBasedMeasureGradient()
    h.GetVoxelBasedSimilarityMeasureGradient(1.0)
    h.ConvolveImage()
    h.VoxelCentricToNodeCentric(np.float32(0.999))
    h.ApproxBendingEnergyGradient(np.float32(0.001))
    g_ops = h.node_gradient()
    scale = np.abs(g_ops).max()
    print(f"full size fused vs per-operation: L2 {l2rel(g_fused, g_ops):.2e}, "
          f"max-abs {np.abs(g_fused - g_ops).max() / scale:.2e}")
    assert l2rel(g_fused, g_ops) <= FULL_GRAD_RTOL
    assert np.abs(g_fused - g_ops).max() <= GRAD_MAX_RTOL * scale


def test_full_size_fused_field(full_size):
    h, cpp = full_size
    fused = h.fused_field()
    h.GetDeformationField()
    per_op = h.deformation_field()  # reference order, bit-exact to the scalar reference build
    refg = h.ref_geom
    err_exact, ref_err = check_field(fused, h.cpp_geom, refg, cpp)
    assert np.array_equal(per_op, O.orc_field(refg, h.cpp_geom, cpp))
    print(f"fused field error vs FP64 {err_exact:.2e} mm; reference-order error vs FP64 {ref_err:.2e} mm")


def test_full_size_deterministic(full_size):
    h, cpp = full_size
    v1, _ = h.objective()
    h.gradient()
    g1 = h.node_gradient()
    h.set_control_points(cpp)
    v2, _ = h.objective()
    h.gradient()
    assert v1 == v2
    assert np.array_equal(g1, h.node_gradient())

