"""Velocity grids with affine parts (reg_f3d2 with an input affine): the oracle against the
reference CPU build (scalar, USE_SSE off) bit for bit, and against closed forms.

reg_f3d2 splits the input affine A into half affines (expm(logm(A) / 2) and the same of A^-1),
carries both images half-way when it builds the symmetric grids, and stores each half in its
grid's NIfTI extensions 0 and 1: extension 0 is applied before the spline and its affine-only
field is removed before the scaling and squaring and put back after it, extension 1 is composed
last; the node gradient maps the grid through extension 0's inverse; the gradient exponentiation
removes the displacement of the opposite content's transformation matrix (A or A^-1) from each
intermediate field.
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd.geometry import Geom, create_control_point_grid, create_symmetric_control_point_grids, \
    symmetric_half_affines

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build absent")


def _affine(angle=0.08, shift=(2.5, -1.5, 1.0), scale=(1.03, 0.97, 1.01)):
    c, s = np.cos(angle), np.sin(angle)
    R = np.array([[c, -s, 0.02], [s, c, -0.03], [0.01, 0.04, 1.0]])
    m = np.eye(4)
    m[:3, :3] = R @ np.diag(scale)
    m[:3, 3] = shift
    return m.astype(np.float32)


def _oblique(dims, origin=(3.0, -2.0, 5.0)):
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.1, 0.9, 1.3])).astype(np.float32)
    m[:3, 3] = origin
    return Geom(dims, m)


def _case(which, amplitude=2.0, seed=0):
    refg = Geom.identity((24, 22, 20)) if which == "identity" else _oblique((22, 20, 18))
    # a grid with margin: the half affine moves the identity by a few millimetres
    cg, ident = create_control_point_grid(refg, [3.0 * p for p in refg.pixdim])
    rng = np.random.default_rng(seed)
    vel = (ident + rng.uniform(-amplitude, amplitude, ident.size)).astype(np.float32)
    hf, hb, _ = O.orc_half_affines(_affine())
    return refg, cg, ident, vel, rng, hf, hb


@needs_ref
def test_half_affines_and_grids_match_reference():
    A = _affine()
    hf, hb, ext = O.orc_half_affines(A)
    for ref, flo, sp in ((Geom.identity((24, 22, 20)), _oblique((20, 26, 16)), (5, 5, 5)),
                         (_oblique((30, 28, 26), (-40.0, 12.0, -3.0)), Geom.identity((18, 20, 22), (1.5, 1.2, 2.0)),
                          (4.0, 6.0, 5.5))):
        o, rf, rb, ef, eb, cnt = O.ref_symmetric_grids_aff(ref, flo, sp, A)
        assert ext and cnt == 2
        assert np.array_equal(hf, ef) and np.array_equal(hb, eb)
        g, fw, bw = O.orc_symmetric_grids_aff(ref, flo, sp, A)
        assert g.dims == (o.nx, o.ny, o.nz)
        assert np.array_equal(g.xyz.ravel(), np.array(o.sform[:], np.float32))
        assert np.array_equal(fw, rf) and np.array_equal(bw, rb)
        # the library's host helpers
        hg, hfw, hbw = create_symmetric_control_point_grids(ref, flo, sp, affine=A)
        assert hg.dims == g.dims and np.array_equal(hg.xyz, g.xyz)
        assert np.array_equal(hfw, fw) and np.array_equal(hbw, bw)
    lf, lb, lext = symmetric_half_affines(A)
    assert lext and np.array_equal(lf, hf) and np.array_equal(lb, hb)


def test_half_affines_closed_forms():
    """half * half = A, the backward half inverts the forward one, the identity has no halves."""
    A = _affine(0.2, (10.0, -4.0, 6.0))
    hf, hb, ext = O.orc_half_affines(A)
    assert ext
    np.testing.assert_allclose(hf.astype(np.float64) @ hf, A, atol=2e-5)
    np.testing.assert_allclose(hb.astype(np.float64) @ hb, np.linalg.inv(A.astype(np.float64)), atol=2e-5)
    np.testing.assert_allclose(hf.astype(np.float64) @ hb, np.eye(4), atol=2e-6)
    # a pure translation halves exactly in its translation column
    T = np.eye(4, dtype=np.float32)
    T[:3, 3] = [4.0, -2.0, 8.0]
    tf, tb, _ = O.orc_half_affines(T)
    np.testing.assert_allclose(tf[:3, 3], [2.0, -1.0, 4.0], atol=1e-6)
    np.testing.assert_allclose(tb[:3, 3], [-2.0, 1.0, -4.0], atol=1e-6)
    _, _, e = O.orc_half_affines(np.eye(4, dtype=np.float32))
    assert not e


@needs_ref
@pytest.mark.parametrize("which", ["identity", "oblique"])
@pytest.mark.parametrize("update", [True, False])
def test_def_field_with_extensions_matches_reference(which, update):
    refg, cg, ident, vel, _, hf, hb = _case(which)
    # (an extension 1 needs an extension 0 before it in a NIfTI extension list)
    for e0, e1 in ((hf, hf), (hf, None), (hf, hb)):
        a, sa = O.orc_vel_def_field_aff(refg, cg, vel, e0, e1, 6, update)
        b, sb = O.ref_vel_def_field_aff(refg, cg, vel, e0, e1, 6, update)
        assert sa == sb
        assert np.array_equal(a, b), np.count_nonzero(a != b)


@needs_ref
@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_intermediate_fields_and_exponentiation_with_affines_match_reference(which):
    refg, cg, ident, vel, rng, hf, hb = _case(which, seed=2)
    a = O.orc_vel_intermediate_aff(refg, cg, vel, hb, hb, 6)
    b = O.ref_vel_intermediate_aff(refg, cg, vel, hb, hb, 6)
    assert np.array_equal(a, b)
    grad = rng.normal(size=3 * refg.nvox).astype(np.float32)
    vfw = (ident + rng.uniform(-2, 2, ident.size)).astype(np.float32)
    Ainv = np.linalg.inv(_affine().astype(np.float64)).astype(np.float32)
    for e0, e1, t in ((hb, hb, Ainv), (None, None, Ainv), (hb, hb, None)):
        x = O.orc_exponentiate_gradient_aff(refg, cg, vel, e0, e1, t, 6, grad)
        y = O.ref_exponentiate_gradient_aff(refg, refg, cg, vfw, vel, e0, e1, t, 6, grad)
        assert np.array_equal(x, y), np.count_nonzero(x != y)


@needs_ref
@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_voxel_to_node_with_extension_matches_reference(which):
    refg, cg, ident, vel, rng, hf, hb = _case(which, seed=3)
    flog = _oblique((20, 18, 16), (1.0, 2.0, -1.0))
    vg = rng.normal(size=3 * refg.nvox).astype(np.float32)
    a = O.orc_voxel_to_node_ext(cg, refg, vg, flog, hf, 0.7)
    b = O.ref_voxel_to_node_ext(cg, refg, vg, flog, hf, 0.7)
    assert np.array_equal(a, b)


def test_identity_velocity_with_extensions_is_the_composed_affine():
    """A velocity grid at its identity positions carries no displacement: the field is B A x."""
    refg = Geom.identity((20, 18, 16))
    cg, ident = create_control_point_grid(refg, [3.0, 3.0, 3.0])
    A = _affine(0.03, (0.8, -0.5, 0.4))
    B = _affine(-0.02, (-0.3, 0.6, 0.2), (0.99, 1.01, 1.0))
    field, _ = O.orc_vel_def_field_aff(refg, cg, ident, A, B, 6, False)
    x, y, z = np.meshgrid(np.arange(20), np.arange(18), np.arange(16), indexing="ij")
    vox = np.stack([x.T.ravel(), y.T.ravel(), z.T.ravel(), np.ones(x.size)]).astype(np.float64)
    want = (B.astype(np.float64) @ A.astype(np.float64) @ refg.xyz.astype(np.float64) @ vox)[:3]
    # interior voxels (the grid reproduces A x exactly away from its sliding border)
    inner = (vox[0] > 2) & (vox[0] < 17) & (vox[1] > 2) & (vox[1] < 15) & (vox[2] > 2) & (vox[2] < 13)
    got = field.reshape(3, -1)
    np.testing.assert_allclose(got[:, inner], want[:, inner], atol=5e-4)
    # swapping the extensions composes in the other order: B A differs from A B
    swapped, _ = O.orc_vel_def_field_aff(refg, cg, ident, B, A, 6, False)
    assert np.abs(swapped.reshape(3, -1)[:, inner] - got[:, inner]).max() > 1e-3
