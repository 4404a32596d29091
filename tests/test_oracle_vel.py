"""The velocity-grid (reg_f3d2) restatement of the oracle, pinned three ways.

* Against the reference CPU build without USE_SSE (oracle/_ref/libnrref_nosse.so, whose
  summation order the oracle follows): flow and scaling-and-squaring exponentiation
  (reg_spline_getDefFieldFromVelocityGrid, with and without the step-number update),
  the intermediate fields, reg_defField_compose, reg_resampleGradient, Compute::ExponentiateGradient,
  UpdateVelocityField, SymmetriseVelocityFields and reg_createSymmetricControlPointGrids, bit for bit.
* Against the known answers of the reference's own compose test (reg-test/reg_test_composeField.cpp:
  identity, a 1.11 scaling undone, and the sliding boundary that turns +1 and +3 into +4).
* Against closed forms: a constant velocity exponentiates to a translation, a linear velocity
  A x to (I + A / 2^s)^(2^s) x (trilinear interpolation and cubic B-splines reproduce linear
  maps, so the squaring is exact algebra up to float rounding), a zero backward velocity leaves
  an exponentiated gradient unchanged, and resampling a gradient through the identity field of
  an oblique image returns it (the Jacobian's rotation and spacing cancel).
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd.geometry import Geom, create_control_point_grid

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build absent")


def _oblique(dims, origin=(3.0, -2.0, 5.0)):
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.1, 0.9, 1.3])).astype(np.float32)
    m[:3, 3] = origin
    return Geom(dims, m)


def _case(which, amplitude=3.0, seed=0):
    refg = Geom.identity((24, 22, 20)) if which == "identity" else _oblique((22, 20, 18))
    cg, ident = create_control_point_grid(refg, [4.0 * p for p in refg.pixdim])
    rng = np.random.default_rng(seed)
    vel = (ident + rng.uniform(-amplitude, amplitude, ident.size)).astype(np.float32)
    return refg, cg, ident, vel, rng


def _identity_field(g):
    x, y, z = np.meshgrid(np.arange(g.dims[0]), np.arange(g.dims[1]), np.arange(g.dims[2]), indexing="ij")
    vox = np.stack([x.T.ravel(), y.T.ravel(), z.T.ravel()]).astype(np.float32)
    m = g.xyz.astype(np.float32)
    # float products and sums in the reference's order
    out = np.empty((3, vox.shape[1]), np.float32)
    for r in range(3):
        out[r] = m[r, 0] * vox[0] + m[r, 1] * vox[1] + m[r, 2] * vox[2] + m[r, 3]
    return out.ravel()


@needs_ref
@pytest.mark.parametrize("which", ["identity", "oblique"])
@pytest.mark.parametrize("update", [True, False])
def test_def_field_from_velocity_matches_reference(which, update):
    refg, cg, ident, vel, _ = _case(which)
    a, sa = O.orc_vel_def_field(refg, cg, vel, 6, update)
    b, sb = O.ref_vel_def_field(refg, cg, vel, 6, update)
    assert sa == sb
    assert np.array_equal(a, b)


@needs_ref
def test_step_number_update_matches_reference():
    # large velocities need more than the minimum 6 squarings; a negative intent keeps its sign
    refg, cg, ident, vel, _ = _case("oblique", amplitude=40.0, seed=3)
    a, sa = O.orc_vel_def_field(refg, cg, vel, 6, True)
    b, sb = O.ref_vel_def_field(refg, cg, vel, 6, True)
    assert sa == sb and sa > 6
    assert np.array_equal(a, b)
    a, sa = O.orc_vel_def_field(refg, cg, vel, -6, True)
    b, sb = O.ref_vel_def_field(refg, cg, vel, -6, True)
    assert sa == sb and sa < -6
    assert np.array_equal(a, b)
    a, sa = O.orc_vel_def_field(refg, cg, vel, 0, False)
    b, sb = O.ref_vel_def_field(refg, cg, vel, 0, False)
    assert sa == sb and np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_intermediate_fields_and_compose_match_reference(which):
    refg, cg, ident, vel, rng = _case(which, seed=1)
    ia = O.orc_vel_intermediate(refg, cg, vel, 6)
    ib = O.ref_vel_intermediate(refg, cg, vel, 6)
    assert np.array_equal(ia, ib)
    # the last intermediate field is the exponentiation itself
    assert np.array_equal(ia[-1], O.orc_vel_def_field(refg, cg, vel, 6, False)[0])
    upd = (ia[3] + rng.uniform(-8, 8, ia[3].size)).astype(np.float32)  # positions partly outside
    assert np.array_equal(O.orc_def_compose(refg, ia[1], upd), O.ref_def_compose(refg, ia[1], upd))


@needs_ref
@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_resample_and_exponentiate_gradient_match_reference(which):
    refg, cg, ident, vel, rng = _case(which, seed=2)
    grad = rng.normal(size=3 * refg.nvox).astype(np.float32)
    fields = O.orc_vel_intermediate(refg, cg, vel, 6)
    for k in (0, 3, 6):
        d = fields[k] if k < 6 else (fields[k] + rng.uniform(-6, 6, fields[k].size)).astype(np.float32)
        for pad in (0.0, 1.5):
            assert np.array_equal(O.orc_resample_gradient(refg, grad, d, pad), O.ref_resample_gradient(refg, grad, d, pad))
    vbw = (ident + rng.uniform(-3, 3, ident.size)).astype(np.float32)
    for steps in (6, 7):
        a = O.orc_exponentiate_gradient(refg, cg, vbw, steps, grad)
        b = O.ref_exponentiate_gradient(refg, refg, cg, vel, vbw, steps, grad)
        assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_velocity_update_and_symmetrise_match_reference(which):
    refg, cg, ident, vel, rng = _case(which, seed=4)
    vbw = (ident + rng.uniform(-3, 3, ident.size)).astype(np.float32)
    g = rng.normal(size=ident.size).astype(np.float32)
    for axes in ((1, 1, 1), (1, 0, 1), (0, 1, 0)):
        a = O.orc_update_velocity(vel, g, 0.37, axes)
        b, _ = O.ref_velocity_update(refg, refg, cg, vel, vbw, g, 0.37, axes, mode=0)
        assert np.array_equal(a, b)
    fa, ba = O.orc_symmetrise_velocity(cg, vel, vbw)
    fb, bb = O.ref_velocity_update(refg, refg, cg, vel, vbw, mode=1)
    assert np.array_equal(fa, fb) and np.array_equal(ba, bb)


@needs_ref
def test_symmetric_grids_match_reference():
    for ref, flo, sp in ((Geom.identity((24, 22, 20)), _oblique((20, 26, 16)), (5, 5, 5)),
                         (_oblique((30, 28, 26), (-40.0, 12.0, -3.0)), Geom.identity((18, 20, 22), (1.5, 1.2, 2.0)),
                          (4.0, 6.0, 5.5))):
        g, fw, bw = O.orc_symmetric_grids(ref, flo, sp)
        o, rf, rb = O.ref_symmetric_grids(ref, flo, sp)
        assert g.dims == (o.nx, o.ny, o.nz)
        assert np.array_equal(g.xyz.ravel(), np.array(o.sform[:], np.float32))
        assert np.array_equal(fw, rf) and np.array_equal(bw, rb)


def test_symmetric_grids_host_helper_matches_oracle():
    from niftyreg_amd.geometry import create_symmetric_control_point_grids
    ref, flo = _oblique((30, 28, 26), (-40.0, 12.0, -3.0)), Geom.identity((18, 20, 22), (1.5, 1.2, 2.0))
    g, fw, bw = create_symmetric_control_point_grids(ref, flo, (4.0, 6.0, 5.5))
    og, ofw, obw = O.orc_symmetric_grids(ref, flo, (4.0, 6.0, 5.5))
    assert g.dims == og.dims and np.array_equal(g.xyz, og.xyz) and np.array_equal(g.ijk, og.ijk)
    assert np.array_equal(fw, ofw) and np.array_equal(bw, obw)


# ---------------------------------------------------------------- known answers and closed forms
@pytest.mark.parametrize("case", ["identity", "scaling", "padding"])
def test_compose_known_answers(case):
    """reg-test/reg_test_composeField.cpp on a 5^3 image."""
    g = Geom.identity((5, 5, 5))
    ident = _identity_field(g)
    if case == "identity":
        d, u, exp = ident, ident, ident
    elif case == "scaling":
        d = (ident * np.float32(1.11)).astype(np.float32)
        u = (ident / np.float32(1.11)).astype(np.float32)
        exp = ident
    else:
        d = (ident + 1).astype(np.float32)
        u = (ident + 3).astype(np.float32)
        exp = (ident + 4).astype(np.float32)
    out = O.orc_def_compose(g, d, u)
    np.testing.assert_allclose(out, exp, atol=1e-5)


def test_constant_velocity_is_a_translation():
    refg, cg, ident, _, _ = _case("oblique")
    t = np.array([1.7, -2.3, 0.9], np.float32)
    vel = (ident.reshape(3, -1) + t[:, None]).astype(np.float32).ravel()
    field, steps = O.orc_vel_def_field(refg, cg, vel, 6, True)
    assert steps == 6 + max(0, int(np.ceil(np.log2(2.3 / 0.28))) - 6)
    np.testing.assert_allclose(field.reshape(3, -1), _identity_field(refg).reshape(3, -1) + t[:, None], atol=2e-4)


def test_linear_velocity_squares_exactly():
    refg = Geom.identity((26, 24, 22), origin=(-13.0, -12.0, -11.0))
    cg, ident = create_control_point_grid(refg, [4.0, 4.0, 4.0])
    A = np.array([[0.02, -0.03, 0.01], [0.015, 0.01, -0.02], [-0.01, 0.025, 0.015]])
    P = ident.reshape(3, -1).astype(np.float64)
    vel = (P + A @ P).astype(np.float32).ravel()
    steps = 6
    field, s = O.orc_vel_def_field(refg, cg, vel, steps, False)
    M = np.linalg.matrix_power(np.eye(3) + A / 2 ** steps, 2 ** steps)
    X = _identity_field(refg).reshape(3, -1).astype(np.float64)
    exp = M @ X
    got = field.reshape(3, -1)
    # interior voxels: trajectories stay inside the field, where interpolation is exact for affine maps
    z, y, x = np.meshgrid(*[np.arange(n) for n in refg.shape_zyx], indexing="ij")
    inner = ((x >= 4) & (x < refg.dims[0] - 4) & (y >= 4) & (y < refg.dims[1] - 4) & (z >= 4) &
             (z < refg.dims[2] - 4)).ravel()
    np.testing.assert_allclose(got[:, inner], exp[:, inner], atol=3e-4)
    # the composition order matters: one squaring too few is far off
    M5 = np.linalg.matrix_power(np.eye(3) + A / 2 ** steps, 2 ** (steps - 1))
    assert np.abs((M5 @ X)[:, inner] - got[:, inner]).max() > 0.05


def test_resample_gradient_through_identity_returns_gradient():
    refg = _oblique((14, 12, 10))
    grad = np.random.default_rng(5).normal(size=3 * refg.nvox).astype(np.float32)
    out = O.orc_resample_gradient(refg, grad, _identity_field(refg))
    np.testing.assert_allclose(out, grad, rtol=1e-4, atol=1e-5)


def test_zero_backward_velocity_leaves_gradient_unchanged():
    refg, cg, ident, _, rng = _case("identity", seed=6)
    grad = rng.normal(size=3 * refg.nvox).astype(np.float32)
    out = O.orc_exponentiate_gradient(refg, cg, ident, 6, grad)
    np.testing.assert_allclose(out, grad, rtol=1e-4, atol=1e-5)


def test_symmetrise_makes_displacements_opposite():
    refg, cg, ident, vel, rng = _case("oblique", seed=7)
    vbw = (ident + rng.uniform(-3, 3, ident.size)).astype(np.float32)
    f, b = O.orc_symmetrise_velocity(cg, vel, vbw)
    df = f.astype(np.float64) - ident
    db = b.astype(np.float64) - ident
    np.testing.assert_allclose(df, -db, atol=2e-5)
    np.testing.assert_allclose(df, 0.5 * ((vel - ident) - (vbw - ident)), atol=2e-5)
