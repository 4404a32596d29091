"""The velocity-grid (reg_f3d2) operations on the device against the oracle, bit for bit.

The oracle's restatement equals the reference CPU build (USE_SSE off) bit for bit
(tests/test_oracle_vel.py); the device kernels follow the same float / double arithmetic, so
every comparison here is exact: the scaling-and-squaring exponentiation with and without the
step-number update, the gradient exponentiation along the backward grid's intermediate fields,
the velocity update, the symmetrisation, the field composition and the gradient resampling.
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import _lib, phantoms
from niftyreg_amd.f3d import F3dContent, default_params
from niftyreg_amd.geometry import Geom, create_control_point_grid, create_symmetric_control_point_grids

pytestmark = pytest.mark.gpu


def _oblique(dims, origin=(3.0, -2.0, 5.0)):
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.1, 0.9, 1.3])).astype(np.float32)
    m[:3, 3] = origin
    return Geom(dims, m)


def _pair(which, amplitude=3.0, seed=0):
    """Forward and backward handles of a symmetric registration with random velocity grids."""
    if which == "identity":
        ref_g, flo_g = Geom.identity((40, 36, 32)), Geom.identity((38, 40, 30), (1.1, 0.9, 1.2))
    elif which == "oblique":
        ref_g, flo_g = _oblique((34, 30, 28)), Geom.identity((32, 34, 26), (1.2, 1.0, 1.3), (-4.0, 2.0, 1.0))
    else:  # a registration-sized case
        ref_g, flo_g = Geom.identity((160, 160, 128)), Geom.identity((160, 160, 128))
    cg, fw_id, bw_id = create_symmetric_control_point_grids(ref_g, flo_g, [5.0, 5.0, 5.0])
    rng = np.random.default_rng(seed)
    vfw = (fw_id + rng.uniform(-amplitude, amplitude, fw_id.size)).astype(np.float32)
    vbw = (bw_id + rng.uniform(-amplitude, amplitude, bw_id.size)).astype(np.float32)
    ref = phantoms.blob_phantom(ref_g.dims, 1).ravel()
    flo = phantoms.blob_phantom(flo_g.dims, 2).ravel()
    fw = F3dContent(ref_g, ref, flo_g, flo, cg, vfw, params=default_params())
    bw = F3dContent(flo_g, flo, ref_g, ref, cg, vbw, params=default_params())
    fw.SetVelocityGrid(6)
    bw.SetVelocityGrid(6)
    return ref_g, flo_g, cg, fw, bw, vfw, vbw, rng


@pytest.mark.parametrize("which", ["identity", "oblique", "large"])
def test_def_field_from_velocity_matches_oracle(gpu, which):
    ref_g, flo_g, cg, fw, bw, vfw, vbw, _ = _pair(which)
    for update, steps in ((True, 6), (False, 7), (False, -6)):
        fw.velocity_steps = steps
        fw.GetDefFieldFromVelocityGrid(update)
        want, s = O.orc_vel_def_field(ref_g, cg, vfw, steps, update)
        assert fw.velocity_steps == s
        got = fw.deformation_field()
        assert np.array_equal(got, want), (update, steps, np.count_nonzero(got != want))
    # the backward grid on the floating geometry
    bw.GetDefFieldFromVelocityGrid(False)
    want, _ = O.orc_vel_def_field(flo_g, cg, vbw, 6, False)
    assert np.array_equal(bw.deformation_field(), want)
    fw.close()
    bw.close()


def test_step_number_update_matches_oracle(gpu):
    ref_g, flo_g, cg, fw, bw, vfw, vbw, _ = _pair("oblique", amplitude=40.0, seed=3)
    fw.velocity_steps = 6
    fw.GetDefFieldFromVelocityGrid(True)
    want, s = O.orc_vel_def_field(ref_g, cg, vfw, 6, True)
    assert s > 6 and fw.velocity_steps == s
    assert np.array_equal(fw.deformation_field(), want)
    fw.close()
    bw.close()


@pytest.mark.parametrize("which", ["identity", "oblique", "large"])
@pytest.mark.parametrize("steps", [6, 7])
def test_exponentiate_gradient_matches_oracle(gpu, which, steps):
    ref_g, flo_g, cg, fw, bw, vfw, vbw, rng = _pair(which, seed=1)
    grad = rng.normal(size=3 * ref_g.nvox).astype(np.float32)
    fw.set_voxel_gradient(grad)
    bw.velocity_steps = steps
    fw.ExponentiateGradient(bw)
    want = O.orc_exponentiate_gradient(ref_g, cg, vbw, steps, grad)
    got = fw.voxel_gradient()
    assert np.array_equal(got, want), np.count_nonzero(got != want)
    fw.close()
    bw.close()


@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_resample_gradient_and_compose_match_oracle(gpu, which):
    ref_g, flo_g, cg, fw, bw, vfw, vbw, rng = _pair(which, seed=2)
    grad = rng.normal(size=3 * ref_g.nvox).astype(np.float32)
    fields = O.orc_vel_intermediate(ref_g, cg, vfw, 6)
    d = (fields[4] + rng.uniform(-5, 5, fields[4].size)).astype(np.float32)
    fw.set_voxel_gradient(grad)
    fw.set_deformation_field(d)
    for pad in (0.0, 2.5):
        assert np.array_equal(fw.ResampleGradient(1, pad), O.orc_resample_gradient(ref_g, grad, d, pad))
    with pytest.raises(_lib.NiftyRegError):
        fw.ResampleGradient(3, 0.0)
    fw.DefFieldCompose(ref_g, fields[2])
    assert np.array_equal(fw.deformation_field(), O.orc_def_compose(ref_g, fields[2], d))
    fw.close()
    bw.close()


@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_velocity_update_and_symmetrise_match_oracle(gpu, which):
    ref_g, flo_g, cg, fw, bw, vfw, vbw, rng = _pair(which, seed=4)
    g = rng.normal(size=vfw.size).astype(np.float32)
    fw.set_node_gradient(g)
    fw.UpdateVelocityField(0.37, True, False, True)
    want = O.orc_update_velocity(vfw, g, 0.37, (1, 0, 1))
    assert np.array_equal(fw.get_control_points(), want)
    fw.SymmetriseVelocityFields(bw)
    wf, wb = O.orc_symmetrise_velocity(cg, want, vbw)
    assert np.array_equal(fw.get_control_points(), wf) and np.array_equal(bw.get_control_points(), wb)
    with pytest.raises(_lib.NiftyRegError):
        fw.BchUpdate(0.5, 0)
    fw.close()
    bw.close()


def test_velocity_handles_refuse_the_displacement_path(gpu):
    ref_g, flo_g, cg, fw, bw, vfw, vbw, _ = _pair("identity")
    with pytest.raises(_lib.NiftyRegError):
        fw.objective()
    fw.close()
    bw.close()


def _affine(angle=0.08, shift=(2.5, -1.5, 1.0), scale=(1.03, 0.97, 1.01)):
    c, s = np.cos(angle), np.sin(angle)
    R = np.array([[c, -s, 0.02], [s, c, -0.03], [0.01, 0.04, 1.0]])
    m = np.eye(4)
    m[:3, :3] = R @ np.diag(scale)
    m[:3, 3] = shift
    return m.astype(np.float32)


@pytest.mark.parametrize("which", ["identity", "oblique", "large"])
def test_velocity_grids_with_affine_parts_match_oracle(gpu, which):
    """Grid extensions (the half affines) on both handles, the transformation matrices A / A^-1:
    exponentiation (with and without the step-number update) and the gradient exponentiation."""
    from niftyreg_amd.geometry import symmetric_half_affines
    ref_g, flo_g, cg, fw, bw, vfw, vbw, rng = _pair(which, seed=6)
    A = _affine()
    hf, hb, ext = symmetric_half_affines(A)
    assert ext
    Ainv = O.orc_nifti_inverse(A)
    fw.set_velocity_affines(hf, hf, A)
    bw.set_velocity_affines(hb, hb, Ainv)
    for update, steps in ((True, 6), (False, 7)):
        fw.velocity_steps = steps
        fw.GetDefFieldFromVelocityGrid(update)
        want, s = O.orc_vel_def_field_aff(ref_g, cg, vfw, hf, hf, steps, update)
        assert fw.velocity_steps == s
        got = fw.deformation_field()
        assert np.array_equal(got, want), (update, steps, np.count_nonzero(got != want))
    grad = rng.normal(size=3 * ref_g.nvox).astype(np.float32)
    fw.set_voxel_gradient(grad)
    bw.velocity_steps = 6
    bw.GetDefFieldFromVelocityGrid(False)  # a chain on bw that the affine case must not reuse
    fw.ExponentiateGradient(bw)
    want = O.orc_exponentiate_gradient_aff(ref_g, cg, vbw, hb, hb, Ainv, 6, grad)
    got = fw.voxel_gradient()
    assert np.array_equal(got, want), np.count_nonzero(got != want)
    fw.close()
    bw.close()
