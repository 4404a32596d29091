"""The linear-energy penalty of the oracle against the reference build and closed forms.

reg_spline_approxLinearEnergy{Value,Gradient}3D remove the rotation of each node's Jacobian by a
polar decomposition (niftilib's Newton iteration) before measuring the symmetric strain; the
oracle restates that arithmetic, including niftilib's long-double reciprocal, so the gradient is
bit-exact and the value agrees to the reference's OpenMP summation order (1e-13).
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import phantoms
from niftyreg_amd.geometry import Geom, create_control_point_grid

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build absent")


def _oblique(dims):
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.1, 0.9, 1.3])).astype(np.float32)
    m[:3, 3] = [3.0, -2.0, 5.0]
    return Geom(dims, m)


@needs_ref
@pytest.mark.parametrize("which", ["identity", "oblique"])
def test_linear_energy_matches_reference(which):
    refg = Geom.identity((40, 36, 32)) if which == "identity" else _oblique((30, 28, 26))
    cg, cpp = create_control_point_grid(refg, [5.0 * p for p in refg.pixdim])
    cpp = (cpp + np.random.default_rng(0).uniform(-3, 3, cpp.size)).astype(np.float32)
    v, rv = O.orc_le(cg, cpp), O.ref_le(cg, cpp)
    assert abs(v - rv) <= 1e-13 * rv
    init = (np.random.default_rng(1).normal(size=3 * cg.nvox) * 1e-3).astype(np.float32)
    assert np.array_equal(O.orc_le_grad(cg, cpp, 0.01, init), O.ref_le_grad(cg, cpp, 0.01, init))


def test_first_order_basis_literals():
    import ctypes
    bx = np.zeros(27, np.float32)
    by = np.zeros(27, np.float32)
    bz = np.zeros(27, np.float32)
    O.orc().orc_first_order_basis(O.fp(bx), O.fp(by), O.fp(bz))
    # the reference's table (reg-lib/cpu/_reg_splineBasis.cpp:182-290) starts and ends with these
    assert bx[0] == np.float32(-0.0138889) and by[1] == np.float32(-0.0555556) and bz[4] == np.float32(-0.222222)
    assert bx[26] == np.float32(0.0138889) and bz[13] == np.float32(0.0)
    assert np.allclose(bx.sum(), 0.0, atol=1e-6) and np.allclose(bz.reshape(3, 9).sum(axis=1), [-0.5, 0, 0.5], atol=1e-5)


def test_linear_energy_of_rigid_and_scaled_grids():
    refg = Geom.identity((30, 30, 30))
    cg, ident = create_control_point_grid(refg, [5.0] * 3)
    G = cg.nvox
    pos = ident.reshape(3, G).astype(np.float64)
    A = phantoms.random_affine(3, scale=(1.0, 1.0), shear=0.0)[:3]
    rigid = (A[:, :3] @ pos + A[:, 3:4]).astype(np.float32).ravel()
    assert O.orc_le(cg, ident) < 1e-10
    assert O.orc_le(cg, rigid) < 1e-8  # the polar decomposition removes the rotation
    scaled = (1.1 * pos).astype(np.float32).ravel()
    # isotropic strain 0.1: sum of the squared symmetric part = 3 x 0.1^2 per interior node, / 3G
    interior = (cg.dims[0] - 2) * (cg.dims[1] - 2) * (cg.dims[2] - 2)
    assert abs(O.orc_le(cg, scaled) - 3 * 0.01 * interior / (3 * G)) < 1e-4 * 0.01
