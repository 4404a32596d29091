"""The linear-energy penalty on the device against the oracle and the reference's reg_f3d.

The device kernels follow reg_spline_approxLinearEnergy{Value,Gradient}3D: the Jacobian of each
node from the first-order basis, its rotation removed by niftilib's polar decomposition, the
symmetric strain squared (value) or gathered back over the 27 neighbours (gradient).  niftilib
inverts a 3x3 through a long-double reciprocal of the determinant, which the device reproduces
(recip_x87), so the gradient is bit-exact; the value is a FP64 sum in a fixed order (1e-12
relative).
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import _lib, phantoms
from niftyreg_amd.f3d import F3dContent, default_params
from niftyreg_amd.geometry import Geom, create_control_point_grid

pytestmark = pytest.mark.gpu


def _oblique(dims):
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.1, 0.9, 1.3])).astype(np.float32)
    m[:3, 3] = [3.0, -2.0, 5.0]
    return Geom(dims, m)


@pytest.mark.parametrize("which", ["identity", "oblique", "large"])
def test_linear_energy_matches_oracle(gpu, which):
    g = {"identity": Geom.identity((40, 36, 32)), "oblique": _oblique((30, 28, 26)),
         "large": Geom.identity((160, 160, 128), spacing=(1.0, 1.0, 1.5))}[which]
    ref = phantoms.blob_phantom(g.dims, 1).ravel()
    cg, cpp = create_control_point_grid(g, [5.0 * p for p in g.pixdim])
    cpp = (cpp + np.random.default_rng(2).uniform(-3, 3, cpp.size)).astype(np.float32)
    h = F3dContent(g, ref, g, ref, cg, cpp, params=default_params(le_weight=0.01))
    v = h.ApproxLinearEnergy()
    want = O.orc_le(cg, cpp)
    assert abs(v - want) <= 1e-12 * want
    init = (np.random.default_rng(3).normal(size=3 * cg.nvox) * 1e-3).astype(np.float32)
    h.set_node_gradient(init)
    h.ApproxLinearEnergyGradient(0.01)
    got, want = h.node_gradient(), O.orc_le_grad(cg, cpp, 0.01, init)
    assert np.array_equal(got, want), np.count_nonzero(got != want)
    h.close()


def _oblique_scaled(dims, scales):
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag(scales)).astype(np.float32)
    m[:3, 3] = [3.0, -2.0, 5.0]
    return Geom(dims, m)


@pytest.mark.parametrize("which", ["axis_anisotropic", "oblique_isotropic", "oblique_anisotropic"])
def test_linear_energy_of_identity_grid(gpu, which):
    """The reference multiplies the grid's mm->voxel matrix with the transposed node Jacobian
    (reg-lib/cpu/_reg_localTrans_regul.cpp:619-621), so an identity grid has zero linear energy
    when that product is a rotation (axis-aligned, or oblique with isotropic spacing) and a
    non-zero one for an oblique anisotropic grid; the device reproduces both."""
    g = {"axis_anisotropic": Geom.identity((30, 28, 26), spacing=(1.1, 0.9, 1.3)),
         "oblique_isotropic": _oblique_scaled((30, 28, 26), [1.1, 1.1, 1.1]),
         "oblique_anisotropic": _oblique_scaled((30, 28, 26), [1.1, 0.9, 1.3])}[which]
    ref = phantoms.blob_phantom(g.dims, 1).ravel()
    cg, cpp = create_control_point_grid(g, [5.0 * p for p in g.pixdim])
    h = F3dContent(g, ref, g, ref, cg, cpp, params=default_params(le_weight=0.01))
    v = h.ApproxLinearEnergy()
    if which == "oblique_anisotropic":
        want = O.orc_le(cg, cpp)
        assert want > 1e-4 and abs(v - want) <= 1e-12 * want
    else:
        assert v < 1e-10
    h.close()


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_objective_and_gradient_with_linear_energy_match_reference_driver(gpu):
    """nrh_f3d_objective / nrh_f3d_gradient with -be 0.001 -le 0.01 against reg_f3d<float> (one
    level): the weighted terms and value to 1e-6, the normalised node gradient to 1e-3."""
    g = Geom.identity((48, 44, 40))
    ref, flo = phantoms.config0_pair(3, g.dims)
    r = O.RefF3d(g, ref, g, flo, be=0.001, le=0.01, spacing=-5.0, levels=1)
    cpp0 = r.get_cpp()
    cpp = (cpp0 + np.random.default_rng(1).uniform(-1.5, 1.5, cpp0.size)).astype(np.float32)
    r.set_cpp(cpp)
    rv, rterms = r.value()
    rgrad, rmax = r.gradient(normalise=True)
    simw, bew, lew = r.weights(with_le=True)
    cg = O.geom_from_nr(r.cpp_geom)
    r.close()
    assert lew > 0 and rterms[2] > 0
    # the exact pass reproduces the build compared with (RefF3d: the default SSE build)
    h = F3dContent(g, ref.ravel(), g, flo.ravel(), cg, cpp,
                   params=default_params(similarity_weight=simw, be_weight=bew, le_weight=lew,
                                         field_order=_lib.FIELD_ORDER_SSE))
    h.rescale_intensities((2, 65), (2, 65))  # NMI, 64 bins: [2, bins + 1]
    v, terms = h.objective()
    assert abs(terms[2] - rterms[2]) <= 1e-6 * abs(rterms[2])
    assert abs(v - rv) <= 1e-6 * abs(rv)
    maxlen = h.gradient(normalise=True)
    grad = h.node_gradient()
    assert abs(maxlen - rmax) <= 1e-4 * rmax
    rel = np.linalg.norm(grad - rgrad) / np.linalg.norm(rgrad)
    assert rel <= 1e-3, rel
    h.close()


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_registration_with_linear_energy_matches_reference(gpu):
    """reg_f3d -be 0.001 -le 0.01, three levels, on both platforms.  Line-search decisions depend
    on the last bits of each value, so parity is the final objective's similarity part, scored
    with the reference's own NMI: the device's may differ from the SSE build's by no more than the
    scalar build's does plus 1e-4, and its linear energy stays of the reference's order."""
    from niftyreg_amd.f3d import register
    ref, flo = phantoms.config0_pair(0)
    g = Geom.identity((64, 64, 64))
    cg, cpp, nmi, its = register(g, ref, g, flo, be=0.001, le=0.01, spacing=-5.0, levels=3, maxit=50)
    rcg, rcpp = O.ref_f3d_run(g, ref, g, flo, be=0.001, le=0.01, spacing=-5.0, levels=3, maxit=50)
    _, scpp = O.ref_f3d_run(g, ref, g, flo, be=0.001, le=0.01, spacing=-5.0, levels=3, maxit=50, scalar=True)
    assert tuple(cg.dims) == tuple(rcg.dims)
    nmi_gpu = O.ref_nmi_at(g, ref, g, flo, cpp, levels=3)
    nmi_ref = O.ref_nmi_at(g, ref, g, flo, rcpp, levels=3)
    nmi_scalar = O.ref_nmi_at(g, ref, g, flo, scpp, levels=3)
    g0, ident = create_control_point_grid(g, [5.0] * 3)
    print(f"LE registration: device {nmi_gpu:.6f}, reference SSE {nmi_ref:.6f}, scalar {nmi_scalar:.6f}")
    assert nmi_gpu > O.ref_nmi_at(g, ref, g, flo, ident) + 0.01
    assert abs(nmi_gpu - nmi_ref) <= abs(nmi_scalar - nmi_ref) + 1e-4
    le_gpu, le_ref = O.orc_le(cg, cpp), O.orc_le(rcg, rcpp)
    assert le_gpu <= 2.0 * le_ref + 1e-6, (le_gpu, le_ref)
