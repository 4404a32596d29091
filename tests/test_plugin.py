"""The HIP platform plugin (niftyreg_amd/plugin) against the reference's own classes.

oracle/_ref/libnrplugin.so drives the calls reg_f3d makes for one objective value and one
gradient through either the reference's CPU Content / Compute / Measure classes or the plugin's
HIP classes.  On the CPU side this pins the driver sequence against the oracle chain; on the GPU
the two platforms must agree: the per-operation device kernels follow the scalar reference
build, so the gradients are bit-exact and the values agree to FP64 reduction order (1e-12).
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import phantoms
from niftyreg_amd.geometry import Geom, create_control_point_grid

needs_plugin = pytest.mark.skipif(not O.plugin_available(), reason="plugin check library not built")


def _case(dims=(40, 36, 32), seed=0, geom=None):
    g = geom if geom is not None else Geom.identity(dims)
    ref = phantoms.blob_phantom(g.dims, seed).ravel()
    flo = phantoms.blob_phantom(g.dims, seed + 3).ravel()
    cg, cpp = create_control_point_grid(g, [5.0 * p for p in g.pixdim])
    cpp = (cpp + np.random.default_rng(seed).uniform(-2, 2, cpp.size)).astype(np.float32)
    return g, ref, flo, cg, cpp


@needs_plugin
@pytest.mark.parametrize("which", ["cube", "oblique"])
def test_reference_classes_path_matches_oracle_chain(which):
    g, ref, flo, cg, cpp = _case(geom=_oblique((34, 30, 28)) if which == "oblique" else None)
    m, be, grad, _ = O.plugin_f3d_eval(False, g, ref, g, flo, cg, cpp)
    r, f = O.orc_rescale(ref, 2, 65), O.orc_rescale(flo, 2, 65)
    want, nmi, _ = O.orc_node_gradient_chain(g, g, cg, r, f, cpp, 68, 68, np.float32(0.999), np.float32(0.001))
    # the bending-energy gradient round-trips the points through displacements in place, and the
    # linear-energy gradient that follows (reg_f3d::GetGradient's order) sees the result
    want = O.orc_le_grad(cg, O.orc_round_trip(cg, cpp), 0.01, want)
    assert m == nmi
    assert np.array_equal(grad, want)
    # the check driver reports bending energy + 1000 x linear energy (both before the gradient)
    assert abs(be - (O.orc_be(cg, cpp) + 1000.0 * O.orc_le(cg, cpp))) <= 1e-13 * be


def _oblique(dims):
    th = 0.2
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.1, 0.9, 1.2])).astype(np.float32)
    m[:3, 3] = [4.0, -6.0, 2.5]
    return Geom(dims, m)


@needs_plugin
@pytest.mark.gpu
@pytest.mark.parametrize("lncc", [False, True])
@pytest.mark.parametrize("which", ["cube", "oblique"])
def test_hip_platform_matches_reference_platform(gpu, lncc, which):
    g, ref, flo, cg, cpp = _case(geom=_oblique((34, 30, 28)) if which == "oblique" else None)
    cm, cbe, cgrad, cvg = O.plugin_f3d_eval(False, g, ref, g, flo, cg, cpp, lncc=lncc)
    hm, hbe, hgrad, hvg = O.plugin_f3d_eval(True, g, ref, g, flo, cg, cpp, lncc=lncc)
    assert abs(hm - cm) <= 1e-12 * abs(cm)
    assert abs(hbe - cbe) <= 1e-12 * abs(cbe)
    assert np.array_equal(hvg, cvg, equal_nan=True)
    assert np.array_equal(hgrad, cgrad)


@needs_plugin
@pytest.mark.gpu
@pytest.mark.parametrize("lncc", [False, True])
@pytest.mark.parametrize("which", ["cube", "oblique"])
def test_hip_platform_fused_route_matches_reference_platform(gpu, monkeypatch, lncc, which):
    """NRH_PLUGIN_FUSED=1: the plugin records the field / warp / gradient stages and runs them as
    nrh_f3d_similarity_value / _gradient: the fused sweeps for NMI, the separable node projection for
    both measures (integer 5-voxel spacing, nodes on voxels, also on the oblique geometry).  For NMI
    the route's node gradient equals the library's own fused evaluation (nrh_f3d_objective +
    nrh_f3d_gradient) on the same inputs bit for bit.  Against the CPU classes: the value within
    1e-6, the node gradient within the fused tolerances (1e-3 L2, 2e-3 of the largest element;
    measured 4e-6 and below).  The convolved voxel gradient, read through the accessor after the
    bending-energy gradient, runs the recorded per-operation stages at the current control points,
    which the bending energy's round trip may have moved in the last bit (not at all on the cube's
    integer node positions): bit-exact on the cube, within 1e-3 (L2) on the oblique grid."""
    g, ref, flo, cg, cpp = _case(geom=_oblique((34, 30, 28)) if which == "oblique" else None)
    cm, cbe, cgrad, cvg = O.plugin_f3d_eval(False, g, ref, g, flo, cg, cpp, lncc=lncc)
    monkeypatch.setenv("NRH_PLUGIN_FUSED", "1")
    hm, hbe, hgrad, hvg = O.plugin_f3d_eval(True, g, ref, g, flo, cg, cpp, lncc=lncc)
    assert abs(hbe - cbe) <= 1e-12 * abs(cbe)
    assert abs(hm - cm) <= 1e-6 * abs(cm)
    if not lncc:
        from niftyreg_amd.f3d import F3dContent, default_params
        h = F3dContent(g, ref, g, flo, cg, cpp, params=default_params(similarity_weight=0.999, be_weight=0.001,
                                                                       le_weight=0.01))
        h.rescale_intensities((2, 65), (2, 65))
        h.objective()
        h.gradient()
        lib = h.node_gradient()
        h.close()
        assert np.array_equal(hgrad, lib)
    rel = np.linalg.norm(hgrad - cgrad) / np.linalg.norm(cgrad)
    print(f"{which} lncc={lncc}: fused route vs CPU classes, node gradient L2 {rel:.2e}")
    assert rel <= 1e-3, rel
    assert np.abs(hgrad - cgrad).max() <= 2e-3 * np.abs(cgrad).max()
    if which == "cube":
        assert np.array_equal(hvg, cvg, equal_nan=True)
    else:
        ok = np.isfinite(cvg) & np.isfinite(hvg)
        assert np.array_equal(np.isfinite(cvg), np.isfinite(hvg))
        vrel = np.linalg.norm(hvg[ok] - cvg[ok]) / np.linalg.norm(cvg[ok])
        assert vrel <= 1e-3, vrel


def _f3d2_case(which, seed=2, amplitude=2.0):
    from niftyreg_amd.geometry import create_symmetric_control_point_grids
    if which == "cube":
        rg_ = fg = Geom.identity((36, 34, 30))
    else:
        rg_, fg = _oblique((34, 30, 28)), Geom.identity((32, 34, 26), (1.1, 1.0, 1.2))
    ref = phantoms.blob_phantom(rg_.dims, seed).ravel()
    flo = phantoms.blob_phantom(fg.dims, seed + 3).ravel()
    cg, fw, bw = create_symmetric_control_point_grids(rg_, fg, [5.0, 5.0, 5.0])
    rng = np.random.default_rng(seed)
    fw = (fw + rng.uniform(-amplitude, amplitude, fw.size)).astype(np.float32)
    bw = (bw + rng.uniform(-amplitude, amplitude, bw.size)).astype(np.float32)
    return rg_, fg, ref, flo, cg, fw, bw


@needs_plugin
@pytest.mark.gpu
@pytest.mark.parametrize("lncc", [False, True])
@pytest.mark.parametrize("which", ["cube", "oblique"])
def test_hip_platform_symmetric_matches_reference_platform(gpu, which, lncc):
    """reg_f3d2's calls (velocity exponentiation, symmetric NMI or LNCC, gradient exponentiation
    along the opposite grid, both directions' convolution, penalties) through both platforms."""
    rg_, fg, ref, flo, cg, fw, bw = _f3d2_case(which)
    c = O.plugin_f3d2_eval(False, rg_, ref, fg, flo, cg, fw, bw, lncc=lncc)
    h = O.plugin_f3d2_eval(True, rg_, ref, fg, flo, cg, fw, bw, lncc=lncc)
    assert abs(h[0] - c[0]) <= 1e-12 * abs(c[0])
    assert abs(h[1] - c[1]) <= 1e-12 * abs(c[1])
    assert h[6] == c[6]
    for k in (2, 3, 4, 5):
        assert np.array_equal(h[k], c[k]), (k, np.count_nonzero(h[k] != c[k]))


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["cube", "oblique"])
def test_hip_platform_symmetric_with_affine_matches_reference_platform(gpu, which):
    """The same calls with an input affine: the half affines in the grids' extensions (read by
    HipF3dContent), the affine and its inverse in the contents."""
    rg_, fg, ref, flo, cg, fw, bw = _f3d2_case(which)
    c_, s_ = np.cos(0.05), np.sin(0.05)
    A = np.eye(4)
    A[:3, :3] = [[c_, -s_, 0.01], [s_, c_, 0.0], [0.0, 0.02, 1.0]]
    A[:3, 3] = [1.2, -0.8, 0.5]
    A = A.astype(np.float32)
    hf, hb, _ = O.orc_half_affines(A)
    c = O.plugin_f3d2_eval(False, rg_, ref, fg, flo, cg, fw, bw, affine=A, halves=(hf, hb))
    h = O.plugin_f3d2_eval(True, rg_, ref, fg, flo, cg, fw, bw, affine=A, halves=(hf, hb))
    assert abs(h[0] - c[0]) <= 1e-12 * abs(c[0])
    assert abs(h[1] - c[1]) <= 1e-12 * abs(c[1])
    assert h[6] == c[6]
    for k in (2, 3, 4, 5):
        assert np.array_equal(h[k], c[k]), (k, np.count_nonzero(h[k] != c[k]))


@needs_plugin
@pytest.mark.parametrize("which", ["cube", "oblique"])
def test_symmetric_check_sequence_matches_reg_f3d2(which):
    """The plugin check's call sequence on the reference CPU classes equals reg_f3d2 itself (one
    level, -be 0.001 -le 0.01): the same transformation gradients of both directions, the same
    rounded grids and step numbers, the similarity term to FP64 order."""
    rg_, fg, ref, flo, cg, fw, bw = _f3d2_case(which)
    r = O.RefF3d2(rg_, ref, fg, flo, be=0.001, le=0.01, spacing=5.0, levels=1)
    assert O.geom_from_nr(r.cpp_geom).dims == cg.dims
    simw, bew, lew = r.weights()
    r.set_cpp(fw, bw)
    rv, rterms = r.value()
    rgf, rgb, _ = r.gradient(normalise=False)
    rf, rb, rsteps = r.get_cpp()
    r.close()
    c = O.plugin_f3d2_eval(False, rg_, ref, fg, flo, cg, fw, bw, simw=simw, bew=bew, lew=lew)
    assert abs(simw * c[0] - rterms[0]) <= 1e-12 * abs(rterms[0])
    assert c[6] == rsteps
    assert np.array_equal(c[2], rgf) and np.array_equal(c[3], rgb)
    assert np.array_equal(c[4], rf) and np.array_equal(c[5], rb)
