"""Pinning the oracle (CPU): golden vectors from the reference build, the scalar reference build
when present, and closed forms that a plausible mistake (a dropped term, a transposed
operand, a wrong sign or index) would break."""
import glob
import os

import numpy as np
import pytest

import oracles as O
from niftyreg_amd import phantoms
from niftyreg_amd.geometry import Geom, create_control_point_grid

HERE = os.path.dirname(os.path.abspath(__file__))
RB = FB = 68


def load(name):
    return dict(np.load(os.path.join(HERE, "golden", name), allow_pickle=False))


def geoms(d):
    refg = Geom(tuple(d["ref_dims"]), d["ref_xyz"], d["ref_ijk"], tuple(float(v) for v in d["ref_pixdim"]))
    cppg = Geom(tuple(d["cpp_dims"]), d["cpp_xyz"], d["cpp_ijk"], tuple(float(v) for v in d["cpp_pixdim"]))
    return refg, cppg


F3D_FIXTURES = sorted(os.path.basename(p) for p in glob.glob(os.path.join(HERE, "golden", "f3d_*.npz")))


@pytest.mark.parametrize("name", F3D_FIXTURES)
def test_golden_chain(name):
    """Each oracle stage reproduces the scalar reference build bit for bit, and the SSE build
    within the stated tolerances (field 1e-4 voxel, NMI 1e-5 relative, gradient 1e-3)."""
    d = load(name)
    refg, cppg = geoms(d)
    cpp = d["cpp"]
    field = O.orc_field(refg, cppg, cpp)
    assert np.array_equal(field, d["scalar_field"])
    vox = min(refg.pixdim)
    assert np.abs(field - d["sse_field"]).max() <= 1e-4 * vox
    ref_r = O.orc_rescale(d["ref"], 2, RB - 3)
    flo_r = O.orc_rescale(d["flo"], 2, FB - 3)
    assert np.array_equal(ref_r, d["scalar_ref_rescaled"]) and np.array_equal(flo_r, d["scalar_flo_rescaled"])
    war = O.orc_resample(refg, flo_r, refg, field)
    assert np.array_equal(war, d["scalar_warped"], equal_nan=True)
    wg = O.orc_image_gradient(refg, flo_r, refg, field)
    assert np.array_equal(wg, d["scalar_warped_grad"])
    nmi, pro, logh, ent = O.orc_nmi(ref_r, war, RB, FB)
    assert nmi == d["scalar_nmi"]
    assert np.array_equal(pro, d["scalar_hist_pro"]) and np.array_equal(logh, d["scalar_hist_log"])
    assert np.array_equal(ent, d["scalar_entropies"])
    assert abs(nmi - d["sse_nmi"]) <= 1e-5 * nmi
    vg = O.orc_nmi_grad(ref_r, war, wg, RB, FB, logh, ent)
    assert np.array_equal(vg, d["scalar_voxel_grad"])
    conv = O.orc_convolve_voxel_gradient(refg, cppg, vg)
    assert np.array_equal(conv, d["scalar_conv_voxel_grad"])
    ng = O.orc_voxel_to_node(cppg, refg, conv, refg, np.float32(0.999))
    assert np.array_equal(ng, d["scalar_node_grad"])
    sse = d["sse_node_grad"]
    assert np.abs(ng - sse).max() <= 1e-3 * np.abs(sse).max()
    # the reference reduces the per-node terms with OpenMP, so its last bits follow its thread
    # count; the oracle sums in node order
    assert abs(O.orc_be(cppg, cpp) - d["scalar_be"]) <= 1e-13 * abs(d["scalar_be"])
    assert np.array_equal(O.orc_be_grad(cppg, cpp, np.float32(0.001)), d["scalar_be_grad"])


def test_golden_objective_driver():
    """The whole reg_f3d objective and normalised gradient (reference driver, SSE build) from the
    oracle's chain: NMI 1e-5 relative, gradient 1e-3 of its max."""
    d = load("objective_config0_small.npz")
    refg = Geom.identity((40, 36, 32))
    cppg = Geom(tuple(d["cpp_dims"]), d["cpp_xyz"], pixdim=tuple(float(v) for v in d["cpp_pixdim"]))
    simw, bew = (np.float32(v) for v in d["weights"])
    ref_r = O.orc_rescale(d["ref"], 2, RB - 3)
    flo_r = O.orc_rescale(d["flo"], 2, FB - 3)
    assert np.array_equal(ref_r, d["ref_rescaled"]) and np.array_equal(flo_r, d["flo_rescaled"])
    ng, nmi, _ = O.orc_node_gradient_chain(refg, refg, cppg, ref_r, flo_r, d["cpp"], RB, FB, simw, bew)
    be = O.orc_be(cppg, d["cpp"])
    value = float(simw) * nmi - float(bew) * be
    assert abs(value - d["value"]) <= 1e-5 * abs(d["value"])
    m = float(O.orc().orc_max_length(O.fp(ng), cppg.nvox))
    assert abs(m - d["max_length"]) <= 1e-4 * m
    O.orc().orc_normalise_gradient(O.fp(ng), cppg.nvox, O.ctypes.c_double(m))
    assert np.abs(ng - d["grad"]).max() <= 1e-3 * np.abs(d["grad"]).max()


# ---------------------------------------------------------------- closed forms
def test_bspline_partition_of_unity():
    vals = np.zeros(4, np.float32)
    d1 = np.zeros(4, np.float32)
    for t in np.linspace(0, 0.999, 37, dtype=np.float32):
        O.orc().orc_bspline_basis_d1(O.ctypes.c_float(t), O.fp(vals), O.fp(d1))
        assert abs(vals.sum() - 1) < 1e-6 and abs(d1.sum()) < 1e-6
        # value at t: B(t+1), B(t), B(1-t), B(2-t) of the centred cubic B-spline
        B = lambda x: (2 / 3 - x * x + abs(x) ** 3 / 2) if abs(x) < 1 else ((2 - abs(x)) ** 3 / 6 if abs(x) < 2 else 0)
        want = [B(t + 1), B(t), B(1 - t), B(2 - t)]
        assert np.allclose(vals, want, atol=2e-7)


@pytest.mark.parametrize("spacing", [5.0, 4.0, 3.3])
def test_field_reproduces_affine_grids(spacing):
    """Identity, translated and scaled control grids give the analytic field (cubic B-splines
    reproduce linear functions), on both the 5-voxel table path and the general path."""
    refg = Geom.identity((31, 27, 23), spacing=(1.0, 1.0, 1.0), origin=(4.0, -2.0, 7.0))
    cppg, ident = create_control_point_grid(refg, [spacing] * 3)
    z, y, x = np.meshgrid(np.arange(23), np.arange(27), np.arange(31), indexing="ij")
    world = np.stack([x + 4.0, y - 2.0, z + 7.0]).reshape(3, -1)
    f = O.orc_field(refg, cppg, ident).reshape(3, -1)
    assert np.abs(f - world).max() < 2e-5 * 40
    G = cppg.nvox
    shifted = ident.copy()
    shifted[:G] += 2.0
    f = O.orc_field(refg, cppg, shifted).reshape(3, -1)
    assert np.abs(f[0] - (world[0] + 2.0)).max() < 1e-4 and np.abs(f[1:] - world[1:]).max() < 1e-4
    f = O.orc_field(refg, cppg, (ident * np.float32(1.1)).astype(np.float32)).reshape(3, -1)
    assert np.abs(f - 1.1 * world).max() < 1e-4


def test_resample_integer_shift_and_constant():
    refg = Geom.identity((20, 18, 16))
    img = phantoms.blob_phantom(refg.dims, 3).ravel()
    z, y, x = np.meshgrid(np.arange(16), np.arange(18), np.arange(20), indexing="ij")
    field = np.stack([x + 1.0, y, z - 2.0]).astype(np.float32).reshape(-1)
    w = O.orc_resample(refg, img, refg, field, pad=0.0).reshape(16, 18, 20)
    im = img.reshape(16, 18, 20)
    assert np.array_equal(w[2:, :, :-1], im[:-2, :, 1:])
    assert np.all(w[:, :, -1] == 0) and np.all(w[:2] == 0)
    wn = O.orc_resample(refg, img, refg, field).reshape(16, 18, 20)
    assert np.all(np.isnan(wn[:, :, -1]))  # NaN padding marks voxels mapped outside
    const = np.full(refg.nvox, 3.25, np.float32)
    rng = np.random.default_rng(0)
    field2 = (np.stack([x, y, z]).reshape(-1) + rng.uniform(-0.4, 0.4, 3 * refg.nvox) + 2.0).astype(np.float32)
    w2 = O.orc_resample(refg, const, refg, field2, pad=3.25)
    assert np.allclose(w2, 3.25, atol=1e-6)


def test_image_gradient_of_ramp():
    refg = Geom.identity((16, 15, 14))
    z, y, x = np.meshgrid(np.arange(14), np.arange(15), np.arange(16), indexing="ij")
    ramp = (0.5 * x - 0.25 * y + 2.0 * z).astype(np.float32).reshape(-1)
    rng = np.random.default_rng(1)
    field = (np.stack([x, y, z]).reshape(-1) * 0.8 + 1.3 + rng.uniform(0, 0.5, 3 * refg.nvox)).astype(np.float32)
    g = O.orc_image_gradient(refg, ramp, refg, field).reshape(3, -1)
    ok = ~(g == 0).all(0)
    assert ok.mean() > 0.5
    assert np.allclose(g[0, ok], 0.5, atol=1e-5) and np.allclose(g[1, ok], -0.25, atol=1e-5)
    assert np.allclose(g[2, ok], 2.0, atol=1e-5)


def test_nmi_identical_images_and_bounds():
    rng = np.random.default_rng(0)
    a = O.orc_rescale(rng.uniform(0, 1, 5000).astype(np.float32), 2, RB - 3)
    v_same, pro, _, ent = O.orc_nmi(a, a, RB, FB)
    b = O.orc_rescale(rng.uniform(0, 1, 5000).astype(np.float32), 2, RB - 3)
    v_ind, _, _, _ = O.orc_nmi(a, b, RB, FB)
    # NMI = (H_r + H_f) / H_rf lies in [1, 2]; identical images are far above independent ones
    assert 1.0 <= v_ind < v_same <= 2.0
    assert ent[3] == pytest.approx(5000.0)  # total Parzen mass = number of counted voxels
    assert abs(pro[:RB * FB].sum() - 1.0) < 1e-12
    # marginals are the row / column sums of the joint histogram
    J = pro[:RB * FB].reshape(FB, RB)
    assert np.allclose(pro[RB * FB:RB * FB + RB], J.sum(0)) and np.allclose(pro[RB * FB + RB:], J.sum(1))


def test_convolution_keeps_constants_and_matches_interior_filter():
    refg = Geom.identity((40, 3, 3))
    data = np.full(refg.nvox, 2.5, np.float32)
    O.orc().orc_kernel_convolution_axis(O.ctypes.byref(O.og(refg)), O.fp(data), O.ctypes.c_float(5.0), 3, None, 0)
    assert np.allclose(data, 2.5, rtol=1e-6)
    rng = np.random.default_rng(0)
    line = rng.normal(size=refg.nvox).astype(np.float32)
    out = line.copy()
    O.orc().orc_kernel_convolution_axis(O.ctypes.byref(O.og(refg)), O.fp(out), O.ctypes.c_float(5.0), 3, None, 0)
    r = np.arange(-10, 11) / 5.0
    k = np.where(np.abs(r) < 1, 2 / 3 - r ** 2 + np.abs(r) ** 3 / 2, np.where(np.abs(r) < 2, (2 - np.abs(r)) ** 3 / 6, 0))
    row = line.reshape(3, 3, 40)[1, 1]
    want = np.convolve(row, k[::-1], mode="same") / k.sum()
    assert np.allclose(out.reshape(3, 3, 40)[1, 1, 10:30], want[10:30], atol=1e-6)


def test_bending_energy_affine_zero_and_gradient_fd():
    refg = Geom.identity((30, 28, 26))
    cppg, ident = create_control_point_grid(refg, [5.0] * 3)
    G = cppg.nvox
    aff = ident.copy().reshape(3, G)
    aff = (np.array([[1.05, 0.02, 0.0], [0.01, 0.97, 0.03], [0.0, -0.02, 1.02]]) @ aff + 3.0).astype(np.float32).reshape(-1)
    rng = np.random.default_rng(2)
    rnd = (ident + rng.uniform(-1, 1, ident.size)).astype(np.float32)
    assert O.orc_be(cppg, aff) < 1e-6 * O.orc_be(cppg, rnd)
    w = np.float32(0.5)
    g = O.orc_be_grad(cppg, rnd, w)
    # d BE / d cp at an interior node: grad = 1.5 * w * dBE/dcp (the value divides by 3G, the
    # gradient scales by w/G and carries the factor 2 of the squares over 2)
    nx, ny, nz = cppg.dims
    for (i, j, k, c) in [(5, 5, 5, 0), (4, 6, 5, 1), (6, 4, 3, 2)]:
        idx = (k * ny + j) * nx + i + c * G
        eps = np.float32(0.05)
        p, m = rnd.copy(), rnd.copy()
        p[idx] += eps
        m[idx] -= eps
        fd = (O.orc_be(cppg, p) - O.orc_be(cppg, m)) / (2 * float(eps))
        assert g[idx] == pytest.approx(1.5 * float(w) * fd, rel=2e-2)


def test_conjugate_gradient_invariants():
    rng = np.random.default_rng(0)
    n = 300
    g1 = rng.normal(size=n).astype(np.float32)
    G, H = np.zeros(n, np.float32), np.zeros(n, np.float32)
    grad = g1.copy()
    O.orc().orc_cg_update(O.fp(grad), O.fp(G), O.fp(H), n, 1)
    assert np.array_equal(G, -g1) and np.array_equal(H, -g1) and np.array_equal(grad, g1)
    # the same gradient again: Polak-Ribiere gamma is 0 and the direction is unchanged
    grad = g1.copy()
    gam = O.orc().orc_cg_update(O.fp(grad), O.fp(G), O.fp(H), n, 0)
    assert gam == 0.0 and np.array_equal(grad, g1)
    # an orthogonal new gradient: gamma = |g2|^2 / |g1|^2
    g2 = rng.normal(size=n).astype(np.float32)
    g2 -= np.float32(g2 @ g1 / (g1 @ g1)) * g1
    grad = g2.copy()
    gam = O.orc().orc_cg_update(O.fp(grad), O.fp(G), O.fp(H), n, 0)
    assert gam == pytest.approx(float(g2 @ g2) / float(g1 @ g1), rel=1e-4)
    assert np.allclose(grad, g2 + gam * g1, rtol=1e-5, atol=1e-6)
