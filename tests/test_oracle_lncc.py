"""The LNCC oracle against the reference's golden vectors and against what the mathematics fixes.

Golden vectors (tests/golden/lncc_*.npz, written by tests/golden/make_golden.py from the
reference build): the oracle follows the scalar build, so its voxel gradient must equal that
build's bit for bit; its value sums the voxels in order where the reference reduces with
OpenMP, so the value agrees to 1e-12.  The default SSE build sums its convolution lines in four
lanes, and agrees to 1e-6.
"""
import os

import numpy as np
import pytest

import oracles as O
from niftyreg_amd import phantoms
from niftyreg_amd.geometry import Geom

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.parametrize("name", ["blobs", "masked"])
def test_lncc_golden(name):
    d = np.load(os.path.join(HERE, "golden", f"lncc_{name}.npz"))
    g = Geom.identity(tuple(int(v) for v in d["dims"]))
    mask = d["mask"] if d["mask"].size else None
    r = O.orc_rescale(d["ref"], 0, 1)
    assert np.array_equal(r, d["ref_rescaled"])
    war, wg = d["warped"], d["warped_grad"]
    v = O.orc_lncc(g, r, war, mask)
    assert abs(v - d["scalar_value"]) <= 1e-12 * abs(d["scalar_value"])
    assert abs(v - d["sse_value"]) <= 1e-6 * abs(d["sse_value"])
    grad = O.orc_lncc_grad(g, r, war, wg, mask)
    assert np.array_equal(grad, d["scalar_grad"])
    assert np.abs(grad - d["sse_grad"]).max() <= 1e-5 * np.abs(d["sse_grad"]).max()


def test_convolution_of_a_constant_is_constant():
    g = Geom.identity((23, 19, 17))
    rng = np.random.default_rng(0)
    mask = (rng.uniform(size=g.nvox) > 0.3).astype(np.int32) - 1
    out = O.orc_convolution(g, np.full(g.nvox, 0.37, np.float32), -5.0, 2, mask)
    keep = mask > -1
    assert np.all(np.isnan(out[~keep]))
    assert np.allclose(out[keep], 0.37, rtol=1e-6)


def test_convolution_matches_textbook_gaussian_in_the_interior():
    g = Geom.identity((48, 40, 36))
    img = phantoms.blob_phantom(g.dims, 2).ravel()
    out = O.orc_convolution(g, img, -2.0, 2).reshape(g.dims[::-1])
    x = np.arange(-6, 7)
    k = np.exp(-x * x / 8.0)
    k /= k.sum()
    a = img.reshape(g.dims[::-1]).astype(np.float64)
    for axis in range(3):
        a = np.apply_along_axis(lambda l: np.convolve(l, k, mode="same"), axis, a)
    inner = (slice(7, -7),) * 3
    assert np.allclose(out[inner], a[inner], rtol=1e-5, atol=1e-6)


def test_lncc_of_an_image_with_itself_and_its_affine_intensity_map_is_one():
    g = Geom.identity((30, 28, 26))
    img = O.orc_rescale(phantoms.blob_phantom(g.dims, 4).ravel(), 0, 1)
    assert abs(O.orc_lncc(g, img, img) - 1.0) < 1e-4
    neg = (1.0 - 0.5 * img).astype(np.float32)  # |local correlation| is invariant to a x + b
    assert abs(O.orc_lncc(g, img, neg) - 1.0) < 1e-4


def test_lncc_gradient_tracks_finite_differences():
    """The reference's voxel gradient is -dLNCC/dW (times the warped-image gradient) in an
    approximate form, so it is checked as a direction: along ten smooth perturbations of the
    warped image its projection correlates with central differences of the value (>= 0.95),
    with the sign of a descent direction."""
    g = Geom.identity((26, 24, 22))
    ref = O.orc_rescale(phantoms.blob_phantom(g.dims, 5).ravel(), 0, 1)
    war = O.orc_rescale(phantoms.blob_phantom(g.dims, 9).ravel(), 0, 1)
    ex = np.zeros(3 * g.nvox, np.float32)
    ex[: g.nvox] = 1.0  # warped gradient (1, 0, 0): the x component is the intensity derivative
    gx = O.orc_lncc_grad(g, ref, war, ex)[: g.nvox].astype(np.float64)
    fds, ans = [], []
    eps = 3e-2
    for seed in range(10):
        d = O.orc_convolution(g, np.random.default_rng(seed).normal(size=g.nvox).astype(np.float32), -2.0, 2)
        d = d.astype(np.float64)
        vp = O.orc_lncc(g, ref, (war + eps * d).astype(np.float32))
        vm = O.orc_lncc(g, ref, (war - eps * d).astype(np.float32))
        fds.append((vp - vm) / (2 * eps))
        ans.append(-float(np.dot(gx, d)))
    assert np.corrcoef(fds, ans)[0, 1] >= 0.95
    assert np.mean(np.sign(fds) == np.sign(ans)) >= 0.8
