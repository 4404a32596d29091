"""The Aladin oracle (block selection, block matching, LTS) against the reference CPU build and
against cases whose answer is fixed by construction.

Block selection and block matching are integer/float work that the oracle restates step for
step, so they must agree bit for bit with the reference.  The LTS fit is FP64 least squares
solved differently (Householder QR here, Eigen's column-pivoting QR in the reference), so the
matrices agree to 1e-5 relative; the kept subset is the same std::nth_element choice.
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import phantoms
from niftyreg_amd.geometry import Geom

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference build absent")


def _oblique(dims):
    th = 0.15
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m = np.eye(4, dtype=np.float32)
    m[:3, :3] = (R @ np.diag([1.2, 0.9, 1.1])).astype(np.float32)
    m[:3, 3] = [-30.0, 12.5, 4.0]
    return Geom(dims, m)


def _warp_by(g, img, A):
    """Resample img (on g) through the affine A (mm), NaN outside, with the oracle."""
    n = g.nvox
    T = (A.astype(np.float64) @ g.xyz.astype(np.float64))
    z, y, x = np.meshgrid(np.arange(g.dims[2]), np.arange(g.dims[1]), np.arange(g.dims[0]), indexing="ij")
    p = np.stack([x.ravel(), y.ravel(), z.ravel(), np.ones(n)])
    f = (T @ p)[:3].astype(np.float32).reshape(-1)
    return O.orc_resample(g, img.ravel(), g, np.ascontiguousarray(f))


CASES = {
    "cube_48": (Geom.identity((48, 48, 48)), None),
    "ragged": (Geom.identity((45, 38, 29), spacing=(1.0, 1.25, 1.5)), None),
    "oblique_masked": (_oblique((40, 36, 33)), "mask"),
}


@pytest.fixture(params=list(CASES))
def aladin_case(request):
    g, extra = CASES[request.param]
    img = phantoms.blob_phantom(g.dims, 3).ravel().copy()
    mask = None
    if extra == "mask":
        mask = np.zeros(g.nvox, np.int32)
        mask[: g.nvox // 5] = -1
        img[7::97] = np.nan
    A = phantoms.random_affine(11, rot_deg=4.0, trans_mm=2.5, scale=(0.97, 1.03), shear=0.02).astype(np.float32)
    warped = _warp_by(g, img.reshape(g.dims[::-1]), A)
    if request.param == "ragged":
        # a NaN slab leaves the blocks under it without any offset of more than 32 voxels
        warped.reshape(g.dims[::-1])[:, :, :13] = np.nan
    return g, img, mask, warped


@needs_ref
def test_active_blocks_match_reference(aladin_case):
    g, img, mask, _ = aladin_case
    for percent in (50, 100, 20):
        tb, n = O.orc_active_blocks(g, img, percent, mask)
        rtb, rn, _, _, _ = O.ref_block_matching(g, img, None, percent, mask)
        assert n == rn
        assert np.array_equal(tb, rtb)


def test_active_block_ties_follow_heap_order():
    # a periodic image gives many blocks of identical variance: the rank order is then set by
    # the heap sort alone, and every rank 0..n-1 is used exactly once
    g = Geom.identity((32, 32, 32))
    x = np.arange(32) % 4
    img = np.broadcast_to((x[None, None, :] * 1.0 + (np.arange(32)[:, None, None] // 16)), (32, 32, 32))
    img = np.ascontiguousarray(img, np.float32).ravel()
    tb, n = O.orc_active_blocks(g, img, 50)
    assert n == 256
    assert sorted(tb[tb >= 0]) == list(range(n))
    if O.ref_available():
        rtb, rn, _, _, _ = O.ref_block_matching(g, img, None, 50)
        assert np.array_equal(tb, rtb)


@needs_ref
def test_block_matching_matches_reference(aladin_case):
    g, img, mask, warped = aladin_case
    tb, n = O.orc_active_blocks(g, img, 50, mask)
    rp, wp, d = O.orc_block_matching(g, img, warped, tb, n, mask)
    _, rn, rrp, rwp, rd = O.ref_block_matching(g, img, warped, 50, mask)
    assert n == rn and d == rd
    assert np.array_equal(rp, rrp)
    assert np.array_equal(wp, rwp, equal_nan=True)


def test_block_matching_recovers_integer_shift():
    g = Geom.identity((40, 40, 40))
    img = phantoms.blob_phantom(g.dims, 5)
    shift = (2, -1, 3)  # voxels (x, y, z)
    warped = np.full_like(img, np.nan)
    warped[3:, :39, 2:] = img[:-3, 1:, :-2]  # warped(p) = img(p - shift)
    tb, n = O.orc_active_blocks(g, img.ravel(), 50)
    rp, wp, d = O.orc_block_matching(g, img.ravel(), warped.ravel(), tb, n)
    disp = (wp - rp).reshape(-1, 3)
    ok = np.isfinite(disp[:, 0])
    assert d == ok.sum() and d > 0.9 * n
    # blocks whose search window stays inside the valid region find the shift exactly
    corner = rp.reshape(-1, 3)
    inner = ok & np.all(corner >= 8, axis=1) & np.all(corner <= 28, axis=1)
    assert inner.sum() > 20
    assert np.array_equal(disp[inner], np.tile(np.array(shift, np.float32), (inner.sum(), 1)))


def _points(n, seed, A, outliers=0.3):
    r = np.random.default_rng(seed)
    p = r.uniform(-60, 60, (n, 3)).astype(np.float32)
    q = (p.astype(np.float64) @ A[:3, :3].T + A[:3, 3]).astype(np.float32)
    k = int(outliers * n)
    q[:k] += r.uniform(-25, 25, (k, 3)).astype(np.float32)
    return p.ravel(), q.ravel()


def test_lts_recovers_affine_with_outliers():
    A = phantoms.random_affine(4).astype(np.float64)
    rp, wp = _points(2000, 0, A, outliers=0.3)
    M = O.orc_lts(rp, wp, 50, True)
    assert np.allclose(M, A, atol=2e-4)


def test_lts_recovers_rigid_with_outliers():
    A = phantoms.random_affine(5, scale=(1.0, 1.0), shear=0.0).astype(np.float64)
    rp, wp = _points(1500, 1, A, outliers=0.25)
    M = O.orc_lts(rp, wp, 50, False)
    assert np.allclose(M, A, atol=2e-4)
    R = M[:3, :3].astype(np.float64)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-5) and np.linalg.det(R) > 0


@needs_ref
@pytest.mark.parametrize("affine", [True, False])
def test_lts_matches_reference(affine):
    A = phantoms.random_affine(6, scale=(0.95, 1.05) if affine else (1.0, 1.0), shear=0.03 if affine else 0.0)
    rp, wp = _points(3000, 2, A, outliers=0.4)
    wp = wp.copy()
    wp[30:60:3] = np.nan  # undefined correspondences are skipped
    start = np.eye(4, dtype=np.float32)
    start[:3, 3] = [0.5, -0.25, 1.0]
    M = O.orc_lts(rp, wp, 50, affine, start)
    R = O.ref_lts(rp, wp, int(np.isfinite(wp.reshape(-1, 3)[:, 0]).sum()), 50, affine, start)
    assert np.allclose(M, R, rtol=1e-5, atol=1e-5)


@needs_ref
def test_lts_on_block_matching_output_matches_reference(aladin_case):
    g, img, mask, warped = aladin_case
    tb, n, rp, wp, d = O.ref_block_matching(g, img, warped, 50, mask)
    M = O.orc_lts(rp, wp, 50, True)
    R = O.ref_lts(rp, wp, d, 50, True)
    assert np.allclose(M, R, rtol=1e-5, atol=1e-5)
