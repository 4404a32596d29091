"""Aladin at the batched configuration's size (128^3 blob phantoms under random affines) and
batched Aladin registrations (nrh_aladin_register_batch, non-symmetric, affine).

At 128^3: the first iteration's warp and block-matching correspondences equal the reference
build's bit for bit, and a registration ends within the stated bounds of the reference's run.
Batched: every pair's final matrix equals its own single registration's bit for bit (the pairs
share launches, not arithmetic), and a pair that cannot be registered (a flat image: no active
block) fails alone with NRH_ERR_NUMERIC while the others finish.
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import _lib, aladin, phantoms
from niftyreg_amd.aladin import AladinContent
from niftyreg_amd.geometry import Geom

pytestmark = pytest.mark.gpu

DIMS = (128, 128, 128)


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
def test_first_iteration_matches_reference_at_128(gpu):
    ref, flo, A = phantoms.config3_pair(0, DIMS)
    g = Geom.identity(DIMS)
    ref, flo = ref.ravel(), flo.ravel()
    h = AladinContent(g, ref, g, flo)
    try:
        eye = np.eye(4, dtype=np.float32)
        h.set_matrix(eye)
        h.GetWarpedImage()
        warped = h.warped()
        h.BlockMatching()
        rp, wp, d = h.positions()
        n_active = h.active_blocks
    finally:
        h.close()
    field = O.ref_affine_field(g, eye)
    rw = O.ref_resample(g, flo, g, field)
    assert np.array_equal(warped, rw, equal_nan=True)
    tb, rn, rrp, rwp, rd = O.ref_block_matching(g, ref, rw, 50)
    assert rn == n_active and rd == d
    assert np.array_equal(rp, rrp)
    assert np.array_equal(wp, rwp, equal_nan=True)


def _pairs(seeds, dims=DIMS):
    out = []
    for s in seeds:
        ref, flo, _ = phantoms.config3_pair(s, dims)
        g = Geom.identity(dims)
        out.append((g, ref.ravel(), g, flo.ravel()))
    return out


def test_batch_matches_single_registrations(gpu):
    pairs = _pairs([0, 1, 2, 3])
    Ms, status, stats = aladin.register_batch(pairs, levels=3, maxit=5, rigid=False, affine=True)
    print(f"batched Aladin: {stats}")
    assert status == [_lib.NRH_OK] * 4
    for q, (g, ref, fg, flo) in enumerate(pairs):
        M1 = aladin.register(g, ref, fg, flo, levels=3, maxit=5, rigid=False, affine=True, symmetric=False)
        assert np.array_equal(Ms[q], M1), q
    if O.ref_available():
        g, ref, fg, flo = pairs[0]
        R = O.ref_aladin_run(g, ref, fg, flo, levels=3, maxit=5, rigid=False, affine=True)
        print("batched\n", Ms[0], "\nreference\n", R)
        assert np.abs(Ms[0][:3, :3] - R[:3, :3]).max() < 5e-3
        assert np.abs(Ms[0][:3, 3] - R[:3, 3]).max() < 0.1


def test_batch_isolates_a_failing_pair(gpu):
    dims = (64, 64, 64)
    pairs = _pairs([5, 6], dims)
    g = Geom.identity(dims)
    flat = np.full(g.nvox, 0.5, np.float32)
    pairs.insert(1, (g, flat, g, pairs[0][3]))
    Ms, status, _ = aladin.register_batch(pairs, levels=2, maxit=4, rigid=False, affine=True)
    assert status[1] == _lib.NRH_ERR_NUMERIC
    assert status[0] == _lib.NRH_OK and status[2] == _lib.NRH_OK
    for q in (0, 2):
        g0, ref, fg, flo = pairs[q]
        M1 = aladin.register(g0, ref, fg, flo, levels=2, maxit=4, rigid=False, affine=True, symmetric=False)
        assert np.array_equal(Ms[q], M1), q
