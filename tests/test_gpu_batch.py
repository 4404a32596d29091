"""Batches of F3D pairs of different sizes (nrh_batch_*, nrh_f3d_register_batch).

A batch runs every kernel of an evaluation once for all pairs (persistent sweeps over a (pair,
tile) work list, item tables for the small kernels).  Each pair's arithmetic and reduction order
are the pair's own, so the batch must give every pair exactly the objective, node gradient and
max length of the pair evaluated alone, with float32 and float16 image storage, for any
persistent grid size; and a batched registration must take every pair through exactly the
decisions of its own registration (same final grid, NMI and iteration count).
"""
import numpy as np
import pytest

from niftyreg_amd import _lib, phantoms
from niftyreg_amd.f3d import F3dBatch, F3dContent, default_params, register, register_batch
from niftyreg_amd.geometry import Geom, create_control_point_grid

pytestmark = pytest.mark.gpu

# eight uneven pairs: every axis differs, tiles are ragged on every axis
SIZES = [(64, 48, 40), (72, 60, 44), (56, 64, 52), (80, 52, 48), (48, 48, 48), (68, 76, 40), (60, 56, 64),
         (88, 44, 36)]


def make_pair(dims, seed, storage=0):
    g = Geom.identity(dims)
    ref = phantoms.blob_phantom(dims, seed).ravel()
    flo = phantoms.bspline_warp(phantoms.blob_phantom(dims, seed).reshape(dims[::-1]), 5, 1.5, seed + 100).ravel()
    cg, cpp = create_control_point_grid(g, [5.0] * 3)
    cpp = (cpp + np.random.default_rng(seed).uniform(-1.5, 1.5, cpp.size)).astype(np.float32)
    h = F3dContent(g, ref, g, flo, cg, cpp, params=default_params(similarity_weight=0.999, be_weight=0.001,
                                                                  image_storage=storage))
    h.rescale_intensities((2, 65), (2, 65))
    return h, cpp


def alone(h, cpp):
    h.set_control_points(cpp)
    value, terms = h.objective()
    maxlen = h.gradient(normalise=True)
    grad = h.node_gradient()
    h.set_control_points(cpp)  # the bending-energy gradient rounds the points through displacements
    return value, terms[0] / np.float32(0.999), maxlen, grad


@pytest.mark.parametrize("storage", [0, 1])
def test_batch_matches_pairs_alone(gpu, storage):
    pairs = [make_pair(d, 10 + i, storage) for i, d in enumerate(SIZES)]
    try:
        want = [alone(h, c) for h, c in pairs]
        b = F3dBatch([h for h, _ in pairs])
        try:
            for blocks in (None, 3, 64):
                if blocks is not None:
                    b.set_sweep_blocks(blocks)
                for h, c in pairs:
                    h.set_control_points(c)
                values, nmi = b.objective()
                maxlens = b.gradient(normalise=True)
                for q, (h, c) in enumerate(pairs):
                    assert values[q] == want[q][0], (blocks, q)
                    assert maxlens[q] == want[q][2], (blocks, q)
                    assert np.array_equal(h.node_gradient(), want[q][3]), (blocks, q)
            k, ms = b.autotune()
            print(f"storage {storage}: autotuned {k} workgroups per XCD, histogram sweep ms {ms}")
            assert k in (32, 64, 128, 256)
        finally:
            b.close()
        # the handles work alone again after the batch
        h, c = pairs[0]
        assert alone(h, c)[0] == want[0][0]
    finally:
        for h, _ in pairs:
            h.close()


def test_half_storage_against_float(gpu):
    """float16 images, float32 arithmetic: the objective and gradient move by the images' rounding
    (half an ulp of 64 is 0.016 of a bin), reported here against float32 storage."""
    errs = []
    for i, d in enumerate(SIZES[:4]):
        h32, c = make_pair(d, 10 + i, 0)
        h16, _ = make_pair(d, 10 + i, 1)
        try:
            v32, n32, m32, g32 = alone(h32, c)
            v16, n16, m16, g16 = alone(h16, c)
        finally:
            h32.close()
            h16.close()
        rel = abs(n16 - n32) / n32
        gl2 = float(np.linalg.norm(g16 - g32) / np.linalg.norm(g32))
        errs.append((rel, gl2))
        print(f"{d}: NMI float32 {n32:.6f} float16 {n16:.6f} (rel {rel:.1e}), normalised gradient L2 {gl2:.1e}")
    assert max(e[0] for e in errs) <= 2e-3
    assert max(e[1] for e in errs) <= 0.1


def test_batch_rejects_unsupported(gpu):
    h, c = make_pair(SIZES[0], 1)
    h2, _ = make_pair(SIZES[1], 2)
    try:
        with pytest.raises(_lib.NiftyRegError):
            F3dBatch([h, h])  # the same handle twice
        h2.set_params(default_params(similarity_weight=0.989, be_weight=0.001, le_weight=0.01))
        with pytest.raises(_lib.NiftyRegError):
            F3dBatch([h, h2])  # linear energy is not batched
        # a failed batch leaves the handles usable on their own streams
        assert np.isfinite(alone(h, c)[0])
    finally:
        h.close()
        h2.close()


@pytest.mark.parametrize("storage", [0, 1])
def test_batched_registration_matches_single(gpu, storage):
    """nrh_f3d_register_batch against nrh_f3d_register on each pair: two levels, 20 iterations."""
    pairs = []
    for i, d in enumerate([(64, 56, 48), (48, 60, 52), (56, 48, 64), (60, 64, 40)]):
        g = Geom.identity(d)
        ref = phantoms.blob_phantom(d, 30 + i)
        flo = phantoms.bspline_warp(ref.reshape(d[::-1]), 5, 2.0, 60 + i)
        pairs.append((g, ref.ravel(), g, flo.ravel()))
    res, stats = register_batch(pairs, levels=2, maxit=20, image_storage=storage)
    print(f"batched registration: {stats}")
    for q, (g, ref, fg, flo) in enumerate(pairs):
        cg, cpp, nmi, its = register(g, ref, fg, flo, levels=2, maxit=20, image_storage=storage)
        bg, bcpp, bnmi, bits = res[q]
        assert bg.dims == cg.dims
        assert its == bits, (q, its, bits)
        assert nmi == bnmi, (q, nmi, bnmi)
        assert np.array_equal(cpp, bcpp), q
    assert stats["graph_replays"] >= 1 and all(k in (32, 64, 128, 256) for k in stats["blocks_per_xcd"])
