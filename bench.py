#!/usr/bin/env python3
"""F3D objective+gradient evaluations per second on one MI355X.

Workload: the 256 x 256 x 160 float32 shell-phantom pair (1 mm, identity orientation), NMI
with 64 bins (68 internal), bending energy 0.001, control-point spacing 5 voxels (grid
55 x 55 x 35), linear interpolation, NaN padding.  One step is one reg_f3d objective value
plus one objective gradient at the same control points (the unit "obj+grad eval"), run
through libnrhip's fused sweeps with the images and control points resident in HBM.

Prints one JSON line.  The roofline entry covers the dominant kernel (the gradient sweep,
k_sweep<float, true>), timed with HIP events on the stream it runs on; its algorithmic bytes are
the reference and floating images read once (8 B per voxel), its measured traffic also counts
the voxel-gradient image it writes for the projection passes.  The CPU baseline times the
reference CPU path (oracle/_ref, OpenMP) on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
DIMS = (256, 256, 160)
BINS = 64
BE = 0.001
SWEEP_TIMING_PERIOD = 4  # HIP-event timing of every 4th fused sweep in the timed region


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--storage", choices=["f32", "f16"], default="f32")
    p.add_argument("--no-secondary", action="store_true", help="skip the Aladin, LNCC and F3D2 lines")
    p.add_argument("--only-f3d2", action="store_true",
                   help="run only the symmetric (reg_f3d2) line and print it (for profiling)")
    p.add_argument("--only", choices=["batch-f3d", "batch-aladin", "serve", "aladin", "lncc"],
                   help="run only one of the batched / serving / secondary lines and print it")
    p.add_argument("--no-batch", action="store_true", help="skip the batched F3D, batched Aladin and serving lines")
    return p.parse_args()


def cpu_baseline(ref, flo, geom):
    """reg_f3d<float> objective + gradient from the reference CPU build, timed on the host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracles
    if not oracles.ref_available():
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    r = oracles.RefF3d(geom, ref, geom, flo, bins=BINS, be=BE, le=0.0, spacing=-5.0, levels=1)
    r.value()
    r.gradient()
    t0 = time.perf_counter()
    n = 0
    while True:
        r.value()
        r.gradient()
        n += 1
        if time.perf_counter() - t0 > 10.0 or n >= 5:
            break
    dt = time.perf_counter() - t0
    r.close()
    return {"value": n / dt, "unit": "evals/s", "cores": threads, "kind": "reference",
            "sample": f"{n} obj+grad evals of reg_f3d<float> at 256x256x160 (one level, -be 0.001 -le 0, 64 bins), "
                      f"OMP_NUM_THREADS={threads}, after one warm-up eval"}


def _sync_time(torch, fn, n):
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / n


def bench_aladin(torch, steps, with_cpu):
    """One Aladin iteration (warp, block matching, LTS) on the 128^3 blob pair with a random
    affine (50% of the 32^3 blocks, capture range 3, 50% LTS inliers), inputs resident in HBM."""
    from niftyreg_amd import phantoms
    from niftyreg_amd.aladin import AladinContent
    from niftyreg_amd.geometry import Geom
    dims = (128, 128, 128)
    ref, flo, _ = phantoms.config3_pair(0, dims)
    g = Geom.identity(dims)
    h = AladinContent(g, ref, g, flo)
    eye = np.eye(4, dtype=np.float32)
    h.set_matrix(eye)
    for _ in range(3):
        h.GetWarpedImage()
        h.BlockMatching()
        h.Lts(True)
    h.set_matrix(eye)
    warp_s = _sync_time(torch, h.GetWarpedImage, steps)
    bm_s = _sync_time(torch, h.BlockMatching, steps)
    lts_s = _sync_time(torch, lambda: (h.set_matrix(eye), h.Lts(True)), steps)
    it_s = warp_s + bm_s + lts_s
    N = g.nvox
    active = h.active_blocks
    flops = active * 343 * 64 * 11.0  # pass 1: 3 adds; pass 2: 2 sub, 3 mul, 3 add (per offset and voxel)
    out = {
        "metric": "Aladin iterations/s at 128^3 (warp + block matching + LTS)",
        "value": 1.0 / it_s, "unit": "iterations/s",
        "ms": {"affine_warp": warp_s * 1e3, "block_matching": bm_s * 1e3, "lts": lts_s * 1e3},
        "config": {"workload": "128^3 float32 blob pair, random affine, 4^3 blocks, 50% kept, capture 3, LTS 50%",
                   "active_blocks": active},
        "roofline": {"bound": "valu", "kernel": "k_block_match", "achieved": flops / bm_s / 1e12,
                     "peak": 157.3, "unit": "TFLOP/s", "frac": flops / bm_s / 1e12 / 157.3,
                     "hbm_GBs": 8.0 * N / bm_s / 1e9},
        "cpu_baseline": None,
    }
    h.close()
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracles
        if oracles.ref_available():
            A = eye.copy()
            t0 = time.perf_counter()
            n = 0
            while n < 3:
                field = oracles.ref_affine_field(g, A)
                w = oracles.ref_resample(g, flo.ravel(), g, field)
                _, _, rp, wp, d = oracles.ref_block_matching(g, ref.ravel(), w, 50)
                oracles.ref_lts(rp, wp, d, 50, True, A)
                n += 1
            dt = time.perf_counter() - t0
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            out["cpu_baseline"] = {"value": n / dt, "unit": "iterations/s", "cores": threads, "kind": "reference",
                                   "sample": f"{n} iterations of reg_affine_getDeformationField + reg_resampleImage + "
                                             "initialise_block_matching_method + block_matching_method + optimize "
                                             f"(reference CPU build) at 128^3, OMP_NUM_THREADS={threads}"}
    return out


def bench_lncc(torch, ref, flo, geom, cppg, cpp, steps, with_cpu):
    """F3D with LNCC (Gaussian sigma 5 voxels): one objective value + one gradient at 256x256x160
    through the per-operation chain."""
    from niftyreg_amd import _lib
    from niftyreg_amd.f3d import F3dContent, default_params
    h = F3dContent(geom, ref, geom, flo, cppg, cpp,
                   params=default_params(similarity_weight=1.0 - BE, be_weight=BE, measure=_lib.MEASURE_LNCC,
                                         lncc_sigma=-5.0))
    h.rescale_intensities((0, 1), (0, 1))
    h.objective()
    h.gradient()
    step_s = _sync_time(torch, lambda: (h.objective(), h.gradient()), steps)
    value, terms = h.objective()
    h.close()
    N = geom.nvox
    # algorithmic bytes: the reference and floating images read once per sweep, two sweeps per
    # obj+grad eval (8 B per voxel each); the design's traffic (the two statistics sets, 6 + 4
    # planar arrays through 3 axis passes, read + write) is reported beside it
    algo_bytes = 2 * 8 * N
    conv_bytes = (6 + 6 + 4) * 2 * 4 * 3 * N
    out = {
        "metric": "F3D LNCC obj+grad evals/s at 256x256x160", "value": 1.0 / step_s, "unit": "evals/s",
        "ms_per_step": step_s * 1e3, "lncc": terms[0] / (1.0 - BE),
        "config": {"workload": "F3D LNCC (Gaussian sigma 5 voxels) objective+gradient, 256x256x160 shell pair, "
                               "BE 0.001, 5-voxel grid"},
        "roofline": {"bound": "hbm", "kernel": "whole eval (algorithmic: images read by two sweeps)",
                     "achieved": algo_bytes / step_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": algo_bytes / step_s / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_eval": algo_bytes,
                     "design_bytes_per_eval": conv_bytes,
                     "design_frac": conv_bytes / step_s / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": None,
    }
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracles
        if oracles.ref_available():
            r = oracles.RefF3d(geom, ref, geom, flo, be=BE, le=0.0, spacing=-5.0, levels=1, measure=1, lncc_sigma=-5.0)
            r.set_cpp(cpp)
            t0 = time.perf_counter()
            r.value()
            r.gradient()
            dt = time.perf_counter() - t0
            r.close()
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "evals/s", "cores": threads, "kind": "reference",
                                   "sample": f"1 LNCC obj+grad eval of reg_f3d<float> at 256x256x160, "
                                             f"OMP_NUM_THREADS={threads}"}
    return out


def bench_f3d2(torch, ref, flo, geom, steps, with_cpu):
    """Symmetric diffeomorphic F3D (reg_f3d2): one objective value + one gradient of the forward /
    backward pair at 256x256x160 (velocity grids exponentiated by scaling and squaring, six
    squarings, both directions warped, symmetric NMI, gradients exponentiated along the opposite
    grid), inputs resident in HBM."""
    from niftyreg_amd.f3d import F3d2Pair, F3dContent, default_params
    from niftyreg_amd.geometry import create_symmetric_control_point_grids
    cg, fw0, bw0 = create_symmetric_control_point_grids(geom, geom, [5.0, 5.0, 5.0])
    rng = np.random.default_rng(11)
    vfw = (fw0 + rng.uniform(-2.0, 2.0, fw0.size)).astype(np.float32)
    vbw = (bw0 + rng.uniform(-2.0, 2.0, bw0.size)).astype(np.float32)
    params = lambda: default_params(bins=BINS, similarity_weight=1.0 - BE, be_weight=BE)
    hf = F3dContent(geom, ref, geom, flo, cg, vfw, params=params())
    hb = F3dContent(geom, flo, geom, ref, cg, vbw, params=params())
    hf.rescale_intensities((2, BINS + 1), (2, BINS + 1))
    hb.rescale_intensities((2, BINS + 1), (2, BINS + 1))
    pair = F3d2Pair(hf, hb, steps=6)
    for _ in range(3):
        pair.objective()
        pair.gradient()
    step_s = _sync_time(torch, lambda: (pair.objective(), pair.gradient()), steps)
    value, terms = pair.objective()
    squarings = hf.velocity_steps
    pair.close()
    N = geom.nvox
    # HBM bytes of one eval at compulsory traffic, per direction, for the passes that run (the
    # gradient finds the objective's exponentiation, warped image and histogram in place, and on
    # this shared geometry each gradient exponentiation reuses the opposite direction's squaring
    # chain): the flow (12 B per voxel written; the forward one, whose step number is chosen from
    # its extremes, also rescaled: +24) and `squarings` compositions (12 B read, 12 B written);
    # the warp (field 12 + floating 4 read, 4 written) and histogram (8); the image gradient
    # (field 12 + floating 4 read, 12 written); the voxel-gradient fill (12) and NMI voxel
    # gradient (4 + 4 + 12 + 12 read, 12 written); `squarings` gradient resamplings (field and
    # gradient read, gradient written: 36); the separable node projection's x pass (12 read).
    s = squarings
    per_dir = 12 + 24 * s + 20 + 8 + 28 + 12 + 44 + 36 * s + 12
    eval_bytes = (2 * per_dir + 24) * N
    # algorithmic bytes: both images read by the objective and the gradient sweep of each direction
    algo_bytes = 2 * 2 * 8 * N
    out = {
        "metric": "F3D2 (symmetric, velocity) obj+grad evals/s at 256x256x160", "value": 1.0 / step_s,
        "unit": "evals/s", "ms_per_step": step_s * 1e3, "squarings": squarings, "nmi_sum": terms[0] / (1.0 - BE),
        "config": {"workload": "reg_f3d2 NMI objective+gradient eval of the forward/backward pair, 256x256x160 "
                               "shell pair, 64 bins, BE 0.001, 5-voxel symmetric velocity grids 55x55x35"},
        "roofline": {"bound": "hbm", "kernel": "whole eval (algorithmic: images read by four sweeps)",
                     "achieved": algo_bytes / step_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": algo_bytes / step_s / 1e9 / HBM_PEAK_GBS, "algorithmic_bytes_per_eval": algo_bytes,
                     "design_bytes_per_eval": eval_bytes,
                     "design_frac": eval_bytes / step_s / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": None,
    }
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracles
        if oracles.ref_available():
            r = oracles.RefF3d2(geom, ref, geom, flo, bins=BINS, be=BE, le=0.0, spacing=-5.0, levels=1, scalar=False)
            r.set_cpp(vfw, vbw)
            t0 = time.perf_counter()
            r.value()
            r.gradient()
            dt = time.perf_counter() - t0
            r.close()
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            out["cpu_baseline"] = {"value": 1.0 / dt, "unit": "evals/s", "cores": threads, "kind": "reference",
                                   "sample": f"1 obj+grad eval of reg_f3d2<float> at 256x256x160 (same grids), "
                                             f"OMP_NUM_THREADS={threads}"}
    return out


def _gen_config2(seed):
    from niftyreg_amd import phantoms
    ref, flo = phantoms.config2_pair(seed)
    return ref, flo


def _gen_config3(seed):
    from niftyreg_amd import phantoms
    ref, flo, _ = phantoms.config3_pair(seed, (128, 128, 128))
    return ref, flo


def _gen_config4(seed):
    from niftyreg_amd import phantoms
    return phantoms.config4_request(seed)


def _generate(fn, seeds):
    """Synthetic inputs on host worker processes (at most 16, the box's CPU share)."""
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=min(16, len(seeds), os.cpu_count() or 1)) as ex:
        return list(ex.map(fn, seeds))


CONFIG2_SEEDS = list(range(1000, 1064))
CONFIG3_SEEDS = list(range(2000, 2128))
CONFIG4_SEEDS = list(range(3000, 3012))


def bench_batch_f3d(torch, imgs, with_cpu, maxit=200):
    """Batched F3D: 64 pairs with axes drawn from U{96..224}, shell phantoms under B-spline warps,
    float16 image storage (float32 arithmetic), NMI 64 bins, BE 0.001, LE 0, 5-voxel grid, three
    levels, 200 iterations; nrh_f3d_register_batch against the same registrations one at a time
    (nrh_f3d_register) and the reference CPU build on one pair of median size."""
    from niftyreg_amd.f3d import register, register_batch
    from niftyreg_amd.geometry import Geom
    from niftyreg_amd import phantoms
    seeds = CONFIG2_SEEDS
    n_pairs = len(seeds)
    pairs = []
    for s, (ref, flo) in zip(seeds, imgs):
        g = Geom.identity(phantoms.config2_dims(s))
        pairs.append((g, ref.ravel(), g, flo.ravel()))
    voxels = [p[0].nvox for p in pairs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res, stats = register_batch(pairs, levels=3, maxit=maxit, image_storage=1)
    t_batch = time.perf_counter() - t0
    # the same registrations one at a time (a sample of 16 pairs: the rate per pair is the same
    # registration; every pair's result is the batch's, bit for bit)
    sample = list(range(0, n_pairs, max(1, n_pairs // 16)))
    t0 = time.perf_counter()
    matched = 0
    for q in sample:
        g, ref, fg, flo = pairs[q]
        cg, cpp, nmi, its = register(g, ref, fg, flo, levels=3, maxit=maxit, image_storage=1)
        matched += int(nmi == res[q][2] and its == res[q][3] and np.array_equal(cpp, res[q][1]))
    t_single = (time.perf_counter() - t0) / len(sample) * n_pairs
    out = {
        "metric": "batched F3D registrations/hour (64 pairs, axes U{96..224})",
        "value": n_pairs / t_batch * 3600.0, "unit": "registrations/hour",
        "seconds": t_batch, "device_seconds": stats["device_seconds"], "graph_replays": stats["graph_replays"],
        "one_at_a_time_gpu": {"value": n_pairs / t_single * 3600.0, "unit": "registrations/hour",
                              "sample": f"{len(sample)} of the {n_pairs} registrations one after another "
                                        "(nrh_f3d_register), scaled to all pairs",
                              "identical_results": f"{matched}/{len(sample)}"},
        "autotuned_blocks_per_xcd": stats["blocks_per_xcd"], "tile": stats["tile"],
        "config": {"workload": f"{n_pairs} F3D registrations, shell phantom pairs with axes U{{96..224}}, float16 "
                               f"storage, NMI 64 bins, BE 0.001, LE 0, 5-voxel grid, 3 levels, {maxit} iterations",
                   "total_voxels": int(sum(voxels)), "mean_final_nmi": float(np.mean([r[2] for r in res]))},
        "cpu_baseline": None,
    }
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracles
        if oracles.ref_available():
            q = int(np.argsort(voxels)[len(voxels) // 2])
            g, ref, fg, flo = pairs[q]
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            t0 = time.perf_counter()
            rcg, rcpp = oracles.ref_f3d_run(g, ref, fg, flo, levels=3, maxit=maxit)
            dt = time.perf_counter() - t0
            nmi_ref = oracles.ref_nmi_at(g, ref, fg, flo, rcpp, levels=3)
            out["cpu_baseline"] = {"value": 3600.0 / dt, "unit": "registrations/hour", "cores": threads,
                                   "kind": "reference",
                                   "sample": f"1 reg_f3d<float> registration of the median pair {g.dims} "
                                             f"(3 levels, {maxit} iterations, float32), OMP_NUM_THREADS={threads}",
                                   "final_nmi": nmi_ref, "device_final_nmi": res[q][2]}
    return out


def bench_batch_aladin(torch, imgs, with_cpu):
    """Batched Aladin: 128 pairs of 128^3 blob phantoms under random affines, affine-only and
    non-symmetric (-noSym -affDirect), 4^3 blocks, 50% blocks, 50% LTS inliers, 3 levels, 5
    iterations; nrh_aladin_register_batch against one-at-a-time nrh_aladin_register and the
    reference CPU build on one pair."""
    from niftyreg_amd import aladin
    from niftyreg_amd.geometry import Geom
    n_pairs = len(imgs)
    g = Geom.identity((128, 128, 128))
    pairs = [(g, r.ravel(), g, f.ravel()) for r, f in imgs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Ms, status, stats = aladin.register_batch(pairs, levels=3, maxit=5, rigid=False, affine=True)
    t_batch = time.perf_counter() - t0
    sample = list(range(0, n_pairs, max(1, n_pairs // 16)))
    t0 = time.perf_counter()
    same = 0
    for q in sample:
        M1 = aladin.register(*pairs[q], levels=3, maxit=5, rigid=False, affine=True, symmetric=False)
        same += int(np.array_equal(M1, Ms[q]))
    t_single = (time.perf_counter() - t0) / len(sample) * n_pairs
    out = {
        "metric": "batched Aladin registrations/hour (128 pairs at 128^3)",
        "value": n_pairs / t_batch * 3600.0, "unit": "registrations/hour", "seconds": t_batch,
        "device_seconds": stats["device_seconds"], "failed": int(sum(1 for s_ in status if s_ != 0)),
        "one_at_a_time_gpu": {"value": n_pairs / t_single * 3600.0, "unit": "registrations/hour",
                              "sample": f"{len(sample)} of the {n_pairs} registrations one after another, scaled",
                              "identical_results": f"{same}/{len(sample)}"},
        "config": {"workload": f"{n_pairs} Aladin affine registrations (non-symmetric, affine only), 128^3 blob "
                               "phantoms under random affines, 4^3 blocks, 50% blocks, 50% LTS inliers, 3 levels, "
                               "5 iterations"},
        "cpu_baseline": None,
    }
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import oracles
        if oracles.ref_available():
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            t0 = time.perf_counter()
            oracles.ref_aladin_run(*pairs[0], levels=3, maxit=5, rigid=False, affine=True)
            dt = time.perf_counter() - t0
            out["cpu_baseline"] = {"value": 3600.0 / dt, "unit": "registrations/hour", "cores": threads,
                                   "kind": "reference",
                                   "sample": f"1 reg_aladin<float> registration of one 128^3 pair (same settings), "
                                             f"OMP_NUM_THREADS={threads}"}
    return out


def bench_serving(torch, reqs, rate_hz=2.0, maxit=100):
    """Serving: Poisson arrivals at rate_hz; each request is an Aladin affine (affine-only, 3
    levels, 5 iterations) followed by F3D (3 levels, `maxit` iterations) from it, on a cubic volume
    log-uniform between 64^3 and 256^3 with NMI (64 bins) or LNCC (sigma 5 voxels) at random;
    admission by a 64 GB device-memory budget, at most 8 requests at once.  Latency is from
    arrival to result; the one-at-a-time form runs the same requests back to back."""
    from niftyreg_amd import _lib
    from niftyreg_amd.geometry import Geom
    from niftyreg_amd.serve import Server, request_config
    n_requests = len(reqs)
    arrivals = np.cumsum(np.random.default_rng(7).exponential(1.0 / rate_hz, n_requests))
    arrivals -= arrivals[0]

    def submit_all(srv, paced):
        ids = []
        t0 = time.perf_counter()
        for k, (dims, ref, flo, measure) in enumerate(reqs):
            if paced:
                delay = arrivals[k] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
            g = Geom.identity(dims)
            m = _lib.MEASURE_LNCC if measure else _lib.MEASURE_NMI
            ids.append(srv.submit(g, ref, g, flo, request_config(levels=3, maxit=maxit, measure=m)))
        return ids

    srv = Server(budget_bytes=64 << 30, max_active=8)
    t0 = time.perf_counter()
    ids = submit_all(srv, True)
    res = [srv.wait(i) for i in ids]
    t_mix = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array([r["finished"] - r["submitted"] for r in res])
    srv = Server(budget_bytes=64 << 30, max_active=1)
    t0 = time.perf_counter()
    ids = submit_all(srv, False)
    seq = [srv.wait(i) for i in ids]
    t_seq = time.perf_counter() - t0
    srv.close()
    same = sum(int(a["status"] == b["status"] and a.get("similarity") == b.get("similarity")) for a, b in zip(res, seq))
    return {
        "metric": "serving registrations/hour and latency (Poisson arrivals)",
        "value": n_requests / t_mix * 3600.0, "unit": "registrations/hour",
        "p50_latency_s": float(np.percentile(lat, 50)), "p99_latency_s": float(np.percentile(lat, 99)),
        "failed": int(sum(1 for r in res if r["status"] != 0)), "peak_concurrent": st["peak_active"],
        "peak_budgeted_bytes": st["peak_bytes"],
        "one_at_a_time_gpu": {"value": n_requests / t_seq * 3600.0, "unit": "registrations/hour",
                              "p50_latency_s_if_queued": None, "identical_results": f"{same}/{n_requests}"},
        "config": {"workload": f"{n_requests} requests, Poisson arrivals at {rate_hz}/s, cubic volumes log-uniform "
                               f"64^3..256^3 (sides {[r[0][0] for r in reqs]}), NMI/LNCC at random "
                               f"({sum(r[3] for r in reqs)} LNCC), Aladin affine (3 levels, 5 it) then F3D (3 levels, "
                               f"{maxit} it), 64 GB budget, at most 8 concurrent"},
    }


def measured_traffic():
    """HBM bytes per gradient-sweep launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (scaled by the copy-kernel calibration recorded there), or None."""
    path = os.path.join(ROOT, "profiles", "f3d_eval_256x256x160_hbm_traffic.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return int(line.split(":")[1].strip())
    except OSError:
        pass
    return None


VALU_PEAK_INSTS = 256 * 4 * 2.4e9 / 4  # wave64 VALU instructions per second: 256 CUs x 4 SIMDs, 4 cycles each


def measured_valu():
    """VALU instructions per gradient-sweep launch from the committed rocprofv3 SQ_INSTS_VALU
    summary, or None."""
    path = os.path.join(ROOT, "profiles", "f3d_sweeps_256x256x160_counters.md")
    try:
        for line in open(path):
            if line.startswith("valu_insts_per_launch"):
                return float(line.split(":")[1].strip())
    except OSError:
        pass
    return None


def main():
    a = parse()
    # stdout carries the one JSON line: whatever else the libraries print there (the reference
    # build's warnings in the CPU baselines) goes to stderr
    json_out = os.fdopen(os.dup(1), "w")
    sys.stdout.flush()
    os.dup2(2, 1)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}): every GPU box here has a single MI355X",
              file=sys.stderr)
        sys.exit(2)
    # synthetic inputs of the batched and serving lines, made on host worker processes before the
    # GPU is touched
    data = {}
    want = [a.only] if a.only else ([] if (a.no_batch or a.no_secondary or a.only_f3d2) else
                                    ["batch-f3d", "batch-aladin", "serve"])
    if "batch-f3d" in want:
        data["batch-f3d"] = _generate(_gen_config2, CONFIG2_SEEDS)
    if "batch-aladin" in want:
        data["batch-aladin"] = _generate(_gen_config3, CONFIG3_SEEDS)
    if "serve" in want:
        data["serve"] = _generate(_gen_config4, CONFIG4_SEEDS)
    import torch
    from niftyreg_amd import phantoms
    from niftyreg_amd.f3d import F3dContent, default_params
    from niftyreg_amd.geometry import Geom, create_control_point_grid
    from niftyreg_amd import _lib

    torch.cuda.init()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    _lib.call("nrh_init", 0)

    geom = Geom.identity(DIMS)
    ref, flo = phantoms.config1_pair(seed=0, dims=DIMS)
    if a.only_f3d2:
        print(json.dumps(bench_f3d2(torch, ref, flo, geom, a.steps, not a.no_cpu_baseline)), file=json_out, flush=True)
        return
    if a.only:
        def lncc_line():
            cg, cp = create_control_point_grid(geom, [5.0, 5.0, 5.0])
            cp = (cp + np.random.default_rng(7).uniform(-2.0, 2.0, cp.size)).astype(np.float32)
            return bench_lncc(torch, ref, flo, geom, cg, cp, a.steps, not a.no_cpu_baseline)
        fn = {"batch-f3d": lambda: bench_batch_f3d(torch, data["batch-f3d"], not a.no_cpu_baseline),
              "batch-aladin": lambda: bench_batch_aladin(torch, data["batch-aladin"], not a.no_cpu_baseline),
              "aladin": lambda: bench_aladin(torch, a.steps, not a.no_cpu_baseline),
              "lncc": lncc_line,
              "serve": lambda: bench_serving(torch, data["serve"])}[a.only]
        print(json.dumps(fn()), file=json_out, flush=True)
        return
    cppg, cpp = create_control_point_grid(geom, [5.0, 5.0, 5.0])
    rng = np.random.default_rng(7)
    cpp = (cpp + rng.uniform(-2.0, 2.0, cpp.size)).astype(np.float32)
    storage = 1 if a.storage == "f16" else 0
    h = F3dContent(geom, ref, geom, flo, cppg, cpp,
                   params=default_params(bins=BINS, similarity_weight=1.0 - BE, be_weight=BE, image_storage=storage))
    h.rescale_intensities((2, BINS + 1), (2, BINS + 1))
    assert h.fused_available()

    for _ in range(a.warmup):
        h.objective()
        h.gradient()
    # HIP events around every 4th sweep of each kind (each timed sweep costs the stream a few
    # microseconds of event overhead)
    h.set_profiling(True, period=SWEEP_TIMING_PERIOD)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        h.objective()
        h.gradient()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms, counts = h.sweep_times()
    value, terms = h.objective()
    h.set_profiling(False)

    N = geom.nvox
    img_bytes = 2 if storage else 4
    algo_bytes = 2 * img_bytes * N  # reference + floating read once per sweep
    grad_ms = ms[1] / max(counts[1], 1)
    hist_ms = ms[0] / max(counts[0], 1)
    achieved = algo_bytes / (grad_ms * 1e-3) / 1e9
    out = {
        "metric": "F3D obj+grad evals/s at 256x256x160",
        "value": a.steps / dt,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": dt / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if not storage else "f32 compute, f16 image storage",
        "data": "synthetic",
        "config": {
            "workload": "F3D NMI objective+gradient eval, 256x256x160 float32 shell phantom pair, "
                        "64 bins, BE 0.001, LE 0, 5-voxel control grid 55x55x35, linear interpolation",
            "voxels": N, "control_points": cppg.nvox, "final_objective": value, "nmi": terms[0] / (1.0 - BE),
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic() if not storage else None,
            "valu_frac": (measured_valu() / (grad_ms * 1e-3) / VALU_PEAK_INSTS) if measured_valu() and not storage
            else None,
            "valu_source": "profiles/f3d_sweeps_256x256x160_counters.md (rocprofv3 SQ_INSTS_VALU per launch) over "
                           "the launch time, against 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction",
            "traffic_source": "profiles/f3d_eval_256x256x160_hbm_traffic.md (rocprofv3 FETCH_SIZE, WRITE_SIZE)",
            "kernel": "k_sweep<float,true> (gradient sweep)",
            "kernel_ms": grad_ms, "histogram_sweep_ms": hist_ms, "algorithmic_bytes_per_launch": algo_bytes,
        },
        "cpu_baseline": None,
    }
    h.close()
    if not a.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(ref, flo, geom)
    if not a.no_secondary:
        out["secondary"] = [bench_aladin(torch, 20, not a.no_cpu_baseline),
                            bench_lncc(torch, ref, flo, geom, cppg, cpp, 10, not a.no_cpu_baseline),
                            bench_f3d2(torch, ref, flo, geom, 10, not a.no_cpu_baseline)]
        if not a.no_batch:
            out["secondary"] += [bench_batch_f3d(torch, data.pop("batch-f3d"), not a.no_cpu_baseline),
                                 bench_batch_aladin(torch, data.pop("batch-aladin"), not a.no_cpu_baseline),
                                 bench_serving(torch, data.pop("serve"))]
    print(json.dumps(out), file=json_out, flush=True)


if __name__ == "__main__":
    main()
