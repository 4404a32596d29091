"""The drop-in, end to end: the reference's own Run() of reg_f3d, reg_f3d2, reg_aladin and
reg_aladin_sym on its HIP platform.

The reference build whose reg-lib/Platform.cpp carries the HIP branches (PlatformType 3; the
copy is written by oracle/patch_platform.py, the reference tree is only read) selects the plugin
classes of niftyreg_amd/plugin (HipComputeFactory, HipContentCreatorFactory, HipKernelFactory,
HipMeasureCreatorFactory, HipConjugateGradient), which drive libnrhip.  Each registration runs
on the HIP platform and on the CPU platform of the same build.  F3D and F3D2 results are scored
with the reference's own objective: the per-operation device kernels the plugin drives reproduce
the build's own summation orders (the SSE build's four-lane field and flow sums and its
convolution), so the whole registration follows the CPU platform's trajectory and the final
scores agree to 1e-9 (relative).  Aladin matrices are held to the bounds of the device Aladin
tests.
"""
import time

import numpy as np
import pytest

import oracles as O
from niftyreg_amd import phantoms
from niftyreg_amd.geometry import Geom

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (O.ref_available() and O.ref_hip_available()),
                                 reason="reference build with the HIP platform absent")]


def test_reg_f3d_run_on_hip_platform(gpu):
    ref, flo = phantoms.config0_pair(0)
    g = Geom.identity((64, 64, 64))
    hcg, hcpp = O.ref_f3d_run(g, ref, g, flo, levels=3, maxit=50, platform=O.PLATFORM_HIP)
    ccg, ccpp = O.ref_f3d_run(g, ref, g, flo, levels=3, maxit=50)
    _, scpp = O.ref_f3d_run(g, ref, g, flo, levels=3, maxit=50, scalar=True)
    assert hcg.dims == ccg.dims
    n_hip, n_cpu, n_scalar = (O.ref_nmi_at(g, ref, g, flo, c, levels=3) for c in (hcpp, ccpp, scpp))
    print(f"reg_f3d Run: HIP platform {n_hip:.9f}, CPU platform {n_cpu:.9f}, scalar build {n_scalar:.9f}")
    assert abs(n_hip - n_cpu) <= 1e-9 * n_cpu


@pytest.mark.parametrize("measure", [0, 1])
def test_reg_f3d_run_on_hip_platform_fused_route(gpu, monkeypatch, measure):
    """NRH_PLUGIN_FUSED=1: the plugin records reg_f3d's field / warp / gradient stages and runs
    them as nrh_f3d_similarity_value / _gradient (the fused sweeps for NMI, the separable node
    projection for both measures).  Their rounding differs from the per-operation kernels' at
    1e-7, so the trajectory leaves the CPU platform's bit path: the final score (NMI, or LNCC) is
    held to the registration bound of the device driver's tests (5e-4 relative) against both
    reference builds' spread."""
    ref, flo = phantoms.config0_pair(0)
    g = Geom.identity((64, 64, 64))
    kw = dict(levels=3, maxit=50, measure=measure)
    monkeypatch.setenv("NRH_PLUGIN_FUSED", "1")
    fcg, fcpp = O.ref_f3d_run(g, ref, g, flo, platform=O.PLATFORM_HIP, **kw)
    monkeypatch.delenv("NRH_PLUGIN_FUSED")
    ccg, ccpp = O.ref_f3d_run(g, ref, g, flo, **kw)
    _, scpp = O.ref_f3d_run(g, ref, g, flo, scalar=True, **kw)
    assert fcg.dims == ccg.dims
    if measure == 0:
        n_f, n_c, n_s = (O.ref_nmi_at(g, ref, g, flo, c, levels=3) for c in (fcpp, ccpp, scpp))
    else:
        n_f, n_c, n_s = (O.ref_lncc_at(g, ref, g, flo, c, levels=3) for c in (fcpp, ccpp, scpp))
    print(f"reg_f3d Run (measure {measure}): HIP fused route {n_f:.9f}, CPU platform {n_c:.9f}, "
          f"scalar build {n_s:.9f}")
    assert abs(n_f - n_c) <= 5e-4 * abs(n_c) + abs(n_c - n_s)


@pytest.mark.skipif(not O.plugin_available(), reason="plugin check library not built")
def test_plugin_evaluation_timing(gpu, monkeypatch):
    """Objective + gradient evaluations of reg_f3d's own calls at 256x256x160 (NMI, 5-voxel grid)
    on the reference CPU classes (scalar build, OpenMP) and on the HIP platform's per-operation and
    fused routes: evaluations per second of each (printed), the fused route faster than the
    per-operation one."""
    from niftyreg_amd.geometry import create_control_point_grid
    dims = (256, 256, 160)
    ref, flo = phantoms.config1_pair(0, dims)
    g = Geom.identity(dims)
    cg, cpp = create_control_point_grid(g, [5.0, 5.0, 5.0])
    cpp = (cpp + np.random.default_rng(0).uniform(-1, 1, cpp.size)).astype(np.float32)
    rates = {}
    for name, hip, fused, iters in (("hip_fused", True, "1", 40), ("hip_ops", True, None, 20), ("cpu", False, None, 2)):
        if fused:
            monkeypatch.setenv("NRH_PLUGIN_FUSED", fused)
        else:
            monkeypatch.delenv("NRH_PLUGIN_FUSED", raising=False)
        rates[name] = iters / O.plugin_f3d_time(hip, g, ref, g, flo, cg, cpp, iters=iters)
    print("reg_f3d evaluation calls at 256x256x160 (evals/s): " + ", ".join(f"{k} {v:.2f}" for k, v in rates.items()))
    assert rates["hip_fused"] > rates["hip_ops"] > rates["cpu"]


def test_reg_f3d2_run_on_hip_platform(gpu):
    ref, flo = phantoms.config0_pair(0)
    g = Geom.identity((64, 64, 64))
    hcg, hfw, hbw, _ = O.ref_f3d2_run(g, ref, g, flo, levels=2, maxit=30, platform=O.PLATFORM_HIP)
    ccg, cfw, cbw, _ = O.ref_f3d2_run(g, ref, g, flo, levels=2, maxit=30)
    _, sfw, sbw, _ = O.ref_f3d2_run(g, ref, g, flo, levels=2, maxit=30, scalar=True)
    assert hcg.dims == ccg.dims and np.array_equal(hcg.xyz, ccg.xyz)
    s_hip = O.ref_nmi_sym_at(g, ref, g, flo, hfw, hbw, levels=2)
    s_cpu = O.ref_nmi_sym_at(g, ref, g, flo, cfw, cbw, levels=2)
    s_scalar = O.ref_nmi_sym_at(g, ref, g, flo, sfw, sbw, levels=2)
    print(f"reg_f3d2 Run: HIP platform {s_hip:.9f}, CPU platform {s_cpu:.9f}, scalar build {s_scalar:.9f}")
    assert abs(s_hip - s_cpu) <= 1e-9 * s_cpu


@pytest.mark.parametrize("rigid,symmetric", [(False, False), (True, False), (False, True), (True, True)])
def test_reg_aladin_run_on_hip_platform(gpu, rigid, symmetric):
    dims = (64, 64, 64)
    ref, flo, _ = phantoms.config3_pair(1, dims)
    g = Geom.identity(dims)
    H = O.ref_aladin_run(g, ref, g, flo, levels=3, maxit=5, rigid=rigid, affine=True, symmetric=symmetric,
                         platform=O.PLATFORM_HIP)
    C = O.ref_aladin_run(g, ref, g, flo, levels=3, maxit=5, rigid=rigid, affine=True, symmetric=symmetric)
    print(f"reg_aladin{'_sym' if symmetric else ''} Run (rigid={rigid}): HIP\n{H}\nCPU\n{C}")
    tol_lin, tol_t = (0.03, 1.0) if rigid else (5e-3, 0.1)
    assert np.abs(H[:3, :3] - C[:3, :3]).max() < tol_lin
    assert np.abs(H[:3, 3] - C[:3, 3]).max() < tol_t
