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
