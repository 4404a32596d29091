"""The fused objective and gradient against reg_f3d<float> of the reference CPU build itself.

One level of reg_f3d (NMI 64 bins, bending energy 0.001, linear energy 0, 5-voxel grid) is
prepared on the reference build; its objective terms and unnormalised node gradient at a grid are
compared with nrh_f3d_objective / nrh_f3d_gradient on the device at the same grid, with the
device reproducing that build's field rounding (field_order).  Cases: the identity grid every
registration starts from (every voxel then maps exactly onto a floating voxel, where the float
field's last bit decides the trilinear cell and, on the border, whether the voxel is inside) and
a random grid, on the 64^3 blob pair and on the 256x256x160 shell pair, against the default SSE
build and the scalar build.

Tolerances (stated for this path): NMI relative 1e-5, node gradient relative 1e-3 in the L2
norm; the largest element-wise difference is reported beside the spread between the two
reference builds measured on the same grid.
"""
import numpy as np
import pytest

import oracles as O
from niftyreg_amd import _lib, phantoms
from niftyreg_amd.f3d import F3dContent, default_params
from niftyreg_amd.geometry import Geom

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not O.ref_available(), reason="reference build absent")]

NMI_RTOL = 1e-5
GRAD_RTOL = 1e-3

CASES = {
    "blob64_identity": dict(dims=(64, 64, 64), amp=0.0),
    "blob64_random": dict(dims=(64, 64, 64), amp=2.0),
    "shell256_identity": dict(dims=(256, 256, 160), amp=0.0),
    "shell256_random": dict(dims=(256, 256, 160), amp=2.0),
}


def l2rel(a, b):
    return float(np.linalg.norm((a - b).astype(np.float64)) / np.linalg.norm(b.astype(np.float64)))


def images(dims):
    if dims == (64, 64, 64):
        return phantoms.config0_pair(0)
    return phantoms.config1_pair(seed=0, dims=dims)


def reference_eval(g, ref, flo, amp, scalar):
    r = O.RefF3d(g, ref, g, flo, bins=64, be=0.001, le=0.0, spacing=-5.0, levels=1, scalar=scalar)
    cpp = r.get_cpp()
    if amp:
        cpp = (cpp + np.random.default_rng(11).uniform(-amp, amp, cpp.size)).astype(np.float32)
    r.set_cpp(cpp)
    value, terms = r.value()
    grad, _ = r.gradient(normalise=False)
    w = r.weights()
    cg = O.geom_from_nr(r.cpp_geom)
    r.close()
    return cpp, cg, w, value, terms, grad


@pytest.mark.parametrize("case", sorted(CASES))
def test_fused_matches_reference_build(gpu, case):
    c = CASES[case]
    g = Geom.identity(c["dims"])
    ref, flo = images(c["dims"])
    ref, flo = ref.ravel(), flo.ravel()
    runs = {}
    for build in ("sse", "scalar"):
        runs[build] = reference_eval(g, ref, flo, c["amp"], build == "scalar")
    spread_nmi = abs(runs["sse"][4][0] - runs["scalar"][4][0]) / runs["scalar"][4][0]
    spread_grad = l2rel(runs["sse"][5], runs["scalar"][5])
    for build, order in (("sse", _lib.FIELD_ORDER_SSE), ("scalar", _lib.FIELD_ORDER_SCALAR)):
        cpp, cg, (simw, bew), rvalue, rterms, rgrad = runs[build]
        h = F3dContent(g, ref, g, flo, cg, cpp,
                       params=default_params(similarity_weight=simw, be_weight=bew, field_order=order))
        try:
            h.rescale_intensities((2, 65), (2, 65))
            assert h.fused_available()
            value, terms = h.objective()
            h.gradient()
            grad = h.node_gradient()
        finally:
            h.close()
        nmi_err = abs(terms[0] - rterms[0]) / rterms[0]
        grad_err = l2rel(grad, rgrad)
        max_err = float(np.abs(grad - rgrad).max() / np.abs(rgrad).max())
        print(f"{case} vs {build} build: NMI rel {nmi_err:.2e}, gradient L2 rel {grad_err:.2e}, max-abs rel "
              f"{max_err:.2e} (between the builds: NMI {spread_nmi:.2e}, gradient L2 {spread_grad:.2e})")
        assert nmi_err <= NMI_RTOL
        assert abs(terms[1] - rterms[1]) <= 1e-9 * max(abs(rterms[1]), 1e-30)
        assert grad_err <= GRAD_RTOL
