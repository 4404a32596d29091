#!/usr/bin/env python3
"""Write the golden vectors of the F3D path from the reference CPU build (oracle/_ref).

Run in a container where the reference sources are present and `make -C oracle ref` has been
built; the fixtures it writes (tests/golden/*.npz) are data only: inputs and the reference's
outputs.  Two builds are used: the default SSE build (the reference's Release configuration on
x86) and the scalar build (USE_SSE off).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracles as O  # noqa: E402
from niftyreg_amd import phantoms  # noqa: E402
from niftyreg_amd.geometry import Geom, create_control_point_grid  # noqa: E402

RB = FB = 68


def case_f3d(name, dims, spacing_vox, seed, geom=None):
    rng = np.random.default_rng(seed)
    refg = geom if geom is not None else Geom.identity(dims)
    ref = phantoms.blob_phantom(refg.dims, seed).ravel()
    flo = phantoms.blob_phantom(refg.dims, seed + 3).ravel()
    cppg, cpp = create_control_point_grid(refg, [spacing_vox * d for d in refg.pixdim])
    cpp = (cpp + rng.uniform(-2, 2, cpp.size)).astype(np.float32)
    out = dict(ref=ref, flo=flo, cpp=cpp, ref_xyz=refg.xyz, ref_ijk=refg.ijk, ref_dims=np.array(refg.dims),
               ref_pixdim=np.array(refg.pixdim, np.float32), cpp_xyz=cppg.xyz, cpp_ijk=cppg.ijk,
               cpp_dims=np.array(cppg.dims), cpp_pixdim=np.array(cppg.pixdim, np.float32))
    for variant, lib in (("sse", "libnrref.so"), ("scalar", "libnrref_nosse.so")):
        O._ref = None
        O.REF = os.path.join(ROOT, "oracle", "_ref", lib)
        field = O.ref_field(refg, cppg, cpp)
        ref_r = O.ref_rescale(refg, ref, 2, RB - 3)
        flo_r = O.ref_rescale(refg, flo, 2, FB - 3)
        war = O.ref_resample(refg, flo_r, refg, field)
        wg = O.ref_image_gradient(refg, flo_r, refg, field)
        nmi, pro, logh, ent = O.ref_nmi(refg, ref_r, war, RB, FB)
        _, vg = O.ref_nmi_grad(refg, ref_r, war, wg, RB, FB)
        ng, conv = O.ref_node_gradient(refg, vg, cppg, refg, np.float32(0.999))
        be = O.ref_be(cppg, cpp)
        beg = O.ref_be_grad(cppg, cpp, np.float32(0.001))
        out.update({f"{variant}_field": field, f"{variant}_ref_rescaled": ref_r, f"{variant}_flo_rescaled": flo_r,
                    f"{variant}_warped": war, f"{variant}_warped_grad": wg, f"{variant}_nmi": np.float64(nmi),
                    f"{variant}_hist_pro": pro, f"{variant}_hist_log": logh, f"{variant}_entropies": ent,
                    f"{variant}_voxel_grad": vg, f"{variant}_conv_voxel_grad": conv, f"{variant}_node_grad": ng,
                    f"{variant}_be": np.float64(be), f"{variant}_be_grad": beg})
    np.savez_compressed(os.path.join(HERE, f"f3d_{name}.npz"), **out)
    print("wrote", name)


def case_objective(name, dims, seed):
    """reg_f3d<float> objective and normalised gradient at one level (full reference driver)."""
    O._ref = None
    O.REF = os.path.join(ROOT, "oracle", "_ref", "libnrref.so")
    refg = Geom.identity(dims)
    ref, flo = phantoms.config0_pair(seed, dims)
    r = O.RefF3d(refg, ref, refg, flo, bins=64, be=0.001, le=0.0, spacing=-5.0, levels=1)
    cpp0 = r.get_cpp()
    rng = np.random.default_rng(seed + 9)
    cpp = (cpp0 + rng.uniform(-1.5, 1.5, cpp0.size)).astype(np.float32)
    r.set_cpp(cpp)
    value, terms = r.value()
    grad, maxlen = r.gradient(normalise=True)
    ref_r, flo_r = r.images()
    cg = O.geom_from_nr(r.cpp_geom)
    out = dict(ref=ref.ravel(), flo=flo.ravel(), ref_rescaled=ref_r, flo_rescaled=flo_r, cpp=cpp,
               cpp_xyz=cg.xyz, cpp_dims=np.array(cg.dims), cpp_pixdim=np.array(cg.pixdim, np.float32),
               value=np.float64(value), terms=terms, grad=grad, max_length=np.float64(maxlen),
               weights=np.array(r.weights(), np.float32))
    r.close()
    np.savez_compressed(os.path.join(HERE, f"objective_{name}.npz"), **out)
    print("wrote", name)


def main():
    case_f3d("lut5", (22, 20, 18), 5.0, 1)
    case_f3d("spacing4", (21, 19, 17), 4.0, 2)
    th = 0.25
    m = np.eye(4, dtype=np.float32)
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    m[:3, :3] = (R @ np.diag([1.25, 1.0, 1.5])).astype(np.float32)
    m[:3, 3] = [2.0, -3.0, 5.0]
    case_f3d("oblique", None, 4.0, 3, geom=Geom((20, 18, 16), m))
    case_objective("config0_small", (40, 36, 32), 0)
    case_lncc("blobs", (32, 28, 24), 5)
    case_lncc("masked", (26, 24, 22), 6, masked=True)


def case_lncc(name, dims, seed, masked=False):
    """reg_lncc value and voxel-based gradient on a B-spline-warped pair rescaled to [0, 1]."""
    rng = np.random.default_rng(seed)
    refg = Geom.identity(dims)
    ref = phantoms.blob_phantom(refg.dims, seed).ravel()
    flo = phantoms.blob_phantom(refg.dims, seed + 3).ravel()
    cppg, cpp = create_control_point_grid(refg, [5.0] * 3)
    cpp = (cpp + rng.uniform(-2, 2, cpp.size)).astype(np.float32)
    mask = None
    if masked:
        mask = (rng.uniform(size=refg.nvox) > 0.1).astype(np.int32) - 1
    out = dict(ref=ref, flo=flo, cpp=cpp, dims=np.array(dims), cpp_dims=np.array(cppg.dims), cpp_xyz=cppg.xyz,
               mask=mask if mask is not None else np.zeros(0, np.int32))
    for variant, lib in (("sse", "libnrref.so"), ("scalar", "libnrref_nosse.so")):
        O._ref = None
        O.REF = os.path.join(ROOT, "oracle", "_ref", lib)
        field = O.ref_field(refg, cppg, cpp)
        ref_r = O.ref_rescale(refg, ref, 0, 1)
        flo_r = O.ref_rescale(refg, flo, 0, 1)
        war = O.ref_resample(refg, flo_r, refg, field, mask)
        wg = O.ref_image_gradient(refg, flo_r, refg, field, mask)
        value = O.ref_lncc(refg, ref_r, war, mask)
        grad = O.ref_lncc_grad(refg, ref_r, war, wg, mask, weight=1.0)
        if variant == "scalar":
            out.update(ref_rescaled=ref_r, warped=war, warped_grad=wg)
        out.update({f"{variant}_value": np.float64(value), f"{variant}_grad": grad})
    np.savez_compressed(os.path.join(HERE, f"lncc_{name}.npz"), **out)
    print("wrote", name)


if __name__ == "__main__":
    main()
