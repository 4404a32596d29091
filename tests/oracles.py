"""ctypes access to the checkers: oracle/liboracle.so (the C restatement) and
oracle/_ref/libnrref.so (the reference CPU path compiled from its own sources).

Test infrastructure only.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libnrref.so")

F = ctypes.POINTER(ctypes.c_float)
D = ctypes.POINTER(ctypes.c_double)
I = ctypes.POINTER(ctypes.c_int)


def fp(a):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(F)


def dp(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(D)


def ip(a):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(I)


class orc_geom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
                ("xyz", ctypes.c_float * 16), ("ijk", ctypes.c_float * 16)]


class nr_geom(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("pixdim", ctypes.c_float * 3),
                ("sform", ctypes.c_float * 16)]


class ref_f3d_params(ctypes.Structure):
    _fields_ = [("bins", ctypes.c_int), ("be", ctypes.c_float), ("le", ctypes.c_float), ("spacing", ctypes.c_float),
                ("levels", ctypes.c_int), ("levels_to_perform", ctypes.c_int), ("maxit", ctypes.c_int),
                ("use_pyramid", ctypes.c_int), ("padding", ctypes.c_float), ("measure", ctypes.c_int),
                ("lncc_sigma", ctypes.c_float), ("affine", ctypes.POINTER(ctypes.c_float))]


def og(g):
    o = orc_geom()
    o.nx, o.ny, o.nz = g.dims
    o.dx, o.dy, o.dz = g.pixdim
    o.xyz[:] = [float(v) for v in g.xyz.reshape(16)]
    o.ijk[:] = [float(v) for v in g.ijk.reshape(16)]
    return o


def rg(g):
    r = nr_geom()
    r.nx, r.ny, r.nz = g.dims
    r.pixdim[:] = [float(v) for v in g.pixdim]
    r.sform[:] = [float(v) for v in g.xyz.reshape(16)]
    return r


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        _orc = ctypes.CDLL(ORACLE)
        _orc.orc_nmi.restype = ctypes.c_double
        _orc.orc_bending_energy.restype = ctypes.c_double
        _orc.orc_max_length.restype = ctypes.c_float
        _orc.orc_cg_update.restype = ctypes.c_double
    return _orc


def ref_available():
    return os.path.exists(REF)


def ref():
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF)
        _ref.ref_last_error.restype = ctypes.c_char_p
        _ref.ref_f3d_create.restype = ctypes.c_void_p
        for n in ("ref_f3d_destroy", "ref_f3d_level_geom", "ref_f3d_level_images", "ref_f3d_set_cpp",
                  "ref_f3d_get_cpp", "ref_f3d_value", "ref_f3d_gradient", "ref_f3d_weights"):
            getattr(_ref, n).argtypes = None
    return _ref


def rcheck(rc):
    if rc != 0:
        raise RuntimeError(ref().ref_last_error().decode())


# ---------------------------------------------------------------- oracle wrappers
def orc_field(refg, cppg, cpp, mask=None, init=None):
    n = refg.nvox
    field = np.zeros(3 * n, np.float32) if init is None else init.astype(np.float32).copy()
    orc().orc_spline_field(ctypes.byref(og(refg)), ctypes.byref(og(cppg)), fp(cpp), ip(mask), fp(field))
    return field


def orc_resample(flog, flo, refg, field, mask=None, pad=float("nan"), init=None):
    w = np.zeros(refg.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    orc().orc_resample_linear(ctypes.byref(og(flog)), fp(flo), ctypes.byref(og(refg)), fp(field), ip(mask),
                              ctypes.c_float(pad), fp(w))
    return w


def orc_image_gradient(flog, flo, refg, field, mask=None, pad=float("nan")):
    g = np.zeros(3 * refg.nvox, np.float32)
    orc().orc_image_gradient_linear(ctypes.byref(og(flog)), fp(flo), ctypes.byref(og(refg)), fp(field), ip(mask),
                                    ctypes.c_float(pad), fp(g))
    return g


def orc_rescale(a, lo, hi):
    a = np.ascontiguousarray(a, np.float32).copy()
    orc().orc_intensity_rescale(fp(a), ctypes.c_size_t(a.size), ctypes.c_float(lo), ctypes.c_float(hi))
    return a


def orc_nmi(refv, war, rb, fb, mask=None):
    tot = rb * fb + rb + fb
    pro = np.zeros(tot, np.float64)
    logh = np.zeros(tot, np.float64)
    ent = np.zeros(4, np.float64)
    v = orc().orc_nmi(fp(refv), fp(war), ip(mask), ctypes.c_size_t(refv.size), rb, fb, dp(pro), dp(logh), dp(ent))
    return v, pro, logh, ent


def orc_nmi_grad(refv, war, wgrad, rb, fb, logh, ent, weight=1.0, mask=None, init=None):
    g = np.zeros(3 * refv.size, np.float32) if init is None else init.astype(np.float32).copy()
    orc().orc_nmi_voxel_gradient(fp(refv), fp(war), fp(wgrad), ip(mask), ctypes.c_size_t(refv.size), rb, fb,
                                 dp(logh), dp(ent), ctypes.c_double(weight), fp(g))
    return g


def orc_convolve_voxel_gradient(refg, cppg, vg):
    vg = vg.astype(np.float32).copy()
    orc().orc_convolve_voxel_gradient(ctypes.byref(og(refg)), ctypes.byref(og(cppg)), fp(vg))
    return vg


def orc_voxel_to_node(cppg, refg, vg, flog, weight):
    ng = np.zeros(3 * cppg.nvox, np.float32)
    ijk = np.ascontiguousarray(flog.ijk, np.float32).reshape(16)
    orc().orc_voxel_to_node(ctypes.byref(og(cppg)), ctypes.byref(og(refg)), fp(vg), fp(ijk), ctypes.c_float(weight), fp(ng))
    return ng


def orc_be(cppg, cpp):
    return orc().orc_bending_energy(ctypes.byref(og(cppg)), fp(cpp))


def orc_be_grad(cppg, cpp, weight, init=None):
    g = np.zeros(3 * cppg.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    orc().orc_bending_energy_gradient(ctypes.byref(og(cppg)), fp(cpp), ctypes.c_float(weight), fp(g))
    return g


def orc_node_gradient_chain(refg, flog, cppg, ref_r, flo_r, cpp, rb, fb, sim_w, be_w, pad=float("nan"), mask=None):
    """reg_f3d::GetObjectiveFunctionGradient on the oracle: (node gradient, nmi, voxel gradient)."""
    field = orc_field(refg, cppg, cpp, mask)
    war = orc_resample(flog, flo_r, refg, field, mask, pad)
    nmi, pro, logh, ent = orc_nmi(ref_r, war, rb, fb, mask)
    wg = orc_image_gradient(flog, flo_r, refg, field, mask, pad)
    vg = orc_nmi_grad(ref_r, war, wg, rb, fb, logh, ent, 1.0, mask)
    cv = orc_convolve_voxel_gradient(refg, cppg, vg)
    ng = orc_voxel_to_node(cppg, refg, cv, flog, sim_w)
    if be_w > 0:
        ng = orc_be_grad(cppg, cpp, be_w, ng)
    return ng, nmi, vg


# ---------------------------------------------------------------- reference wrappers
def ref_field(refg, cppg, cpp, mask=None, force_no_lut=False, init=None):
    field = np.zeros(3 * refg.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    rcheck(ref().ref_spline_field(ctypes.byref(rg(refg)), ctypes.byref(rg(cppg)), fp(cpp), ip(mask),
                                  int(force_no_lut), fp(field)))
    return field


def ref_resample(flog, flo, refg, field, mask=None, pad=float("nan"), init=None):
    w = np.zeros(refg.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    rcheck(ref().ref_resample(ctypes.byref(rg(flog)), fp(flo), ctypes.byref(rg(refg)), fp(field), ip(mask), 1,
                              ctypes.c_float(pad), fp(w)))
    return w


def ref_image_gradient(flog, flo, refg, field, mask=None, pad=float("nan")):
    g = np.zeros(3 * refg.nvox, np.float32)
    rcheck(ref().ref_image_gradient(ctypes.byref(rg(flog)), fp(flo), ctypes.byref(rg(refg)), fp(field), ip(mask), 1,
                                    ctypes.c_float(pad), fp(g)))
    return g


def ref_rescale(g, a, lo, hi):
    a = np.ascontiguousarray(a, np.float32).copy()
    rcheck(ref().ref_intensity_rescale(ctypes.byref(rg(g)), fp(a), ctypes.c_float(lo), ctypes.c_float(hi)))
    return a


def ref_nmi(g, refv, war, rb, fb, mask=None):
    tot = rb * fb + rb + fb
    pro = np.zeros(tot, np.float64)
    logh = np.zeros(tot, np.float64)
    ent = np.zeros(4, np.float64)
    v = ctypes.c_double()
    rcheck(ref().ref_nmi_value(ctypes.byref(rg(g)), fp(refv), fp(war), ip(mask), rb, fb, 1, ctypes.byref(v),
                               dp(pro), dp(logh), dp(ent)))
    return v.value, pro, logh, ent


def ref_nmi_grad(g, refv, war, wgrad, rb, fb, mask=None, init=None):
    vg = np.zeros(3 * g.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    v = ctypes.c_double()
    rcheck(ref().ref_nmi_gradient(ctypes.byref(rg(g)), fp(refv), fp(war), fp(wgrad), ip(mask), rb, fb,
                                  ctypes.byref(v), fp(vg)))
    return v.value, vg


def ref_node_gradient(refg, vg, cppg, flog, weight):
    vg = vg.astype(np.float32).copy()
    ng = np.zeros(3 * cppg.nvox, np.float32)
    rcheck(ref().ref_node_gradient(ctypes.byref(rg(refg)), fp(vg), ctypes.byref(rg(cppg)), ctypes.byref(rg(flog)),
                                   ctypes.c_float(weight), fp(ng)))
    return ng, vg


def ref_kernel_convolution(g, data, sigma, kernel, axes, mask=None):
    data = np.ascontiguousarray(data, np.float32).copy()
    nt = data.size // g.nvox
    s = np.full(nt, sigma, np.float32)
    ax = np.asarray(axes, np.int32)
    rcheck(ref().ref_kernel_convolution(ctypes.byref(rg(g)), nt, fp(data), fp(s), kernel, ip(mask), ip(ax)))
    return data


def ref_be(cppg, cpp):
    v = ctypes.c_double()
    rcheck(ref().ref_be_value(ctypes.byref(rg(cppg)), fp(cpp), ctypes.byref(v)))
    return v.value


def ref_be_grad(cppg, cpp, weight, init=None):
    g = np.zeros(3 * cppg.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    rcheck(ref().ref_be_gradient(ctypes.byref(rg(cppg)), fp(cpp), ctypes.c_float(weight), fp(g), None))
    return g


def ref_create_cpp(refg, spacing):
    out = nr_geom()
    sp = np.asarray(spacing, np.float32)
    rcheck(ref().ref_create_cpp(ctypes.byref(rg(refg)), fp(sp), ctypes.byref(out), None))
    n = out.nx * out.ny * out.nz
    data = np.zeros(3 * n, np.float32)
    rcheck(ref().ref_create_cpp(ctypes.byref(rg(refg)), fp(sp), ctypes.byref(out), fp(data)))
    return (out.nx, out.ny, out.nz), np.array(out.sform[:], np.float32).reshape(4, 4), data, tuple(out.pixdim[:])


def orc_affine_field(refg, affine, mask=None):
    field = np.zeros(3 * refg.nvox, np.float32)
    a = np.ascontiguousarray(affine, np.float32).reshape(16)
    orc().orc_affine_field(ctypes.byref(og(refg)), fp(a), ip(mask), fp(field))
    return field


def ref_affine_field(refg, affine, mask=None):
    field = np.zeros(3 * refg.nvox, np.float32)
    a = np.ascontiguousarray(affine, np.float32).reshape(16)
    rcheck(ref().ref_affine_field(ctypes.byref(rg(refg)), fp(a), ip(mask), fp(field)))
    return field


class RefF3d:
    """reg_f3d<float> prepared at one level for single objective / gradient evaluations.
    scalar=True uses the reference built without USE_SSE, else the default SSE build."""

    def __init__(self, refg, ref_img, flog, flo_img, bins=64, be=0.001, le=0.0, spacing=-5.0, levels=1,
                 level=0, maxit=150, padding=float("nan"), measure=0, lncc_sigma=-5.0, scalar=False):
        p = ref_f3d_params()
        p.bins, p.be, p.le, p.spacing = bins, be, le, spacing
        p.levels, p.levels_to_perform, p.maxit, p.use_pyramid, p.padding = levels, levels, maxit, 1, padding
        p.measure, p.lncc_sigma = measure, lncc_sigma
        self.p = p
        self.lib = r = ref_scalar() if scalar else ref()
        r.ref_f3d_create.restype = ctypes.c_void_p
        r.ref_f3d_create.argtypes = [ctypes.c_void_p] * 6
        h = r.ref_f3d_create(ctypes.byref(rg(refg)), fp(np.ascontiguousarray(ref_img, np.float32).reshape(-1)),
                             ctypes.byref(rg(flog)), fp(np.ascontiguousarray(flo_img, np.float32).reshape(-1)),
                             ctypes.byref(p), ctypes.c_void_p(level))
        if not h:
            raise RuntimeError(r.ref_last_error().decode())
        self.h = ctypes.c_void_p(h)
        cg, rgm, fgm = nr_geom(), nr_geom(), nr_geom()
        self._check(r.ref_f3d_level_geom(self.h, ctypes.byref(cg), ctypes.byref(rgm), ctypes.byref(fgm)))
        self.cpp_geom, self.ref_geom, self.flo_geom = cg, rgm, fgm
        self.G = cg.nx * cg.ny * cg.nz

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.lib.ref_last_error().decode())

    def images(self):
        nr_ = self.ref_geom.nx * self.ref_geom.ny * self.ref_geom.nz
        nf = self.flo_geom.nx * self.flo_geom.ny * self.flo_geom.nz
        a, b = np.zeros(nr_, np.float32), np.zeros(nf, np.float32)
        self._check(self.lib.ref_f3d_level_images(self.h, fp(a), fp(b)))
        return a, b

    def weights(self, with_le=False):
        """(similarity, bending-energy[, linear-energy]) weights of the prepared level."""
        w = np.zeros(3, np.float32)
        self._check(self.lib.ref_f3d_weights(self.h, fp(w)))
        return (float(w[0]), float(w[1]), float(w[2])) if with_le else (float(w[0]), float(w[1]))

    def get_cpp(self):
        c = np.zeros(3 * self.G, np.float32)
        self._check(self.lib.ref_f3d_get_cpp(self.h, fp(c)))
        return c

    def set_cpp(self, c):
        self._check(self.lib.ref_f3d_set_cpp(self.h, fp(np.ascontiguousarray(c, np.float32))))

    def value(self):
        v = ctypes.c_double()
        t = np.zeros(3, np.float64)
        self._check(self.lib.ref_f3d_value(self.h, ctypes.byref(v), dp(t)))
        return v.value, t

    def gradient(self, normalise=False):
        g = np.zeros(3 * self.G, np.float32)
        m = ctypes.c_double()
        self._check(self.lib.ref_f3d_gradient(self.h, int(normalise), ctypes.byref(m), fp(g), None))
        return g, m.value

    def close(self):
        if self.h:
            self.lib.ref_f3d_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def geom_from_nr(g):
    from niftyreg_amd.geometry import Geom
    return Geom((g.nx, g.ny, g.nz), np.array(g.sform[:], np.float32).reshape(4, 4), pixdim=tuple(g.pixdim[:]))


# ---------------------------------------------------------------- FP64 references (numpy)
def _axis_weights_fp64(n, gvs, nodes):
    """Dense (n x nodes) cubic B-spline weights in float64 at the exact cell positions."""
    W = np.zeros((n, nodes), np.float64)
    for v in range(n):
        pre = int(np.floor(v / gvs))
        t = v / gvs - pre
        b = ((1 - t) ** 3 / 6, (3 * t ** 3 - 6 * t ** 2 + 4) / 6, (-3 * t ** 3 + 3 * t ** 2 + 3 * t + 1) / 6, t ** 3 / 6)
        for a in range(4):
            W[v, pre + a] = b[a]
    return W


def fp64_field(refg, cppg, cpp):
    """The cubic B-spline deformation field in float64 (separable dense contraction)."""
    nx, ny, nz = refg.dims
    gx, gy, gz = cppg.dims
    gvs = [cppg.pixdim[i] / refg.pixdim[i] for i in range(3)]
    Wx, Wy, Wz = (_axis_weights_fp64(n, g, m) for n, g, m in zip((nx, ny, nz), gvs, (gx, gy, gz)))
    c = np.asarray(cpp, np.float64).reshape(3, gz, gy, gx)
    out = np.empty((3, nz, ny, nx), np.float64)
    for k in range(3):
        t = np.einsum("kji,xi->kjx", c[k], Wx)
        t = np.einsum("kjx,yj->kyx", t, Wy)
        out[k] = np.einsum("kyx,zk->zyx", t, Wz)
    return out.reshape(-1)


def ref_f3d_run(refg, ref_img, flog, flo_img, bins=64, be=0.001, le=0.0, spacing=-5.0, levels=3, maxit=150,
                padding=float("nan"), measure=0, lncc_sigma=-5.0, affine=None, scalar=False, levels_to_perform=0,
                platform=None):
    """reg_f3d<float>::Run on the reference CPU build (scalar=True: the build without USE_SSE), or
    with platform=PLATFORM_HIP on the reference build whose Platform selects libnrhip; returns
    (grid geom, control points)."""
    p = ref_f3d_params()
    p.bins, p.be, p.le, p.spacing = bins, be, le, spacing
    p.levels, p.levels_to_perform, p.maxit, p.use_pyramid, p.padding = (levels, levels_to_perform or levels, maxit, 1,
                                                                        padding)
    p.measure, p.lncc_sigma = measure, lncc_sigma
    aff = None if affine is None else np.ascontiguousarray(affine, np.float32).reshape(16)
    p.affine = None if aff is None else fp(aff)
    out = nr_geom()
    sp = [abs(spacing) * d if spacing < 0 else spacing for d in refg.pixdim]
    cap = 3 * int(np.prod([int(np.ceil(n * d / s + 4)) for n, d, s in zip(refg.dims, refg.pixdim, sp)]))
    buf = np.zeros(cap, np.float32)
    args = (ctypes.byref(rg(refg)), fp(np.ascontiguousarray(ref_img, np.float32).reshape(-1)), ctypes.byref(rg(flog)),
            fp(np.ascontiguousarray(flo_img, np.float32).reshape(-1)), ctypes.byref(p), ctypes.byref(out), fp(buf), cap)
    if platform is not None:
        lib = ref_hip()
        _scheck(lib, lib.ref_f3d_run_on(int(platform), *args))
    else:
        lib = ref_scalar() if scalar else ref()
        _scheck(lib, lib.ref_f3d_run(*args))
    n = 3 * out.nx * out.ny * out.nz
    return geom_from_nr(out), buf[:n].copy()


def nmi_at(refg, flog, cppg, cpp, ref_img, flo_img, bins=64, fp64=False):
    """NMI of the pair warped through a control-point grid, on the oracle (images rescaled here).
    fp64: evaluate the field in double (rounded once to float) instead of the reference's float
    sums; the two differ in whether voxels mapped exactly onto the last voxel plane (identity
    grids) fall outside the floating image."""
    rb = bins + 4
    r = orc_rescale(np.asarray(ref_img, np.float32).ravel(), 2, rb - 3)
    f = orc_rescale(np.asarray(flo_img, np.float32).ravel(), 2, rb - 3)
    field = fp64_field(refg, cppg, cpp).astype(np.float32) if fp64 else orc_field(refg, cppg, cpp)
    war = orc_resample(flog, f, refg, field)
    return orc_nmi(r, war, rb, rb)[0]


def ref_nmi_at(refg, ref_img, flog, flo_img, cpp, bins=64, spacing=-5.0, levels=1, scalar=False):
    """NMI of reg_f3d<float> itself (the reference build, no penalty terms) at the control points
    cpp of the finest level of a `levels`-level run: the reference's objective as a registration
    score."""
    r = RefF3d(refg, ref_img, flog, flo_img, bins=bins, be=0.0, le=0.0, spacing=spacing, levels=levels,
               level=levels - 1, scalar=scalar)
    try:
        assert 3 * r.G == np.asarray(cpp).size, "grid of another geometry"
        r.set_cpp(cpp)
        _, terms = r.value()
        return terms[0] / r.weights()[0]
    finally:
        r.close()


def ref_lncc_at(refg, ref_img, flog, flo_img, cpp, spacing=-5.0, levels=1, lncc_sigma=-5.0, scalar=False):
    """LNCC of reg_f3d<float> itself (the reference build, no penalty terms) at the control points
    cpp of the finest level of a `levels`-level run."""
    r = RefF3d(refg, ref_img, flog, flo_img, be=0.0, le=0.0, spacing=spacing, levels=levels, level=levels - 1,
               measure=1, lncc_sigma=lncc_sigma, scalar=scalar)
    try:
        assert 3 * r.G == np.asarray(cpp).size, "grid of another geometry"
        r.set_cpp(cpp)
        _, terms = r.value()
        return terms[0] / r.weights()[0]
    finally:
        r.close()


def ref_nmi_sym_at(refg, ref_img, flog, flo_img, fw, bw, bins=64, spacing=-5.0, levels=1, affine=None,
                   scalar=False):
    """Summed forward / backward NMI of reg_f3d2<float> itself (no penalty terms) at the velocity
    grids fw, bw of the finest level of a `levels`-level run, with the input affine's half affines
    on the grids when given."""
    r = RefF3d2(refg, ref_img, flog, flo_img, bins=bins, be=0.0, le=0.0, spacing=spacing, levels=levels,
                level=levels - 1, scalar=scalar, affine=affine)
    try:
        assert 3 * r.G == np.asarray(fw).size, "grid of another geometry"
        r.set_cpp(fw, bw)
        _, terms = r.value()
        return terms[0] / r.weights()[0]
    finally:
        r.close()


# ---------------------------------------------------------------- Aladin
class ref_aladin_params(ctypes.Structure):
    _fields_ = [("levels", ctypes.c_int), ("levels_to_perform", ctypes.c_int), ("max_iterations", ctypes.c_int),
                ("block_percent", ctypes.c_int), ("inlier_lts", ctypes.c_int), ("rigid", ctypes.c_int),
                ("affine", ctypes.c_int), ("align_centre", ctypes.c_int), ("symmetric", ctypes.c_int)]


def block_count(g):
    nb = [(n + 3) // 4 for n in g.dims]
    return nb, nb[0] * nb[1] * nb[2]


def orc_active_blocks(g, img, percent=50, mask=None):
    _, total = block_count(g)
    tb = np.zeros(total, np.int32)
    n = orc().orc_active_blocks(ctypes.byref(og(g)), fp(img), ip(mask), percent, ip(tb))
    return tb, n


def orc_block_matching(g, img, warped, tb, n_active, mask=None):
    rp = np.zeros(3 * n_active, np.float32)
    wp = np.zeros(3 * n_active, np.float32)
    d = orc().orc_block_matching(ctypes.byref(og(g)), fp(img), fp(warped), ip(mask), ip(tb), fp(rp), fp(wp))
    return rp, wp, d


def orc_lts(rp, wp, percent=50, affine=True, matrix=None):
    m = np.eye(4, dtype=np.float32) if matrix is None else np.array(matrix, np.float32).copy()
    rc = orc().orc_lts(fp(rp), fp(wp), len(rp) // 3, percent, int(affine), fp(m))
    if rc != 0:
        raise RuntimeError("too few correspondences")
    return m


def ref_block_matching(g, img, warped=None, percent=50, mask=None):
    _, total = block_count(g)
    tb = np.zeros(total, np.int32)
    counts = np.zeros(3, np.int32)
    n_max = total
    rp = np.zeros(3 * n_max, np.float32)
    wp = np.zeros(3 * n_max, np.float32)
    rcheck(ref().ref_block_matching(ctypes.byref(rg(g)), fp(img), None if warped is None else fp(warped), ip(mask),
                                    percent, 50, ip(tb), fp(rp), fp(wp), ip(counts)))
    n = int(counts[1])
    return tb, n, rp[:3 * n], wp[:3 * n], int(counts[2])


def ref_mat44_avg2(A, B):
    out = np.zeros(16, np.float32)
    rcheck(ref().ref_mat44_avg2(fp(np.ascontiguousarray(A, np.float32).ravel()),
                                fp(np.ascontiguousarray(B, np.float32).ravel()), fp(out)))
    return out.reshape(4, 4)


def ref_mat44_inverse(A):
    out = np.zeros(16, np.float32)
    rcheck(ref().ref_mat44_inverse(fp(np.ascontiguousarray(A, np.float32).ravel()), fp(out)))
    return out.reshape(4, 4)


def ref_lts(rp, wp, n_defined, percent=50, affine=True, matrix=None):
    m = np.eye(4, dtype=np.float32) if matrix is None else np.array(matrix, np.float32).copy()
    rcheck(ref().ref_lts(fp(rp), fp(wp), len(rp) // 3, n_defined, percent, int(affine), fp(m)))
    return m


def ref_aladin_run(refg, ref_img, flog, flo_img, levels=3, maxit=5, block_percent=50, inlier=50, rigid=False,
                   affine=True, align_centre=True, symmetric=False, platform=None):
    """reg_aladin<float>::Run (reg_aladin_sym with symmetric=True) on the reference CPU build, or with
    platform=PLATFORM_HIP on the HIP-platform build; returns the 4x4 affine."""
    p = ref_aladin_params(levels, levels, maxit, block_percent, inlier, int(rigid), int(affine), int(align_centre),
                          int(symmetric))
    m = np.zeros(16, np.float32)
    args = (ctypes.byref(rg(refg)), fp(np.ascontiguousarray(ref_img, np.float32).ravel()), ctypes.byref(rg(flog)),
            fp(np.ascontiguousarray(flo_img, np.float32).ravel()), ctypes.byref(p), fp(m))
    if platform is not None:
        lib = ref_hip()
        _scheck(lib, lib.ref_aladin_run_on(int(platform), *args))
    else:
        rcheck(ref().ref_aladin_run(*args))
    return m.reshape(4, 4)


# ---------------------------------------------------------------- LNCC
def orc_convolution(g, data, sigma, kernel=2, mask=None):
    out = np.ascontiguousarray(data, np.float32).copy()
    orc().orc_kernel_convolution(ctypes.byref(og(g)), fp(out), ctypes.c_float(sigma), kernel, ip(mask))
    return out


def orc_lncc(g, refv, war, mask=None, sigma=-5.0):
    orc().orc_lncc_value.restype = ctypes.c_double
    return orc().orc_lncc_value(ctypes.byref(og(g)), fp(refv), fp(war), ip(mask), ctypes.c_float(sigma))


def orc_lncc_grad(g, refv, war, wgrad, mask=None, sigma=-5.0, weight=1.0, init=None):
    grad = np.zeros(3 * g.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    orc().orc_lncc_gradient(ctypes.byref(og(g)), fp(refv), fp(war), fp(wgrad), ip(mask), ctypes.c_float(sigma),
                            ctypes.c_double(weight), fp(grad))
    return grad


def ref_lncc(g, refv, war, mask=None, sigma=-5.0):
    v = ctypes.c_double()
    rcheck(ref().ref_lncc_value(ctypes.byref(rg(g)), fp(refv), fp(war), ip(mask), ctypes.c_float(sigma),
                                ctypes.byref(v)))
    return v.value


def ref_lncc_grad(g, refv, war, wgrad, mask=None, sigma=-5.0, weight=1.0, init=None):
    grad = np.zeros(3 * g.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    rcheck(ref().ref_lncc_gradient(ctypes.byref(rg(g)), fp(refv), fp(war), fp(wgrad), ip(mask), ctypes.c_float(sigma),
                                   ctypes.c_double(weight), fp(grad)))
    return grad


# ---------------------------------------------------------------- HIP platform plugin check
PLUGIN = os.path.join(ROOT, "oracle", "_ref", "libnrplugin.so")
_plugin = None


def plugin_available():
    return os.path.exists(PLUGIN)


def plugin_f3d_eval(hip, refg, ref_img, flog, flo_img, cppg, cpp, bins=64, lncc=False, simw=0.999, bew=0.001,
                    lew=0.01):
    """One objective value + gradient through the reference's own Content / Compute / Measure
    calls, on the reference CPU classes (hip=False) or the HIP platform plugin (hip=True)."""
    global _plugin
    if _plugin is None:
        _plugin = ctypes.CDLL(PLUGIN)
        _plugin.plugin_last_error.restype = ctypes.c_char_p
    m, be = ctypes.c_double(), ctypes.c_double()
    grad = np.zeros(3 * cppg.nvox, np.float32)
    vg = np.zeros(3 * refg.nvox, np.float32)
    rc = _plugin.plugin_f3d_eval(int(hip), ctypes.byref(rg(refg)), fp(np.ascontiguousarray(ref_img, np.float32).ravel()),
                                 ctypes.byref(rg(flog)), fp(np.ascontiguousarray(flo_img, np.float32).ravel()),
                                 ctypes.byref(rg(cppg)), fp(np.ascontiguousarray(cpp, np.float32)), bins, int(lncc),
                                 ctypes.c_float(simw), ctypes.c_float(bew), ctypes.c_float(lew), ctypes.byref(m),
                                 ctypes.byref(be),
                                 fp(grad), fp(vg))
    if rc != 0:
        raise RuntimeError(_plugin.plugin_last_error().decode())
    return m.value, be.value, grad, vg


def plugin_f3d_time(hip, refg, ref_img, flog, flo_img, cppg, cpp, iters=5, bins=64):
    """Seconds for `iters` objective + gradient evaluations through the reference's own calls on
    the CPU classes (hip=False) or the HIP plugin (hip=True), after one untimed evaluation."""
    global _plugin
    if _plugin is None:
        _plugin = ctypes.CDLL(PLUGIN)
        _plugin.plugin_last_error.restype = ctypes.c_char_p
    sec = ctypes.c_double()
    rc = _plugin.plugin_f3d_time(int(hip), ctypes.byref(rg(refg)), fp(np.ascontiguousarray(ref_img, np.float32).ravel()),
                                 ctypes.byref(rg(flog)), fp(np.ascontiguousarray(flo_img, np.float32).ravel()),
                                 ctypes.byref(rg(cppg)), fp(np.ascontiguousarray(cpp, np.float32)), bins, int(iters),
                                 ctypes.byref(sec))
    if rc != 0:
        raise RuntimeError(_plugin.plugin_last_error().decode())
    return sec.value


# ---------------------------------------------------------------- linear energy
def orc_round_trip(cppg, cpp):
    """The control points after the reference's bending-energy gradient: (p - node) + node."""
    c = np.ascontiguousarray(cpp, np.float32).copy()
    orc().orc_displacement_round_trip(ctypes.byref(og(cppg)), fp(c))
    return c


def orc_le(cppg, cpp):
    orc().orc_linear_energy.restype = ctypes.c_double
    return orc().orc_linear_energy(ctypes.byref(og(cppg)), fp(cpp))


def orc_le_grad(cppg, cpp, weight, init=None):
    g = np.zeros(3 * cppg.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    orc().orc_linear_energy_gradient(ctypes.byref(og(cppg)), fp(cpp), ctypes.c_float(weight), fp(g))
    return g


def ref_le(cppg, cpp):
    v = ctypes.c_double()
    rcheck(ref().ref_le_value(ctypes.byref(rg(cppg)), fp(cpp), ctypes.byref(v)))
    return v.value


def ref_le_grad(cppg, cpp, weight, init=None):
    g = np.zeros(3 * cppg.nvox, np.float32) if init is None else init.astype(np.float32).copy()
    rcheck(ref().ref_le_gradient(ctypes.byref(rg(cppg)), fp(cpp), ctypes.c_float(weight), fp(g)))
    return g


# ---------------------------------------------------------------- symmetric Aladin matrices
def orc_nifti_inverse(A):
    out = np.zeros(16, np.float32)
    orc().orc_nifti_mat44_inverse(fp(np.ascontiguousarray(A, np.float32).ravel()), fp(out))
    return out.reshape(4, 4)


def orc_avg2(A, B):
    out = np.zeros(16, np.float32)
    ok = orc().orc_mat44_avg2(fp(np.ascontiguousarray(A, np.float32).ravel()),
                              fp(np.ascontiguousarray(B, np.float32).ravel()), fp(out))
    assert ok, "matrix logarithm failed"
    return out.reshape(4, 4)


def orc_symmetric_update(fw, bw):
    f = np.ascontiguousarray(fw, np.float32).ravel().copy()
    b = np.ascontiguousarray(bw, np.float32).ravel().copy()
    assert orc().orc_symmetric_update(fp(f), fp(b)), "matrix logarithm failed"
    return f.reshape(4, 4), b.reshape(4, 4)


# ---------------------------------------------------------------- file formats (niftilib)
class ref_nifti_fields(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int * 8), ("pixdim", ctypes.c_float * 8), ("datatype", ctypes.c_int),
                ("qform_code", ctypes.c_int), ("sform_code", ctypes.c_int), ("quatern", ctypes.c_float * 3),
                ("qoffset", ctypes.c_float * 3), ("qfac", ctypes.c_float), ("qto_xyz", ctypes.c_float * 16),
                ("sto_xyz", ctypes.c_float * 16), ("scl_slope", ctypes.c_float), ("scl_inter", ctypes.c_float),
                ("intent_code", ctypes.c_int), ("intent_p1", ctypes.c_float), ("intent_name", ctypes.c_char * 16)]


def ref_nifti_read(path, with_data=True):
    f = ref_nifti_fields()
    data = None
    if with_data:
        f0 = ref_nifti_fields()
        rcheck(ref().ref_nifti_read(str(path).encode(), ctypes.byref(f0), None, ctypes.c_size_t(0)))
        n = int(np.prod([f0.dim[i] for i in range(1, 8)]))
        data = np.zeros(n, np.float32)
        rcheck(ref().ref_nifti_read(str(path).encode(), ctypes.byref(f), fp(data), ctypes.c_size_t(n)))
    else:
        rcheck(ref().ref_nifti_read(str(path).encode(), ctypes.byref(f), None, ctypes.c_size_t(0)))
    return f, data


def ref_control_point_grid(ref_path, spacing, grid_path=None):
    f = ref_nifti_fields()
    sp = (ctypes.c_float * 3)(*[float(v) for v in spacing])
    rcheck(ref().ref_control_point_grid(str(ref_path).encode(), sp,
                                        None if grid_path is None else str(grid_path).encode(), ctypes.byref(f)))
    return f


def ref_affine_write(path, m):
    rcheck(ref().ref_affine_write(str(path).encode(), fp(np.ascontiguousarray(m, np.float32).ravel())))


def ref_affine_read(path):
    m = np.zeros(16, np.float32)
    rcheck(ref().ref_affine_read(str(path).encode(), fp(m)))
    return m.reshape(4, 4)


# ---------------------------------------------------------------- velocity fields (F3D2)
REF_SCALAR = os.path.join(ROOT, "oracle", "_ref", "libnrref_nosse.so")
_ref_scalar = None


REF_HIP = os.path.join(ROOT, "oracle", "_ref", "libnrref_hip.so")
_ref_hip = None
PLATFORM_CPU, PLATFORM_HIP = 0, 3


def ref_hip_available():
    return os.path.exists(REF_HIP)


def ref_hip():
    """The default reference build with the HIP platform (PlatformType 3) in its Platform.cpp."""
    global _ref_hip
    if _ref_hip is None:
        _ref_hip = ctypes.CDLL(REF_HIP)
        _ref_hip.ref_last_error.restype = ctypes.c_char_p
    return _ref_hip


def ref_scalar():
    """The reference built without USE_SSE: its summation order is the one the oracle follows."""
    global _ref_scalar
    if _ref_scalar is None:
        _ref_scalar = ctypes.CDLL(REF_SCALAR)
        _ref_scalar.ref_last_error.restype = ctypes.c_char_p
    return _ref_scalar


def _scheck(lib, rc):
    if rc != 0:
        raise RuntimeError(lib.ref_last_error().decode())


def orc_vel_flow(refg, cppg, vel):
    flow = np.zeros(3 * refg.nvox, np.float32)
    orc().orc_vel_flow(ctypes.byref(og(refg)), ctypes.byref(og(cppg)), fp(vel), fp(flow))
    return flow


def orc_vel_def_field(refg, cppg, vel, steps, update=False):
    field = np.zeros(3 * refg.nvox, np.float32)
    s = ctypes.c_int(int(steps))
    orc().orc_vel_def_field(ctypes.byref(og(refg)), ctypes.byref(og(cppg)), fp(vel), ctypes.byref(s),
                            int(bool(update)), fp(field))
    return field, s.value


def ref_vel_def_field(refg, cppg, vel, steps, update=False, scalar=True):
    lib = ref_scalar() if scalar else ref()
    field = np.zeros(3 * refg.nvox, np.float32)
    s = ctypes.c_int(int(steps))
    _scheck(lib, lib.ref_vel_def_field(ctypes.byref(rg(refg)), ctypes.byref(rg(cppg)), fp(vel), ctypes.byref(s),
                                       int(bool(update)), fp(field)))
    return field, s.value


def orc_vel_intermediate(refg, cppg, vel, steps):
    out = np.zeros((abs(steps) + 1) * 3 * refg.nvox, np.float32)
    orc().orc_vel_intermediate(ctypes.byref(og(refg)), ctypes.byref(og(cppg)), fp(vel), int(steps), fp(out))
    return out.reshape(abs(steps) + 1, 3 * refg.nvox)


def ref_vel_intermediate(refg, cppg, vel, steps):
    lib = ref_scalar()
    out = np.zeros((abs(steps) + 1) * 3 * refg.nvox, np.float32)
    _scheck(lib, lib.ref_vel_intermediate(ctypes.byref(rg(refg)), ctypes.byref(rg(cppg)), fp(vel), int(steps),
                                          fp(out)))
    return out.reshape(abs(steps) + 1, 3 * refg.nvox)


def orc_def_compose(g, d, upd):
    u = np.ascontiguousarray(upd, np.float32).copy()
    orc().orc_def_field_compose(ctypes.byref(og(g)), fp(np.ascontiguousarray(d, np.float32)), fp(u))
    return u


def ref_def_compose(g, d, upd):
    lib = ref_scalar()
    u = np.ascontiguousarray(upd, np.float32).copy()
    _scheck(lib, lib.ref_def_field_compose(ctypes.byref(rg(g)), fp(np.ascontiguousarray(d, np.float32)), fp(u)))
    return u


def orc_resample_gradient(g, grad, d, pad=0.0):
    out = np.zeros(3 * g.nvox, np.float32)
    orc().orc_resample_gradient(ctypes.byref(og(g)), fp(grad), fp(d), ctypes.c_float(pad), fp(out))
    return out


def ref_resample_gradient(g, grad, d, pad=0.0):
    lib = ref_scalar()
    out = np.zeros(3 * g.nvox, np.float32)
    _scheck(lib, lib.ref_resample_gradient(ctypes.byref(rg(g)), fp(grad), fp(d), ctypes.c_float(pad), fp(out)))
    return out


def orc_exponentiate_gradient(refg, cppg, vel_bw, steps, grad):
    g = np.ascontiguousarray(grad, np.float32).copy()
    orc().orc_exponentiate_gradient(ctypes.byref(og(refg)), ctypes.byref(og(cppg)), fp(vel_bw), int(steps), fp(g))
    return g


def ref_exponentiate_gradient(refg, flog, cppg, vel_fw, vel_bw, steps, grad):
    lib = ref_scalar()
    g = np.ascontiguousarray(grad, np.float32).copy()
    _scheck(lib, lib.ref_exponentiate_gradient(ctypes.byref(rg(refg)), ctypes.byref(rg(flog)), ctypes.byref(rg(cppg)),
                                               fp(vel_fw), fp(vel_bw), int(steps), fp(g)))
    return g


def orc_update_velocity(grid, grad, scale, axes=(1, 1, 1)):
    g = np.ascontiguousarray(grid, np.float32).copy()
    orc().orc_update_velocity(fp(g), fp(grad), ctypes.c_float(scale), ctypes.c_size_t(g.size // 3), *map(int, axes))
    return g


def orc_symmetrise_velocity(cppg, fw, bw):
    f = np.ascontiguousarray(fw, np.float32).copy()
    b = np.ascontiguousarray(bw, np.float32).copy()
    orc().orc_symmetrise_velocity(ctypes.byref(og(cppg)), fp(f), fp(b))
    return f, b


def ref_velocity_update(refg, flog, cppg, fw, bw, grad=None, scale=0.0, axes=(1, 1, 1), mode=0):
    lib = ref_scalar()
    f = np.ascontiguousarray(fw, np.float32).copy()
    b = np.ascontiguousarray(bw, np.float32).copy()
    gr = np.zeros_like(f) if grad is None else np.ascontiguousarray(grad, np.float32)
    ax = (ctypes.c_int * 3)(*map(int, axes))
    _scheck(lib, lib.ref_velocity_update(ctypes.byref(rg(refg)), ctypes.byref(rg(flog)), ctypes.byref(rg(cppg)),
                                         fp(f), fp(b), fp(gr), ctypes.c_float(scale), ax, int(mode)))
    return f, b


def orc_symmetric_grids(refg, flog, spacing):
    from niftyreg_amd.geometry import Geom
    sp = np.asarray(spacing, np.float32)
    o = orc_geom()
    orc().orc_symmetric_grids(ctypes.byref(og(refg)), ctypes.byref(og(flog)), fp(sp), ctypes.byref(o), None, None)
    G = o.nx * o.ny * o.nz
    fw = np.zeros(3 * G, np.float32)
    bw = np.zeros(3 * G, np.float32)
    orc().orc_symmetric_grids(ctypes.byref(og(refg)), ctypes.byref(og(flog)), fp(sp), ctypes.byref(o), fp(fw), fp(bw))
    g = Geom((o.nx, o.ny, o.nz), np.array(o.xyz[:], np.float32).reshape(4, 4),
             np.array(o.ijk[:], np.float32).reshape(4, 4), (o.dx, o.dy, o.dz))
    return g, fw, bw


def ref_symmetric_grids(refg, flog, spacing):
    lib = ref_scalar()
    sp = np.asarray(spacing, np.float32)
    o = nr_geom()
    _scheck(lib, lib.ref_symmetric_grids(ctypes.byref(rg(refg)), ctypes.byref(rg(flog)), fp(sp), ctypes.byref(o),
                                         None, None, 0))
    G = o.nx * o.ny * o.nz
    fw = np.zeros(3 * G, np.float32)
    bw = np.zeros(3 * G, np.float32)
    _scheck(lib, lib.ref_symmetric_grids(ctypes.byref(rg(refg)), ctypes.byref(rg(flog)), fp(sp), ctypes.byref(o),
                                         fp(fw), fp(bw), 3 * G))
    return o, fw, bw


class RefF3d2:
    """reg_f3d2<float> (symmetric, velocity grids) prepared at one level for single evaluations.

    scalar=True uses the reference built without USE_SSE (the oracle's summation order)."""

    def __init__(self, refg, ref_img, flog, flo_img, bins=64, be=0.001, le=0.0, spacing=-5.0, levels=1,
                 level=0, maxit=150, padding=float("nan"), scalar=True, measure=0, lncc_sigma=-5.0, affine=None):
        p = ref_f3d_params()
        p.bins, p.be, p.le, p.spacing = bins, be, le, spacing
        p.levels, p.levels_to_perform, p.maxit, p.use_pyramid, p.padding = levels, levels, maxit, 1, padding
        p.measure, p.lncc_sigma = measure, lncc_sigma
        self._affine = None if affine is None else np.ascontiguousarray(affine, np.float32).reshape(16)
        p.affine = None if affine is None else fp(self._affine)
        self.p = p
        self.lib = r = ref_scalar() if scalar else ref()
        r.ref_f3d2_create.restype = ctypes.c_void_p
        r.ref_f3d2_create.argtypes = [ctypes.c_void_p] * 6
        h = r.ref_f3d2_create(ctypes.byref(rg(refg)), fp(np.ascontiguousarray(ref_img, np.float32).reshape(-1)),
                              ctypes.byref(rg(flog)), fp(np.ascontiguousarray(flo_img, np.float32).reshape(-1)),
                              ctypes.byref(p), ctypes.c_void_p(level))
        if not h:
            raise RuntimeError(r.ref_last_error().decode())
        self.h = ctypes.c_void_p(h)
        for n in ("ref_f3d2_destroy", "ref_f3d2_level_geom", "ref_f3d2_level_images", "ref_f3d2_set_cpp",
                  "ref_f3d2_get_cpp", "ref_f3d2_value", "ref_f3d2_gradient", "ref_f3d2_weights"):
            getattr(r, n).argtypes = None
        cg, rgm, fgm = nr_geom(), nr_geom(), nr_geom()
        _scheck(r, r.ref_f3d2_level_geom(self.h, ctypes.byref(cg), ctypes.byref(rgm), ctypes.byref(fgm)))
        self.cpp_geom, self.ref_geom, self.flo_geom = cg, rgm, fgm
        self.G = cg.nx * cg.ny * cg.nz

    def images(self):
        nr_ = self.ref_geom.nx * self.ref_geom.ny * self.ref_geom.nz
        nf = self.flo_geom.nx * self.flo_geom.ny * self.flo_geom.nz
        a, b = np.zeros(nr_, np.float32), np.zeros(nf, np.float32)
        _scheck(self.lib, self.lib.ref_f3d2_level_images(self.h, fp(a), fp(b)))
        return a, b

    def weights(self):
        w = np.zeros(3, np.float32)
        _scheck(self.lib, self.lib.ref_f3d2_weights(self.h, fp(w)))
        return float(w[0]), float(w[1]), float(w[2])

    def get_cpp(self):
        f, b = np.zeros(3 * self.G, np.float32), np.zeros(3 * self.G, np.float32)
        s = np.zeros(2, np.int32)
        _scheck(self.lib, self.lib.ref_f3d2_get_cpp(self.h, fp(f), fp(b), ip(s)))
        return f, b, (int(s[0]), int(s[1]))

    def set_cpp(self, fw, bw):
        _scheck(self.lib, self.lib.ref_f3d2_set_cpp(self.h, fp(np.ascontiguousarray(fw, np.float32)),
                                                    fp(np.ascontiguousarray(bw, np.float32))))

    def value(self):
        v = ctypes.c_double()
        t = np.zeros(3, np.float64)
        _scheck(self.lib, self.lib.ref_f3d2_value(self.h, ctypes.byref(v), dp(t)))
        return v.value, t

    def gradient(self, normalise=False):
        gf, gb = np.zeros(3 * self.G, np.float32), np.zeros(3 * self.G, np.float32)
        m = ctypes.c_double()
        _scheck(self.lib, self.lib.ref_f3d2_gradient(self.h, int(normalise), ctypes.byref(m), fp(gf), fp(gb)))
        return gf, gb, m.value

    def close(self):
        if self.h:
            self.lib.ref_f3d2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def plugin_f3d2_eval(hip, refg, ref_img, flog, flo_img, cppg, fw, bw, bins=64, simw=0.989, bew=0.001, lew=0.01,
                     lncc=False, affine=None, halves=None):
    """One reg_f3d2 objective value + gradient through the reference's own Content / Compute /
    Measure calls (forward and backward contents), on the reference CPU classes (hip=False) or
    the HIP platform plugin (hip=True).  Returns (measure, be, grad_fw, grad_bw, cpp_fw, cpp_bw, steps)."""
    global _plugin
    if _plugin is None:
        _plugin = ctypes.CDLL(PLUGIN)
        _plugin.plugin_last_error.restype = ctypes.c_char_p
    m, be = ctypes.c_double(), ctypes.c_double()
    f = np.ascontiguousarray(fw, np.float32).copy()
    b = np.ascontiguousarray(bw, np.float32).copy()
    gf = np.zeros(3 * cppg.nvox, np.float32)
    gb = np.zeros(3 * cppg.nvox, np.float32)
    steps = np.zeros(2, np.int32)
    rc = _plugin.plugin_f3d2_eval(int(hip), ctypes.byref(rg(refg)), fp(np.ascontiguousarray(ref_img, np.float32).ravel()),
                                  ctypes.byref(rg(flog)), fp(np.ascontiguousarray(flo_img, np.float32).ravel()),
                                  ctypes.byref(rg(cppg)), fp(f), fp(b), bins, int(lncc), ctypes.c_float(simw),
                                  ctypes.c_float(bew),
                                  ctypes.c_float(lew), ctypes.byref(m), ctypes.byref(be), fp(gf), fp(gb), ip(steps),
                                  _fpn(_m16(affine)), *[_fpn(_m16(h)) for h in (halves or (None, None))])
    if rc != 0:
        raise RuntimeError(_plugin.plugin_last_error().decode())
    return m.value, be.value, gf, gb, f, b, (int(steps[0]), int(steps[1]))


def ref_f3d2_run(refg, ref_img, flog, flo_img, bins=64, be=0.001, le=0.01, spacing=-5.0, levels=3, maxit=150,
                 padding=float("nan"), affine=None, scalar=False, platform=None):
    """reg_f3d2<float>'s run on the reference CPU build (scalar=True: the build without USE_SSE), or
    with platform=PLATFORM_HIP on the HIP-platform build; returns (grid geom, fw, bw, steps)."""
    p = ref_f3d_params()
    p.bins, p.be, p.le, p.spacing = bins, be, le, spacing
    p.levels, p.levels_to_perform, p.maxit, p.use_pyramid, p.padding = levels, levels, maxit, 1, padding
    p.measure, p.lncc_sigma = 0, -5.0
    aff = None if affine is None else np.ascontiguousarray(affine, np.float32).reshape(16)
    p.affine = None if aff is None else fp(aff)
    out = nr_geom()
    cap = 3 * 200 ** 3 // 8
    fw = np.zeros(cap, np.float32)
    bw = np.zeros(cap, np.float32)
    steps = np.zeros(2, np.int32)
    args = (ctypes.byref(rg(refg)), fp(np.ascontiguousarray(ref_img, np.float32).reshape(-1)), ctypes.byref(rg(flog)),
            fp(np.ascontiguousarray(flo_img, np.float32).reshape(-1)), ctypes.byref(p), ctypes.byref(out), fp(fw),
            fp(bw), cap, ip(steps))
    if platform is not None:
        lib = ref_hip()
        _scheck(lib, lib.ref_f3d2_run_on(int(platform), *args))
    else:
        lib = ref_scalar() if scalar else ref()
        _scheck(lib, lib.ref_f3d2_run(*args))
    n = 3 * out.nx * out.ny * out.nz
    return geom_from_nr(out), fw[:n].copy(), bw[:n].copy(), (int(steps[0]), int(steps[1]))


def nmi_sym_at(refg, flog, cppg, fw, bw, ref_img, flo_img, bins=64):
    """NMI(reference, floating warped by exp(fw)) + NMI(floating, reference warped by exp(bw)) on
    the oracle, each exponentiation's step number chosen from its flow (images rescaled here)."""
    rb = bins + 4
    r = orc_rescale(np.asarray(ref_img, np.float32).ravel(), 2, rb - 3)
    f = orc_rescale(np.asarray(flo_img, np.float32).ravel(), 2, rb - 3)
    ff, _ = orc_vel_def_field(refg, cppg, np.ascontiguousarray(fw, np.float32), 6, True)
    fb, _ = orc_vel_def_field(flog, cppg, np.ascontiguousarray(bw, np.float32), 6, True)
    return orc_nmi(r, orc_resample(flog, f, refg, ff), rb, rb)[0] + orc_nmi(f, orc_resample(refg, r, flog, fb), rb, rb)[0]


# ---------------------------------------------------------------- velocity grids with affine parts
def _m16(m):
    return None if m is None else np.ascontiguousarray(m, np.float32).reshape(16)


def _fpn(a):
    return None if a is None else fp(a)


def orc_half_affines(A):
    hf, hb = np.zeros(16, np.float32), np.zeros(16, np.float32)
    ext = orc().orc_half_affines(fp(_m16(A)), fp(hf), fp(hb))
    assert ext >= 0
    return hf.reshape(4, 4), hb.reshape(4, 4), bool(ext)


def orc_symmetric_grids_aff(refg, flog, spacing, A):
    from niftyreg_amd.geometry import Geom
    sp = np.asarray(spacing, np.float32)
    a = _m16(A)
    o = orc_geom()
    orc().orc_symmetric_grids_aff(ctypes.byref(og(refg)), ctypes.byref(og(flog)), fp(sp), fp(a), ctypes.byref(o),
                                  None, None)
    G = o.nx * o.ny * o.nz
    fw, bw = np.zeros(3 * G, np.float32), np.zeros(3 * G, np.float32)
    orc().orc_symmetric_grids_aff(ctypes.byref(og(refg)), ctypes.byref(og(flog)), fp(sp), fp(a), ctypes.byref(o),
                                  fp(fw), fp(bw))
    g = Geom((o.nx, o.ny, o.nz), np.array(o.xyz[:], np.float32).reshape(4, 4),
             np.array(o.ijk[:], np.float32).reshape(4, 4), (o.dx, o.dy, o.dz))
    return g, fw, bw


def ref_symmetric_grids_aff(refg, flog, spacing, A):
    """(geometry, fw, bw, extension 0 of the forward grid, of the backward grid, extension count)"""
    lib = ref_scalar()
    sp = np.asarray(spacing, np.float32)
    a = _m16(A)
    o = nr_geom()
    ef, eb, cnt = np.zeros(16, np.float32), np.zeros(16, np.float32), ctypes.c_int()
    _scheck(lib, lib.ref_symmetric_grids_aff(ctypes.byref(rg(refg)), ctypes.byref(rg(flog)), fp(sp), fp(a),
                                             ctypes.byref(o), None, None, 0, fp(ef), fp(eb), ctypes.byref(cnt)))
    G = o.nx * o.ny * o.nz
    fw, bw = np.zeros(3 * G, np.float32), np.zeros(3 * G, np.float32)
    _scheck(lib, lib.ref_symmetric_grids_aff(ctypes.byref(rg(refg)), ctypes.byref(rg(flog)), fp(sp), fp(a),
                                             ctypes.byref(o), fp(fw), fp(bw), 3 * G, fp(ef), fp(eb), ctypes.byref(cnt)))
    return o, fw, bw, ef.reshape(4, 4), eb.reshape(4, 4), cnt.value


def orc_vel_def_field_aff(refg, cppg, vel, ext0, ext1, steps, update=False):
    field = np.zeros(3 * refg.nvox, np.float32)
    s = ctypes.c_int(int(steps))
    e0, e1 = _m16(ext0), _m16(ext1)
    orc().orc_vel_def_field_aff(ctypes.byref(og(refg)), ctypes.byref(og(cppg)), fp(vel), _fpn(e0), _fpn(e1),
                                ctypes.byref(s), int(bool(update)), fp(field))
    return field, s.value


def ref_vel_def_field_aff(refg, cppg, vel, ext0, ext1, steps, update=False):
    lib = ref_scalar()
    field = np.zeros(3 * refg.nvox, np.float32)
    s = ctypes.c_int(int(steps))
    e0, e1 = _m16(ext0), _m16(ext1)
    _scheck(lib, lib.ref_vel_def_field_aff(ctypes.byref(rg(refg)), ctypes.byref(rg(cppg)), fp(vel), _fpn(e0),
                                           _fpn(e1), ctypes.byref(s), int(bool(update)), fp(field)))
    return field, s.value


def orc_vel_intermediate_aff(refg, cppg, vel, ext0, ext1, steps):
    out = np.zeros((abs(steps) + 1) * 3 * refg.nvox, np.float32)
    e0, e1 = _m16(ext0), _m16(ext1)
    orc().orc_vel_intermediate_aff(ctypes.byref(og(refg)), ctypes.byref(og(cppg)), fp(vel), _fpn(e0), _fpn(e1),
                                   int(steps), fp(out))
    return out.reshape(abs(steps) + 1, -1)


def ref_vel_intermediate_aff(refg, cppg, vel, ext0, ext1, steps):
    lib = ref_scalar()
    out = np.zeros((abs(steps) + 1) * 3 * refg.nvox, np.float32)
    e0, e1 = _m16(ext0), _m16(ext1)
    _scheck(lib, lib.ref_vel_intermediate_aff(ctypes.byref(rg(refg)), ctypes.byref(rg(cppg)), fp(vel), _fpn(e0),
                                              _fpn(e1), int(steps), fp(out)))
    return out.reshape(abs(steps) + 1, -1)


def orc_exponentiate_gradient_aff(refg, cppg, vel_bw, ext0, ext1, trans_bw, steps, grad):
    g = np.ascontiguousarray(grad, np.float32).copy()
    e0, e1, t = _m16(ext0), _m16(ext1), _m16(trans_bw)
    orc().orc_exponentiate_gradient_aff(ctypes.byref(og(refg)), ctypes.byref(og(cppg)), fp(vel_bw), _fpn(e0),
                                        _fpn(e1), _fpn(t), int(steps), fp(g))
    return g


def ref_exponentiate_gradient_aff(refg, flog, cppg, vel_fw, vel_bw, ext0, ext1, trans_bw, steps, grad):
    lib = ref_scalar()
    g = np.ascontiguousarray(grad, np.float32).copy()
    e0, e1, t = _m16(ext0), _m16(ext1), _m16(trans_bw)
    _scheck(lib, lib.ref_exponentiate_gradient_aff(ctypes.byref(rg(refg)), ctypes.byref(rg(flog)),
                                                   ctypes.byref(rg(cppg)), fp(vel_fw), fp(vel_bw), _fpn(e0), _fpn(e1),
                                                   _fpn(t), int(steps), fp(g)))
    return g


def orc_voxel_to_node_ext(cppg, refg, vg, flog, ext0, weight):
    ng = np.zeros(3 * cppg.nvox, np.float32)
    ijk = np.ascontiguousarray(flog.ijk, np.float32).reshape(16)
    orc().orc_voxel_to_node_ext(ctypes.byref(og(cppg)), ctypes.byref(og(refg)), fp(vg), fp(ijk), fp(_m16(ext0)),
                                ctypes.c_float(weight), fp(ng))
    return ng


def ref_voxel_to_node_ext(cppg, refg, vg, flog, ext0, weight):
    lib = ref_scalar()
    ng = np.zeros(3 * cppg.nvox, np.float32)
    _scheck(lib, lib.ref_voxel_to_node_ext(ctypes.byref(rg(cppg)), ctypes.byref(rg(refg)), ctypes.byref(rg(flog)),
                                           fp(np.ascontiguousarray(vg, np.float32)), fp(_m16(ext0)),
                                           ctypes.c_float(weight), fp(ng)))
    return ng
