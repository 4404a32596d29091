"""F3D objective / gradient on the GPU through libnrhip's C ABI.

F3dContent mirrors the reference's F3dContent + Compute pair for one image pair
(reg-lib/F3dContent.h, reg-lib/Compute.h): the images, control points and every
intermediate stay on the device; methods named after the reference's Compute / reg_nmi
virtuals run one operation, objective() / gradient() run the fused sweeps.
"""
import ctypes

import numpy as np

from . import _lib
from .geometry import Geom


def default_params(bins=64, padding=float("nan"), similarity_weight=0.999, be_weight=0.001, image_storage=0,
                   measure=_lib.MEASURE_NMI, lncc_sigma=-5.0, le_weight=0.0, field_order=_lib.FIELD_ORDER_SCALAR):
    """Handle parameters.  field_order: the reference build whose float-field rounding decides
    floors near voxel planes (FIELD_ORDER_SCALAR: USE_SSE off, the order of the oracle's
    restatement; FIELD_ORDER_SSE: the default x86 build)."""
    p = _lib.nrh_f3d_params()
    p.field_order = field_order
    p.measure = measure
    p.lncc_sigma = lncc_sigma
    p.ref_bins = bins + 4          # the reference adds 4 bins for the spline support
    p.flo_bins = bins + 4
    p.padding = padding
    p.similarity_weight = similarity_weight
    p.be_weight = be_weight
    p.le_weight = le_weight
    p.image_storage = image_storage
    return p


class F3dContent:
    def __init__(self, ref_geom: Geom, ref, flo_geom: Geom, flo, cpp_geom: Geom, cpp=None, mask=None, params=None):
        self.ref_geom, self.flo_geom, self.cpp_geom = ref_geom, flo_geom, cpp_geom
        self.N, self.Nf, self.G = ref_geom.nvox, flo_geom.nvox, cpp_geom.nvox
        self.params = params if params is not None else default_params()
        ref = np.ascontiguousarray(ref, np.float32).reshape(-1)
        flo = np.ascontiguousarray(flo, np.float32).reshape(-1)
        assert ref.size == self.N and flo.size == self.Nf
        m = None
        if mask is not None:
            m = np.ascontiguousarray(mask, np.int32).reshape(-1)
            assert m.size == self.N
        self._rg, self._fg, self._cg = ref_geom.to_c(), flo_geom.to_c(), cpp_geom.to_c()
        h = ctypes.c_void_p()
        _lib.call("nrh_f3d_create", ctypes.byref(self._rg), _lib.fptr(ref), ctypes.byref(self._fg), _lib.fptr(flo),
                  _lib.iptr(m) if m is not None else None, ctypes.byref(self._cg), ctypes.byref(self.params),
                  ctypes.byref(h))
        self.h = h
        if cpp is not None:
            self.set_control_points(cpp)

    def close(self):
        if self.h:
            _lib.call("nrh_f3d_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _op(self, name, *args):
        _lib.call(name, self.h, *args)

    def set_params(self, params):
        self.params = params
        self._op("nrh_f3d_set_params", ctypes.byref(params))

    # --- control points
    def set_control_points(self, cpp):
        cpp = np.ascontiguousarray(cpp, np.float32).reshape(-1)
        assert cpp.size == 3 * self.G
        self._op("nrh_f3d_set_control_points", _lib.fptr(cpp))

    def get_control_points(self):
        out = np.empty(3 * self.G, np.float32)
        self._op("nrh_f3d_get_control_points", _lib.fptr(out))
        return out

    def rescale_intensities(self, ref_range, flo_range):
        self._op("nrh_f3d_rescale_intensities", ctypes.c_float(ref_range[0]), ctypes.c_float(ref_range[1]),
                 ctypes.c_float(flo_range[0]), ctypes.c_float(flo_range[1]))

    def get_images(self):
        r = np.empty(self.N, np.float32)
        f = np.empty(self.Nf, np.float32)
        self._op("nrh_f3d_get_images", _lib.fptr(r), _lib.fptr(f))
        return r, f

    # --- per-operation (Compute / reg_nmi)
    def GetDeformationField(self):
        self._op("nrh_f3d_deformation_field")

    def ResampleImage(self):
        self._op("nrh_f3d_resample")

    def GetImageGradient(self):
        self._op("nrh_f3d_image_gradient")

    def GetSimilarityMeasureValue(self):
        nmi = ctypes.c_double()
        ent = np.zeros(4, np.float64)
        self._op("nrh_f3d_nmi_value", ctypes.byref(nmi), _lib.dptr(ent))
        return nmi.value, ent

    def GetLnccValue(self):
        """reg_lncc::GetSimilarityMeasureValueFw on the current warped image."""
        v = ctypes.c_double()
        self._op("nrh_f3d_lncc_value", ctypes.byref(v))
        return v.value

    def GetVoxelBasedLnccGradient(self, weight=1.0):
        """reg_lncc::GetVoxelBasedSimilarityMeasureGradientFw (voxel gradient -= ...)."""
        self._op("nrh_f3d_lncc_voxel_gradient", ctypes.c_double(weight))

    def ZeroVoxelBasedMeasureGradient(self):
        self._op("nrh_f3d_zero_voxel_gradient")

    def GetVoxelBasedSimilarityMeasureGradient(self, weight=1.0):
        self._op("nrh_f3d_nmi_voxel_gradient", ctypes.c_double(weight))

    def ConvolveImage(self):
        self._op("nrh_f3d_convolve_voxel_gradient")

    def VoxelCentricToNodeCentric(self, weight):
        self._op("nrh_f3d_voxel_to_node", ctypes.c_float(weight))

    def ZeroTransformationGradient(self):
        self._op("nrh_f3d_zero_node_gradient")

    def ApproxBendingEnergy(self):
        v = ctypes.c_double()
        self._op("nrh_f3d_bending_energy", ctypes.byref(v))
        return v.value

    def ApproxLinearEnergy(self):
        v = ctypes.c_double()
        self._op("nrh_f3d_linear_energy", ctypes.byref(v))
        return v.value

    def ApproxLinearEnergyGradient(self, weight):
        self._op("nrh_f3d_linear_energy_gradient", ctypes.c_float(weight))

    def ApproxBendingEnergyGradient(self, weight):
        self._op("nrh_f3d_bending_energy_gradient", ctypes.c_float(weight))

    def GetMaximalLength(self):
        v = ctypes.c_double()
        self._op("nrh_f3d_max_length", ctypes.byref(v))
        return v.value

    def NormaliseGradient(self, max_length):
        self._op("nrh_f3d_normalise_gradient", ctypes.c_double(max_length))

    # --- downloads
    # --- velocity grids (reg_f3d2; Compute's velocity virtuals)
    def SetVelocityGrid(self, steps=6):
        """Mark the control points as a velocity grid (reg_f3d2::Initialise, intent_p2 = steps)."""
        self._op("nrh_f3d_set_velocity_grid", int(steps))

    def set_velocity_affines(self, ext0=None, ext1=None, transformation=None):
        """The grid's extensions 0 / 1 and the content's transformation matrix (4x4 each or None)."""
        mats = [None if m is None else np.ascontiguousarray(m, np.float32).reshape(16) for m in (ext0, ext1, transformation)]
        self._keep_affines = mats
        self._op("nrh_f3d_set_velocity_affines", *[None if m is None else _lib.fptr(m) for m in mats])

    @property
    def velocity_steps(self):
        s = ctypes.c_int()
        self._op("nrh_f3d_get_velocity_steps", ctypes.byref(s))
        return s.value

    @velocity_steps.setter
    def velocity_steps(self, steps):
        self._op("nrh_f3d_set_velocity_steps", int(steps))

    def GetDefFieldFromVelocityGrid(self, update_step_number=True):
        self._op("nrh_f3d_def_field_from_velocity", int(bool(update_step_number)))

    def ExponentiateGradient(self, backward):
        self._op("nrh_f3d_exponentiate_gradient", backward.h)

    def UpdateVelocityField(self, scale, optimise_x=True, optimise_y=True, optimise_z=True):
        self._op("nrh_f3d_update_velocity_field", ctypes.c_float(scale), int(optimise_x), int(optimise_y),
                 int(optimise_z))

    def SymmetriseVelocityFields(self, backward):
        self._op("nrh_f3d_symmetrise_velocity_fields", backward.h)

    def BchUpdate(self, scale, bch_update_value):
        self._op("nrh_f3d_bch_update", ctypes.c_float(scale), int(bch_update_value))

    def DefFieldCompose(self, geom: Geom, field):
        f = np.ascontiguousarray(field, np.float32).reshape(-1)
        assert f.size == 3 * geom.nvox
        g = geom.to_c()
        self._op("nrh_f3d_compose_field", ctypes.byref(g), _lib.fptr(f))

    def ResampleGradient(self, interpolation=1, padding=0.0):
        out = np.empty(3 * self.N, np.float32)
        self._op("nrh_f3d_resample_gradient", int(interpolation), ctypes.c_float(padding), _lib.fptr(out))
        return out

    def set_deformation_field(self, field):
        f = np.ascontiguousarray(field, np.float32).reshape(-1)
        assert f.size == 3 * self.N
        self._op("nrh_f3d_upload_field", _lib.fptr(f))

    def set_voxel_gradient(self, grad):
        g = np.ascontiguousarray(grad, np.float32).reshape(-1)
        assert g.size == 3 * self.N
        self._op("nrh_f3d_upload_voxel_gradient", _lib.fptr(g))

    def _dl(self, name, n, dtype=np.float32):
        out = np.empty(n, dtype)
        self._op(name, _lib.fptr(out))
        return out

    def deformation_field(self):
        return self._dl("nrh_f3d_download_field", 3 * self.N)

    def warped(self):
        return self._dl("nrh_f3d_download_warped", self.N)

    def warped_gradient(self):
        return self._dl("nrh_f3d_download_warped_gradient", 3 * self.N)

    def voxel_gradient(self):
        return self._dl("nrh_f3d_download_voxel_gradient", 3 * self.N)

    def node_gradient(self):
        return self._dl("nrh_f3d_download_node_gradient", 3 * self.G)

    def set_node_gradient(self, g):
        g = np.ascontiguousarray(g, np.float32).reshape(-1)
        self._op("nrh_f3d_upload_node_gradient", _lib.fptr(g))

    def histogram(self):
        rb, fb = self.params.ref_bins, self.params.flo_bins
        tot = rb * fb + rb + fb
        pro = np.empty(tot, np.float64)
        logh = np.empty(tot, np.float64)
        self._op("nrh_f3d_download_histogram", _lib.dptr(pro), _lib.dptr(logh))
        return pro, logh

    # --- fused evaluation
    def fused_available(self):
        v = ctypes.c_int()
        self._op("nrh_f3d_fused_available", ctypes.byref(v))
        return bool(v.value)

    def set_profiling(self, enable=True, period=1):
        """HIP-event timing of every period-th fused sweep of each kind (enable=False stops it)."""
        self._op("nrh_f3d_set_profiling", ctypes.c_int(int(period) if enable else 0))

    def sweep_times(self):
        ms = np.zeros(2, np.float64)
        cnt = (ctypes.c_long * 2)()
        self._op("nrh_f3d_sweep_times", _lib.dptr(ms), cnt)
        return ms, np.array(cnt[:])

    def fused_field(self):
        self._op("nrh_f3d_fused_field")
        return self.deformation_field()

    def objective(self):
        v = ctypes.c_double()
        terms = np.zeros(3, np.float64)
        self._op("nrh_f3d_objective", ctypes.byref(v), _lib.dptr(terms))
        return v.value, terms

    def gradient(self, normalise=False):
        if normalise:
            m = ctypes.c_double()
            self._op("nrh_f3d_gradient", ctypes.byref(m))
            return m.value
        self._op("nrh_f3d_gradient", None)
        return None

    # --- optimiser state
    def store_best(self):
        self._op("nrh_f3d_store_best")

    def restore_best(self):
        self._op("nrh_f3d_restore_best")

    def conjugate_gradient(self, first):
        self._op("nrh_f3d_conjugate_gradient", ctypes.c_int(1 if first else 0))

    def update_control_points(self, scale):
        self._op("nrh_f3d_update_control_points", ctypes.c_float(scale))


def register(ref_geom: Geom, ref, flo_geom: Geom, flo, bins=64, be=0.001, le=0.0, spacing=-5.0, levels=3,
             levels_to_perform=0, maxit=150, padding=float("nan"), use_pyramid=True, image_storage=0,
             measure=_lib.MEASURE_NMI, lncc_sigma=-5.0, affine=None, field_order=_lib.FIELD_ORDER_SSE):
    """Full reg_f3d registration of one pair on the device (reg_base<float>::Run).

    affine: optional reference -> floating 4x4 initialisation (-aff).
    field_order: the reference build whose field rounding is reproduced (default: the SSE build,
    the reference's default on x86).
    Returns (grid Geom, control points planar float32, final similarity value, iterations)."""
    cfg = _lib.nrh_f3d_config()
    aff = None if affine is None else np.ascontiguousarray(affine, np.float32).reshape(16)
    cfg.affine = None if aff is None else _lib.fptr(aff)
    cfg.bins, cfg.be_weight, cfg.le_weight, cfg.spacing = bins, be, le, spacing
    cfg.levels, cfg.levels_to_perform, cfg.max_iterations = levels, levels_to_perform, maxit
    cfg.padding, cfg.use_pyramid, cfg.image_storage = padding, int(use_pyramid), image_storage
    cfg.measure, cfg.lncc_sigma = measure, lncc_sigma
    cfg.field_order = field_order
    ref = np.ascontiguousarray(ref, np.float32).reshape(-1)
    flo = np.ascontiguousarray(flo, np.float32).reshape(-1)
    rg, fg = ref_geom.to_c(), flo_geom.to_c()
    out = _lib.nrh_geom()
    # the final grid is at most the grid of the full-resolution image at the final spacing
    cap = 3 * int(np.prod([int(np.ceil(n * d / (abs(spacing) * d if spacing < 0 else spacing) + 4))
                           for n, d in zip(ref_geom.dims, ref_geom.pixdim)]))
    cpp = np.zeros(cap, np.float32)
    nmi = ctypes.c_double()
    its = ctypes.c_int()
    _lib.call("nrh_f3d_register", ctypes.byref(rg), _lib.fptr(ref), ctypes.byref(fg), _lib.fptr(flo), ctypes.byref(cfg),
              ctypes.byref(out), _lib.fptr(cpp), ctypes.c_size_t(cap), ctypes.byref(nmi), ctypes.byref(its))
    g = Geom.from_c(out)
    return g, cpp[:3 * g.nvox].copy(), nmi.value, its.value


class F3dBatch:
    """Many F3dContent handles evaluated by the same launches (nrh_batch_*).  The handles stay
    owned by the caller; while the batch exists their work runs on the batch's stream."""

    def __init__(self, contents):
        self.contents = list(contents)
        arr = (ctypes.c_void_p * len(self.contents))(*[c.h.value if isinstance(c.h, ctypes.c_void_p) else c.h
                                                       for c in self.contents])
        b = ctypes.c_void_p()
        _lib.call("nrh_batch_create", arr, len(self.contents), ctypes.byref(b))
        self.b = b

    def objective(self):
        """(values, nmi) of every pair."""
        n = len(self.contents)
        v, m = (ctypes.c_double * n)(), (ctypes.c_double * n)()
        _lib.call("nrh_batch_objective", self.b, v, m)
        return np.array(v[:]), np.array(m[:])

    def gradient(self, normalise=False):
        """Node gradients land in each handle (node_gradient()); returns the max lengths when
        normalising."""
        n = len(self.contents)
        m = (ctypes.c_double * n)()
        _lib.call("nrh_batch_gradient", self.b, m if normalise else None)
        return np.array(m[:]) if normalise else None

    def set_sweep_blocks(self, blocks_per_xcd):
        _lib.call("nrh_batch_set_sweep_blocks", self.b, int(blocks_per_xcd))

    def autotune(self):
        k = ctypes.c_int()
        ms = (ctypes.c_double * 4)()
        _lib.call("nrh_batch_autotune", self.b, ctypes.byref(k), ms)
        return k.value, list(ms)

    def close(self):
        if self.b:
            _lib.call("nrh_batch_destroy", self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def register_batch(pairs, bins=64, be=0.001, le=0.0, spacing=-5.0, levels=3, levels_to_perform=0, maxit=150,
                   padding=float("nan"), use_pyramid=True, image_storage=0, field_order=_lib.FIELD_ORDER_SSE):
    """reg_f3d registrations of many pairs at once (nrh_f3d_register_batch).

    pairs: list of (ref_geom, ref, flo_geom, flo).  Returns ([(grid Geom, control points, final
    NMI, iterations)] per pair, stats dict)."""
    n = len(pairs)
    cfg = _lib.nrh_f3d_config()
    cfg.affine = None
    cfg.bins, cfg.be_weight, cfg.le_weight, cfg.spacing = bins, be, le, spacing
    cfg.levels, cfg.levels_to_perform, cfg.max_iterations = levels, levels_to_perform, maxit
    cfg.padding, cfg.use_pyramid, cfg.image_storage = padding, int(use_pyramid), image_storage
    cfg.measure, cfg.lncc_sigma = _lib.MEASURE_NMI, -5.0
    cfg.field_order = field_order
    refs = (_lib.nrh_geom * n)(*[p[0].to_c() for p in pairs])
    flos = (_lib.nrh_geom * n)(*[p[2].to_c() for p in pairs])
    keep = [(np.ascontiguousarray(p[1], np.float32).reshape(-1), np.ascontiguousarray(p[3], np.float32).reshape(-1))
            for p in pairs]
    rdata = (ctypes.POINTER(ctypes.c_float) * n)(*[_lib.fptr(k[0]) for k in keep])
    fdata = (ctypes.POINTER(ctypes.c_float) * n)(*[_lib.fptr(k[1]) for k in keep])
    caps, outs = [], []
    for p in pairs:
        g = p[0]
        cap = 3 * int(np.prod([int(np.ceil(nv * d / (abs(spacing) * d if spacing < 0 else spacing) + 4))
                               for nv, d in zip(g.dims, g.pixdim)]))
        caps.append(cap)
        outs.append(np.zeros(cap, np.float32))
    cpp_out = (ctypes.POINTER(ctypes.c_float) * n)(*[_lib.fptr(o) for o in outs])
    capacity = (ctypes.c_size_t * n)(*caps)
    grids = (_lib.nrh_geom * n)()
    nmi = (ctypes.c_double * n)()
    its = (ctypes.c_int * n)()
    stats = _lib.nrh_batch_stats()
    _lib.call("nrh_f3d_register_batch", n, refs, rdata, flos, fdata, ctypes.byref(cfg), grids, cpp_out, capacity, nmi,
              its, ctypes.byref(stats))
    res = []
    for q in range(n):
        g = Geom.from_c(grids[q])
        res.append((g, outs[q][:3 * g.nvox].copy(), nmi[q], its[q]))
    st = {"seconds": stats.seconds, "device_seconds": stats.device_seconds, "graph_replays": stats.graph_replays,
          "blocks_per_xcd": list(stats.blocks_per_xcd[:levels]), "tile": list(stats.tile)}
    return res, st


class F3d2Pair:
    """reg_f3d2's forward / backward pair (reg-lib/_reg_f3d2.h): two F3dContent handles whose
    control points are velocity grids of one symmetric grid geometry, the backward handle with
    the images (and bin counts) swapped.  objective() / gradient() run the symmetric evaluation
    with the two directions on the two handles' streams."""

    def __init__(self, forward: F3dContent, backward: F3dContent, steps=6):
        self.fw, self.bw = forward, backward
        self.fw.SetVelocityGrid(steps)
        self.bw.SetVelocityGrid(steps)

    def objective(self):
        v = ctypes.c_double()
        t = (ctypes.c_double * 3)()
        _lib.call("nrh_f3d2_objective", self.fw.h, self.bw.h, ctypes.byref(v), t)
        return v.value, np.array(t[:])

    def gradient(self, normalise=False):
        m = ctypes.c_double()
        _lib.call("nrh_f3d2_gradient", self.fw.h, self.bw.h, ctypes.byref(m) if normalise else None)
        return m.value if normalise else None

    def close(self):
        self.fw.close()
        self.bw.close()


def register_symmetric(ref_geom: Geom, ref, flo_geom: Geom, flo, bins=64, be=0.001, le=0.01, spacing=-5.0,
                       levels=3, levels_to_perform=0, maxit=150, padding=float("nan"), use_pyramid=True,
                       measure=_lib.MEASURE_NMI, lncc_sigma=-5.0, affine=None, field_order=_lib.FIELD_ORDER_SSE):
    """Full reg_f3d2 registration of one pair on the device (symmetric velocity grids, NMI or LNCC).

    affine: optional reference -> floating 4x4 (-aff), split into half affines on both grids.
    Returns (grid Geom, forward grid, backward grid, final similarity sum of both directions,
    iterations)."""
    cfg = _lib.nrh_f3d_config()
    cfg.bins, cfg.be_weight, cfg.le_weight, cfg.spacing = bins, be, le, spacing
    cfg.levels, cfg.levels_to_perform, cfg.max_iterations = levels, levels_to_perform, maxit
    cfg.padding, cfg.use_pyramid, cfg.image_storage = padding, int(use_pyramid), 0
    cfg.measure, cfg.lncc_sigma = measure, lncc_sigma
    cfg.field_order = field_order
    aff = None if affine is None else np.ascontiguousarray(affine, np.float32).reshape(16)
    cfg.affine = None if aff is None else _lib.fptr(aff)
    ref = np.ascontiguousarray(ref, np.float32).reshape(-1)
    flo = np.ascontiguousarray(flo, np.float32).reshape(-1)
    rg, fg = ref_geom.to_c(), flo_geom.to_c()
    out = _lib.nrh_geom()
    # the final grid spans both images at the final spacing (mm or voxels of the mean voxel size)
    ext = [max(n * d, m * e) + 2 * abs(o) for n, d, m, e, o in
           zip(ref_geom.dims, ref_geom.pixdim, flo_geom.dims, flo_geom.pixdim,
               np.abs(ref_geom.xyz[:3, 3] - flo_geom.xyz[:3, 3]))]
    sp = [abs(spacing) * (d + e) / 2 if spacing < 0 else spacing for d, e in zip(ref_geom.pixdim, flo_geom.pixdim)]
    grow = 1.0 if affine is None else 2.0  # the half affines may rotate the bounding box
    cap = 3 * int(np.prod([int(np.ceil(grow * x / s + 8)) for x, s in zip(ext, sp)]))
    fw = np.zeros(cap, np.float32)
    bw = np.zeros(cap, np.float32)
    nmi = ctypes.c_double()
    its = ctypes.c_int()
    _lib.call("nrh_f3d2_register", ctypes.byref(rg), _lib.fptr(ref), ctypes.byref(fg), _lib.fptr(flo),
              ctypes.byref(cfg), ctypes.byref(out), _lib.fptr(fw), _lib.fptr(bw), ctypes.c_size_t(cap),
              ctypes.byref(nmi), ctypes.byref(its))
    g = Geom.from_c(out)
    return g, fw[:3 * g.nvox].copy(), bw[:3 * g.nvox].copy(), nmi.value, its.value
