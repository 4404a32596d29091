// C ABI of libnrhip: library state, host helpers and the F3D handle entry points.

#include "f3d_handle.hpp"
#include <cmath>
#include <memory>
#include <mutex>

namespace nrh {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }

void check_geom(const nrh_geom *g, const char *what) {
    NRH_REQUIRE(g != nullptr, NRH_ERR_ARG, std::string(what) + ": null geometry");
    NRH_REQUIRE(g->nx > 0 && g->ny > 0 && g->nz > 0, NRH_ERR_ARG, std::string(what) + ": non-positive dimension");
    NRH_REQUIRE(g->dx > 0 && g->dy > 0 && g->dz > 0, NRH_ERR_ARG, std::string(what) + ": non-positive voxel size");
}

// The inverse of an affine 4x4 with the adjugate expressions of nifti_mat44_inverse
// (reg-io/niftilib/nifti1_io.c:1680-1717), in double with the determinant's reciprocal in long
// double (as niftilib), stored in float.
void mat44_inverse(const float *R, float *Q) {
    const double r11 = R[0], r12 = R[1], r13 = R[2], v1 = R[3];
    const double r21 = R[4], r22 = R[5], r23 = R[6], v2 = R[7];
    const double r31 = R[8], r32 = R[9], r33 = R[10], v3 = R[11];
    double deti = r11 * r22 * r33 - r11 * r32 * r23 - r21 * r12 * r33 + r21 * r32 * r13 + r31 * r12 * r23 - r31 * r22 * r13;
    if (deti != 0.0) deti = (double)(1.0L / (long double)deti);
    Q[0] = (float)(deti * (r22 * r33 - r32 * r23));
    Q[1] = (float)(deti * (-r12 * r33 + r32 * r13));
    Q[2] = (float)(deti * (r12 * r23 - r22 * r13));
    Q[3] = (float)(deti * (-r12 * r23 * v3 + r12 * v2 * r33 + r22 * r13 * v3 - r22 * v1 * r33 - r32 * r13 * v2 + r32 * v1 * r23));
    Q[4] = (float)(deti * (-r21 * r33 + r31 * r23));
    Q[5] = (float)(deti * (r11 * r33 - r31 * r13));
    Q[6] = (float)(deti * (-r11 * r23 + r21 * r13));
    Q[7] = (float)(deti * (r11 * r23 * v3 - r11 * v2 * r33 - r21 * r13 * v3 + r21 * v1 * r33 + r31 * r13 * v2 - r31 * v1 * r23));
    Q[8] = (float)(deti * (r21 * r32 - r31 * r22));
    Q[9] = (float)(deti * (-r11 * r32 + r31 * r12));
    Q[10] = (float)(deti * (r11 * r22 - r21 * r12));
    Q[11] = (float)(deti * (-r11 * r22 * v3 + r11 * r32 * v2 + r21 * r12 * v3 - r21 * r32 * v1 - r31 * r12 * v2 + r31 * r22 * v1));
    Q[12] = Q[13] = Q[14] = 0.f;
    Q[15] = deti == 0.0 ? 0.f : 1.f;
}

void create_grid(const nrh_geom& ref, const float spacing[3], nrh_geom& grid, float *identity) {
    // reg_createControlPointGrid (reg-lib/cpu/_reg_localTrans.cpp:21-128)
    auto ceil_int = [](float x) { const int i = (int)x; return i + (x > (float)i); };
    grid = nrh_geom{};
    grid.nx = ceil_int((float)ref.nx * ref.dx / spacing[0] + 3.f);
    grid.ny = ceil_int((float)ref.ny * ref.dy / spacing[1] + 3.f);
    grid.nz = ref.nz > 1 ? ceil_int((float)ref.nz * ref.dz / spacing[2] + 3.f) : 1;
    grid.dx = spacing[0];
    grid.dy = spacing[1];
    grid.dz = ref.nz > 1 ? spacing[2] : 1.f;
    grid.sform = ref.sform;
    const float ratio[3] = { grid.dx / ref.dx, grid.dy / ref.dy, grid.dz / ref.dz };
    float m[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 3; ++c) m[r * 4 + c] = r < 3 ? ref.xyz[r * 4 + c] * ratio[c] : ref.xyz[r * 4 + c];
    for (int r = 0; r < 4; ++r) m[r * 4 + 3] = ref.xyz[r * 4 + 3];
    // origin shifted by one node: Mat44Mul (double accumulation) of (-1, -1, -1)
    const float o[3] = { -1.f, -1.f, ref.nz > 1 ? -1.f : 0.f };
    float origin[3];
    for (int r = 0; r < 3; ++r)
        origin[r] = (float)((double)m[r * 4] * o[0] + (double)m[r * 4 + 1] * o[1] + (double)m[r * 4 + 2] * o[2] +
                            (double)m[r * 4 + 3]);
    for (int r = 0; r < 3; ++r) m[r * 4 + 3] = origin[r];
    std::memcpy(grid.xyz, m, sizeof m);
    mat44_inverse(grid.xyz, grid.ijk);
    if (identity) {
        // reg_getDeformationFromDisplacement_3D of a zero displacement (reg-lib/cpu/_reg_tools.cpp:2107-2150)
        const size_t G = voxels(grid);
        size_t idx = 0;
        for (int z = 0; z < grid.nz; ++z)
            for (int y = 0; y < grid.ny; ++y)
                for (int x = 0; x < grid.nx; ++x, ++idx)
                    for (int r = 0; r < 3; ++r) {
                        volatile float v = m[r * 4] * (float)x;
                        v = v + m[r * 4 + 1] * (float)y;
                        v = v + m[r * 4 + 2] * (float)z;
                        v = v + m[r * 4 + 3];
                        identity[idx + r * G] = v;
                    }
    }
}

void ensure_field(nrh_f3d *h) { h->field.alloc(3 * h->N); }
void ensure_warped(nrh_f3d *h) { h->warped.alloc(h->N); }
void ensure_warped_grad(nrh_f3d *h) { h->warpedGrad.alloc(3 * h->N); }
void ensure_voxel_grad(nrh_f3d *h) {
    if (h->voxelGrad.empty()) {
        h->voxelGrad.alloc(3 * h->N);
        NRH_HIP(hipMemsetAsync(h->voxelGrad.p, 0, h->voxelGrad.bytes(), h->stream));
    }
}
const float *ref_ptr(nrh_f3d *h) {
    NRH_REQUIRE(!h->refImg.empty(), NRH_ERR_UNSUPPORTED, "operation needs float32 image storage");
    return h->refImg.p;
}

void alloc_histogram(nrh_f3d *h) {
    const int rb = h->params.ref_bins, fb = h->params.flo_bins;
    h->histCount.alloc((size_t)rb * fb);
    // every histogram accumulates into zeroed counts; each finalisation zeroes them again
    NRH_HIP(hipMemsetAsync(h->histCount.p, 0, h->histCount.bytes(), h->stream));
    NRH_HIP(hipStreamSynchronize(h->stream));
    h->histPro.alloc(h->totalBins());
    h->histLog.alloc(h->totalBins());
    h->histLogF.alloc(h->totalBins());
}

void check_params(const nrh_f3d_params *p) {
    NRH_REQUIRE(p != nullptr, NRH_ERR_ARG, "null parameters");
    NRH_REQUIRE(p->ref_bins >= 5 && p->flo_bins >= 5 && p->ref_bins <= 260 && p->flo_bins <= 260, NRH_ERR_ARG,
                "bin counts must lie in [5, 260]");
    NRH_REQUIRE(p->le_weight >= 0.f && p->be_weight >= 0.f, NRH_ERR_ARG, "penalty weights must be >= 0");
    NRH_REQUIRE(p->image_storage == 0 || p->image_storage == 1, NRH_ERR_ARG, "image_storage must be 0 or 1");
    NRH_REQUIRE(p->measure == NRH_MEASURE_NMI || p->measure == NRH_MEASURE_LNCC, NRH_ERR_ARG, "unknown measure");
    NRH_REQUIRE(p->measure != NRH_MEASURE_LNCC || p->lncc_sigma != 0.f, NRH_ERR_ARG, "LNCC needs a non-zero sigma");
    NRH_REQUIRE(p->field_order == NRH_FIELD_ORDER_SCALAR || p->field_order == NRH_FIELD_ORDER_SSE, NRH_ERR_ARG,
                "field_order must be NRH_FIELD_ORDER_SCALAR or NRH_FIELD_ORDER_SSE");
}

}  // namespace nrh

using namespace nrh;

extern "C" {

const char *nrh_last_error(void) { return g_last_error.c_str(); }

int nrh_version(void) { return 10000; }

int nrh_device_count(int *count) {
    return guard([&] {
        NRH_REQUIRE(count, NRH_ERR_ARG, "null count");
        int c = 0;
        hipError_t e = hipGetDeviceCount(&c);
        *count = e == hipSuccess ? c : 0;
    });
}

int nrh_init(int device) {
    return guard([&] {
        int c = 0;
        NRH_HIP(hipGetDeviceCount(&c));
        NRH_REQUIRE(device >= 0 && device < c, NRH_ERR_ARG, "invalid device index " + std::to_string(device));
        NRH_HIP(hipSetDevice(device));
        lts_static_scratch();
    });
}

int nrh_mat44_inverse(const float in[16], float out[16]) {
    return guard([&] {
        NRH_REQUIRE(in && out, NRH_ERR_ARG, "null matrix");
        mat44_inverse(in, out);
    });
}

int nrh_create_control_point_grid(const nrh_geom *ref, const float spacing_mm[3], nrh_geom *grid, float *identity) {
    return guard([&] {
        check_geom(ref, "reference");
        NRH_REQUIRE(spacing_mm && grid, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(spacing_mm[0] > 0 && spacing_mm[1] > 0 && spacing_mm[2] > 0, NRH_ERR_ARG, "spacing must be positive (mm)");
        create_grid(*ref, spacing_mm, *grid, identity);
    });
}

int nrh_f3d_create(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                   const int *mask, const nrh_geom *cpp, const nrh_f3d_params *params, nrh_f3d **out) {
    return guard([&] {
        check_geom(ref, "reference");
        check_geom(flo, "floating");
        check_geom(cpp, "control point grid");
        check_params(params);
        NRH_REQUIRE(ref_data && flo_data && out, NRH_ERR_ARG, "null image data or output");
        NRH_REQUIRE(ref->nz > 1, NRH_ERR_UNSUPPORTED, "2-D images are not supported");
        std::unique_ptr<nrh_f3d> h(new nrh_f3d);
        h->ref = *ref; h->flo = *flo; h->cpp = *cpp; h->params = *params;
        h->N = voxels(*ref); h->Nf = voxels(*flo); h->G = voxels(*cpp);
        NRH_HIP(hipGetDevice(&h->device));
        NRH_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        NRH_HIP(hipHostMalloc((void **)&h->pinned, 16 * sizeof(double), hipHostMallocDefault));
        NRH_HIP(hipHostMalloc((void **)&h->published, 16 * sizeof(double), hipHostMallocCoherent | hipHostMallocMapped));
        NRH_HIP(hipHostGetDevicePointer((void **)&h->publishedDev, h->published, 0));
        std::memset(h->published, 0, 16 * sizeof(double));  // published[15]: no sequence number yet
        // Images: float32 always kept (per-operation path); float16 copies added for the fused
        // path when requested.
        h->refImg.alloc(h->N);
        h->floImg.alloc(h->Nf);
        NRH_HIP(hipMemcpyAsync(h->refImg.p, ref_data, h->N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        NRH_HIP(hipMemcpyAsync(h->floImg.p, flo_data, h->Nf * sizeof(float), hipMemcpyHostToDevice, h->stream));
        if (mask) {
            bool any = false;
            for (size_t i = 0; i < h->N && !any; ++i) any = mask[i] < 0;
            if (any) {
                h->mask.alloc(h->N);
                NRH_HIP(hipMemcpyAsync(h->mask.p, mask, h->N * sizeof(int), hipMemcpyHostToDevice, h->stream));
            }
        }
        h->cppPos.alloc(3 * h->G);
        h->best.alloc(3 * h->G);
        h->nodeGrad.alloc(3 * h->G);
        h->cgG.alloc(3 * h->G);
        h->cgH.alloc(3 * h->G);
        NRH_HIP(hipMemsetAsync(h->cppPos.p, 0, h->cppPos.bytes(), h->stream));
        NRH_HIP(hipMemsetAsync(h->nodeGrad.p, 0, h->nodeGrad.bytes(), h->stream));
        alloc_histogram(h.get());
        h->ent.alloc(8);
        h->scalars.alloc(16);
        h->maxBits.alloc(8);
        NRH_HIP(hipMemsetAsync(h->maxBits.p, 0, h->maxBits.bytes(), h->stream));  // [6]: BE reduction ticket
        h->partials.alloc(4096);
        NRH_HIP(hipStreamSynchronize(h->stream));
        *out = h.release();
    });
}

int nrh_f3d_destroy(nrh_f3d *h) {
    return guard([&] {
        if (!h) return;
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        if (h->pinned) (void)hipHostFree(h->pinned);
        if (h->published) (void)hipHostFree(h->published);
        fused_release(h);
        for (int i = 0; i < 4; ++i)
            if (h->ev[i]) (void)hipEventDestroy(h->ev[i]);
        if (h->evPair) (void)hipEventDestroy(h->evPair);
        if (h->side) (void)hipStreamSynchronize(h->side);
        if (h->evFork) (void)hipEventDestroy(h->evFork);
        if (h->evJoin) (void)hipEventDestroy(h->evJoin);
        if (h->side) (void)hipStreamDestroy(h->side);
        hipStream_t s = h->stream;
        delete h;
        if (s) (void)hipStreamDestroy(s);
    });
}

int nrh_f3d_set_params(nrh_f3d *h, const nrh_f3d_params *params) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        check_params(params);
        const bool bins = params->ref_bins != h->params.ref_bins || params->flo_bins != h->params.flo_bins;
        h->params = *params;
        if (bins) { alloc_histogram(h); h->histVersion = 0; }
    });
}

int nrh_f3d_set_control_points(nrh_f3d *h, const float *planar) {
    return guard([&] {
        NRH_REQUIRE(h && planar, NRH_ERR_ARG, "null argument");
        NRH_HIP(hipMemcpyAsync(h->cppPos.p, planar, h->cppPos.bytes(), hipMemcpyHostToDevice, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
        h->cppVersion++;
    });
}

int nrh_f3d_get_control_points(nrh_f3d *h, float *planar) {
    return guard([&] {
        NRH_REQUIRE(h && planar, NRH_ERR_ARG, "null argument");
        NRH_HIP(hipMemcpyAsync(planar, h->cppPos.p, h->cppPos.bytes(), hipMemcpyDeviceToHost, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

int nrh_f3d_rescale_intensities(nrh_f3d *h, float ref_lo, float ref_hi, float flo_lo, float flo_hi) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        op_rescale(h, h->refImg.p, h->N, ref_lo, ref_hi);
        op_rescale(h, h->floImg.p, h->Nf, flo_lo, flo_hi);
        NRH_HIP(hipStreamSynchronize(h->stream));
        h->histVersion = 0;
        h->halfDirty = true;
    });
}

int nrh_f3d_get_images(nrh_f3d *h, float *ref, float *flo) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        if (ref) NRH_HIP(hipMemcpyAsync(ref, h->refImg.p, h->N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        if (flo) NRH_HIP(hipMemcpyAsync(flo, h->floImg.p, h->Nf * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

#define NRH_HANDLE_OP(name, body)                     \
    int name {                                        \
        return guard([&] {                            \
            NRH_REQUIRE(h, NRH_ERR_ARG, "null handle"); \
            body;                                     \
            NRH_HIP(hipStreamSynchronize(h->stream)); \
        });                                           \
    }

NRH_HANDLE_OP(nrh_f3d_deformation_field(nrh_f3d *h), op_spline_field(h))
NRH_HANDLE_OP(nrh_f3d_resample(nrh_f3d *h), op_resample(h))
NRH_HANDLE_OP(nrh_f3d_image_gradient(nrh_f3d *h), op_image_gradient(h))
NRH_HANDLE_OP(nrh_f3d_zero_voxel_gradient(nrh_f3d *h),
              ensure_voxel_grad(h); NRH_HIP(hipMemsetAsync(h->voxelGrad.p, 0, h->voxelGrad.bytes(), h->stream)))
NRH_HANDLE_OP(nrh_f3d_nmi_voxel_gradient(nrh_f3d *h, double weight), op_nmi_voxel_gradient(h, weight))
NRH_HANDLE_OP(nrh_f3d_convolve_voxel_gradient(nrh_f3d *h), op_convolve_voxel_gradient(h))
NRH_HANDLE_OP(nrh_f3d_voxel_to_node(nrh_f3d *h, float weight), op_voxel_to_node(h, weight))
NRH_HANDLE_OP(nrh_f3d_zero_node_gradient(nrh_f3d *h),
              NRH_HIP(hipMemsetAsync(h->nodeGrad.p, 0, h->nodeGrad.bytes(), h->stream)))
NRH_HANDLE_OP(nrh_f3d_bending_energy_gradient(nrh_f3d *h, float weight), op_bending_energy_gradient(h, weight))
NRH_HANDLE_OP(nrh_f3d_normalise_gradient(nrh_f3d *h, double max_length), op_normalise(h, max_length))
NRH_HANDLE_OP(nrh_f3d_store_best(nrh_f3d *h),
              NRH_HIP(hipMemcpyAsync(h->best.p, h->cppPos.p, h->cppPos.bytes(), hipMemcpyDeviceToDevice, h->stream)))
NRH_HANDLE_OP(nrh_f3d_restore_best(nrh_f3d *h),
              NRH_HIP(hipMemcpyAsync(h->cppPos.p, h->best.p, h->cppPos.bytes(), hipMemcpyDeviceToDevice, h->stream));
              h->cppVersion++)
NRH_HANDLE_OP(nrh_f3d_conjugate_gradient(nrh_f3d *h, int first), op_cg(h, first))
NRH_HANDLE_OP(nrh_f3d_update_control_points(nrh_f3d *h, float scale), op_update_dof(h, scale))

int nrh_f3d_nmi_value(nrh_f3d *h, double *nmi, double entropies[4]) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        NRH_REQUIRE(!h->warped.empty(), NRH_ERR_STATE, "NMI value before resampling");
        op_joint_histogram(h, h->warped.p);
        op_nmi_finalize(h);
        NRH_HIP(hipMemcpyAsync(h->pinned, h->ent.p, 5 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
        if (nmi) *nmi = h->pinned[4];
        if (entropies) for (int i = 0; i < 4; ++i) entropies[i] = h->pinned[i];
    });
}

int nrh_f3d_linear_energy(nrh_f3d *h, double *value) {
    return guard([&] {
        NRH_REQUIRE(h && value, NRH_ERR_ARG, "null argument");
        op_linear_energy(h, h->scalars.p + 8);
        *value = read_scalar(h, h->scalars.p + 8);
    });
}

NRH_HANDLE_OP(nrh_f3d_linear_energy_gradient(nrh_f3d *h, float weight), op_linear_energy_gradient(h, weight))

int nrh_f3d_bending_energy(nrh_f3d *h, double *value) {
    return guard([&] {
        NRH_REQUIRE(h && value, NRH_ERR_ARG, "null argument");
        op_bending_energy(h, h->scalars.p + 0);
        *value = read_scalar(h, h->scalars.p + 0);
    });
}

int nrh_f3d_max_length(nrh_f3d *h, double *value) {
    return guard([&] {
        NRH_REQUIRE(h && value, NRH_ERR_ARG, "null argument");
        op_max_length(h);
        *value = read_scalar(h, h->scalars.p + 1);
    });
}

static void download(nrh_f3d *h, const void *src, void *dst, size_t bytes, const char *what) {
    NRH_REQUIRE(dst, NRH_ERR_ARG, "null destination");
    NRH_REQUIRE(src, NRH_ERR_STATE, std::string(what) + " has not been computed");
    NRH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, h->stream));
    NRH_HIP(hipStreamSynchronize(h->stream));
}

int nrh_f3d_download_field(nrh_f3d *h, float *p) {
    return guard([&] { NRH_REQUIRE(h, NRH_ERR_ARG, "null handle"); download(h, h->field.p, p, 3 * h->N * sizeof(float), "deformation field"); });
}
int nrh_f3d_download_warped(nrh_f3d *h, float *p) {
    return guard([&] { NRH_REQUIRE(h, NRH_ERR_ARG, "null handle"); download(h, h->warped.p, p, h->N * sizeof(float), "warped image"); });
}
int nrh_f3d_download_warped_gradient(nrh_f3d *h, float *p) {
    return guard([&] { NRH_REQUIRE(h, NRH_ERR_ARG, "null handle"); download(h, h->warpedGrad.p, p, 3 * h->N * sizeof(float), "warped gradient"); });
}
int nrh_f3d_download_voxel_gradient(nrh_f3d *h, float *p) {
    return guard([&] { NRH_REQUIRE(h, NRH_ERR_ARG, "null handle"); download(h, h->voxelGrad.p, p, 3 * h->N * sizeof(float), "voxel-based gradient"); });
}
int nrh_f3d_download_node_gradient(nrh_f3d *h, float *p) {
    return guard([&] { NRH_REQUIRE(h, NRH_ERR_ARG, "null handle"); download(h, h->nodeGrad.p, p, 3 * h->G * sizeof(float), "node gradient"); });
}
int nrh_f3d_upload_node_gradient(nrh_f3d *h, const float *p) {
    return guard([&] {
        NRH_REQUIRE(h && p, NRH_ERR_ARG, "null argument");
        NRH_HIP(hipMemcpyAsync(h->nodeGrad.p, p, 3 * h->G * sizeof(float), hipMemcpyHostToDevice, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}
int nrh_f3d_download_histogram(nrh_f3d *h, double *pro, double *logh) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        if (pro) download(h, h->histPro.p, pro, h->totalBins() * sizeof(double), "histogram");
        if (logh) download(h, h->histLog.p, logh, h->totalBins() * sizeof(double), "log histogram");
    });
}

}  // extern "C"
