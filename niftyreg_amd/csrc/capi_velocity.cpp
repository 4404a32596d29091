// C ABI of the velocity-grid (reg_f3d2) operations and the symmetric grid helper.

#include "f3d_handle.hpp"
#include "host.hpp"
#include "m44.hpp"
#include <cstring>

namespace nrh {

// operator*(mat44, mat44) of reg-lib/cpu/Maths.hpp:237-246: double sums, float results.
void mat44_product(const float *A, const float *B, float *R) {
    float T[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            volatile double acc = (double)A[i * 4] * (double)B[j];
            acc = acc + (double)A[i * 4 + 1] * (double)B[4 + j];
            acc = acc + (double)A[i * 4 + 2] * (double)B[8 + j];
            acc = acc + (double)A[i * 4 + 3] * (double)B[12 + j];
            T[i * 4 + j] = (float)acc;
        }
    std::memcpy(R, T, sizeof T);
}

// The half affines of reg_createSymmetricControlPointGrids (reg-lib/cpu/_reg_localTrans.cpp:
// 150-166): NiftyReg::Mat44Logm (float -> double logarithm -> float), the float matrix times 0.5,
// NiftyReg::Mat44Expm (float -> double exponential -> float); the backward half from
// nifti_mat44_inverse of the affine.  Returns 1 when the reference stores them in the grids'
// extensions (both halves differ from the identity, :320-353), 0 when it stores none.
int half_affines(const float A[16], float hf[16], float hb[16]) {
    static const GaussLegendre8 gl = gauss_legendre8();
    auto half = [&](const float *M, float *out) {
        double m[16], l[16], h[16], e[16];
        for (int i = 0; i < 16; ++i) m[i] = M[i];
        NRH_REQUIRE(d4_logm(m, gl, l), NRH_ERR_NUMERIC,
                    "the matrix logarithm of the affine failed (a matrix with a non-positive determinant)");
        for (int i = 0; i < 16; ++i) h[i] = (double)(float)((double)(float)l[i] * 0.5);
        d4_expm(h, e);
        for (int i = 0; i < 16; ++i) out[i] = (float)e[i];
    };
    float inv[16];
    nifti_inverse(A, inv);
    half(A, hf);
    half(inv, hb);
    auto identity = [](const float *M) {
        for (int i = 0; i < 16; ++i)
            if (M[i] != ((i % 5) == 0 ? 1.f : 0.f)) return false;
        return true;
    };
    return !identity(hf) && !identity(hb);
}

// reg_createSymmetricControlPointGrids<float> (reg-lib/cpu/_reg_localTrans.cpp:132-358): both
// images' spaces carried half-way by the half affines (identity when NULL), their corners in mm
// (Mat44Mul<float>), the grid spans the bounding box with one node of margin, axis-aligned, sform
// only.
void symmetric_grids(const nrh_geom& ref, const nrh_geom& flo, const float spacing[3], nrh_geom& grid, float *fw,
                     float *bw, const float *halfFw, const float *halfBw) {
    static const float eye[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    float refSpace[16], floSpace[16];
    mat44_product(halfFw ? halfFw : eye, ref.xyz, refSpace);
    mat44_product(halfBw ? halfBw : eye, flo.xyz, floSpace);
    auto mul = [](const float *m, const float in[3], float out[3]) {
        for (int r = 0; r < 3; ++r)
            out[r] = (float)((double)m[r * 4] * (double)in[0] + (double)m[r * 4 + 1] * (double)in[1] +
                             (double)m[r * 4 + 2] * (double)in[2] + (double)m[r * 4 + 3]);
    };
    float mn[3] = { 0, 0, 0 }, mx[3] = { 0, 0, 0 };
    for (int c = 0; c < 8; ++c) {
        const float rc[3] = { (c & 1) ? (float)ref.nx : 0.f, (c & 2) ? (float)ref.ny : 0.f, (c & 4) ? (float)ref.nz : 0.f };
        const float fc[3] = { (c & 1) ? (float)flo.nx : 0.f, (c & 2) ? (float)flo.ny : 0.f, (c & 4) ? (float)flo.nz : 0.f };
        float ro[3], fo[3];
        mul(refSpace, rc, ro);
        mul(floSpace, fc, fo);
        for (int r = 0; r < 3; ++r) {
            if (c == 0) {
                mn[r] = fo[r] < ro[r] ? fo[r] : ro[r];
                mx[r] = ro[r] < fo[r] ? fo[r] : ro[r];
            } else {
                if (ro[r] < mn[r]) mn[r] = ro[r];
                if (fo[r] < mn[r]) mn[r] = fo[r];
                if (mx[r] < ro[r]) mx[r] = ro[r];
                if (mx[r] < fo[r]) mx[r] = fo[r];
            }
        }
    }
    grid = nrh_geom{};
    int *dims[3] = { &grid.nx, &grid.ny, &grid.nz };
    for (int r = 0; r < 3; ++r) {
        volatile float v = (mx[r] - mn[r]) / spacing[r];
        v = v + 3.f;
        const int t = (int)v;
        *dims[r] = t + (v > (float)t);
    }
    grid.dx = spacing[0];
    grid.dy = spacing[1];
    grid.dz = spacing[2];
    grid.sform = 1;
    for (int r = 0; r < 3; ++r) {
        grid.xyz[r * 4 + r] = spacing[r];
        grid.xyz[r * 4 + 3] = mn[r] - spacing[r];
    }
    grid.xyz[15] = 1.f;
    mat44_inverse(grid.xyz, grid.ijk);
    if (!fw && !bw) return;
    const size_t G = voxels(grid);
    size_t idx = 0;
    for (int z = 0; z < grid.nz; ++z)
        for (int y = 0; y < grid.ny; ++y)
            for (int x = 0; x < grid.nx; ++x, ++idx)
                for (int r = 0; r < 3; ++r) {
                    // reg_getDeformationFromDisplacement_3D (reg-lib/cpu/_reg_tools.cpp:2107-2150)
                    volatile float v = grid.xyz[r * 4] * (float)x;
                    v = v + grid.xyz[r * 4 + 1] * (float)y;
                    v = v + grid.xyz[r * 4 + 2] * (float)z;
                    v = v + grid.xyz[r * 4 + 3];
                    if (fw) fw[idx + r * G] = v;
                    if (bw) bw[idx + r * G] = v;
                }
}

}  // namespace nrh

using namespace nrh;

extern "C" {

int nrh_create_symmetric_control_point_grids(const nrh_geom *ref, const nrh_geom *flo, const float spacing_mm[3],
                                             const float *affine, nrh_geom *grid, float *fw, float *bw) {
    return guard([&] {
        check_geom(ref, "reference");
        check_geom(flo, "floating");
        NRH_REQUIRE(spacing_mm && grid, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(spacing_mm[0] > 0 && spacing_mm[1] > 0 && spacing_mm[2] > 0, NRH_ERR_ARG,
                    "grid spacing must be positive (mm)");
        NRH_REQUIRE(ref->nz > 1 && flo->nz > 1, NRH_ERR_UNSUPPORTED, "symmetric grids are 3-D only");
        float hf[16], hb[16];
        if (affine) half_affines(affine, hf, hb);
        symmetric_grids(*ref, *flo, spacing_mm, *grid, fw, bw, affine ? hf : nullptr, affine ? hb : nullptr);
    });
}

int nrh_symmetric_half_affines(const float affine[16], float half_fw[16], float half_bw[16], int *extensions) {
    return guard([&] {
        NRH_REQUIRE(affine && half_fw && half_bw, NRH_ERR_ARG, "null argument");
        const int ext = half_affines(affine, half_fw, half_bw);
        if (extensions) *extensions = ext;
    });
}

int nrh_f3d_set_velocity_affines(nrh_f3d *h, const float *ext0, const float *ext1, const float *transformation) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        h->velExt0 = ext0 != nullptr;
        h->velExt1 = ext1 != nullptr;
        h->velTrans = transformation != nullptr;
        if (ext0) std::memcpy(h->velExt0M, ext0, sizeof h->velExt0M);
        if (ext1) std::memcpy(h->velExt1M, ext1, sizeof h->velExt1M);
        if (transformation) std::memcpy(h->velTransM, transformation, sizeof h->velTransM);
        h->velFieldVersion = 0;
        h->velChainVersion = 0;
        h->pairStateVersion = 0;
    });
}

int nrh_f3d_set_velocity_grid(nrh_f3d *h, int steps) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        NRH_REQUIRE(h->ref.nz > 1, NRH_ERR_UNSUPPORTED, "velocity grids are 3-D only");
        h->velocity = true;
        h->velSteps = steps;
    });
}

int nrh_f3d_get_velocity_steps(nrh_f3d *h, int *steps) {
    return guard([&] {
        NRH_REQUIRE(h && steps, NRH_ERR_ARG, "null argument");
        *steps = h->velSteps;
    });
}

int nrh_f3d_set_velocity_steps(nrh_f3d *h, int steps) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        NRH_REQUIRE(h->velocity, NRH_ERR_STATE, "the control points are not a velocity grid");
        h->velSteps = steps;
    });
}

int nrh_f3d_def_field_from_velocity(nrh_f3d *h, int update_step_number) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        op_vel_def_field(h, update_step_number != 0);
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

int nrh_f3d_exponentiate_gradient(nrh_f3d *h, nrh_f3d *bw) {
    return guard([&] {
        NRH_REQUIRE(h && bw, NRH_ERR_ARG, "null handle");
        op_exponentiate_gradient(h, bw);
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

int nrh_f3d_update_velocity_field(nrh_f3d *h, float scale, int optimise_x, int optimise_y, int optimise_z) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        op_update_velocity(h, scale, optimise_x != 0, optimise_y != 0, optimise_z != 0);
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

int nrh_f3d_symmetrise_velocity_fields(nrh_f3d *fw, nrh_f3d *bw) {
    return guard([&] {
        NRH_REQUIRE(fw && bw, NRH_ERR_ARG, "null handle");
        op_symmetrise_velocity(fw, bw);
    });
}

int nrh_f3d_bch_update(nrh_f3d *h, float, int) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        NRH_REQUIRE(false, NRH_ERR_UNSUPPORTED,
                    "BCH update: the reference's compute_BCH_update is disabled (\"needs updating\")");
    });
}

int nrh_f3d_compose_field(nrh_f3d *h, const nrh_geom *geom, const float *field) {
    return guard([&] {
        NRH_REQUIRE(h && field, NRH_ERR_ARG, "null argument");
        check_geom(geom, "field");
        DevBuf<float> d;
        d.alloc(3 * voxels(*geom));
        NRH_HIP(hipMemcpyAsync(d.p, field, d.bytes(), hipMemcpyHostToDevice, h->stream));
        op_compose_field(h, *geom, d.p);
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

int nrh_f3d_resample_gradient(nrh_f3d *h, int interpolation, float padding, float *out) {
    return guard([&] {
        NRH_REQUIRE(h && out, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(interpolation == 1, NRH_ERR_UNSUPPORTED, "only linear interpolation is supported");
        h->velTmp.alloc(3 * h->N);
        op_resample_gradient(h, padding, h->velTmp.p);
        NRH_HIP(hipMemcpyAsync(out, h->velTmp.p, 3 * h->N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

int nrh_f3d_upload_field(nrh_f3d *h, const float *planar) {
    return guard([&] {
        NRH_REQUIRE(h && planar, NRH_ERR_ARG, "null argument");
        ensure_field(h);
        h->velFieldVersion = 0;
        h->pairStateVersion = 0;
        NRH_HIP(hipMemcpyAsync(h->field.p, planar, 3 * h->N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

int nrh_f3d_upload_voxel_gradient(nrh_f3d *h, const float *planar) {
    return guard([&] {
        NRH_REQUIRE(h && planar, NRH_ERR_ARG, "null argument");
        ensure_voxel_grad(h);
        NRH_HIP(hipMemcpyAsync(h->voxelGrad.p, planar, 3 * h->N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

}  // extern "C"
