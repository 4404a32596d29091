// Device-side helpers shared by the F3D and Aladin kernels.
//
// The scalar helpers reproduce the reference's rounding: float where the reference computes
// in float, double where it promotes, and no contraction into fused multiply-adds inside the
// helpers that feed comparisons with the reference (floor, basis values, matrix products).
#pragma once

#include <hip/hip_runtime.h>
#include "common.hpp"

namespace nrh {

// Floor<int>(float) (reg-lib/cpu/Maths.hpp:63-66)
__device__ __forceinline__ int floor_int(float x) {
    const int i = (int)x;
    return i - (x < (float)i);
}

// Mat44Mul<float>: double accumulation, float result (reg-lib/cpu/Maths.hpp:272-285)
__device__ __forceinline__ float3 affine_apply_d(const Affine& a, float x, float y, float z) {
#pragma clang fp contract(off)
    float3 r;
    r.x = (float)((double)a.m[0] * (double)x + (double)a.m[1] * (double)y + (double)a.m[2] * (double)z + (double)a.m[3]);
    r.y = (float)((double)a.m[4] * (double)x + (double)a.m[5] * (double)y + (double)a.m[6] * (double)z + (double)a.m[7]);
    r.z = (float)((double)a.m[8] * (double)x + (double)a.m[9] * (double)y + (double)a.m[10] * (double)z + (double)a.m[11]);
    return r;
}

// Same product in float with fused multiply-adds (used by the fused kernels, where the
// result is compared against the reference within tolerance).
__device__ __forceinline__ float3 affine_apply_f(const Affine& a, float x, float y, float z) {
    float3 r;
    r.x = fmaf(a.m[0], x, fmaf(a.m[1], y, fmaf(a.m[2], z, a.m[3])));
    r.y = fmaf(a.m[4], x, fmaf(a.m[5], y, fmaf(a.m[6], z, a.m[7])));
    r.z = fmaf(a.m[8], x, fmaf(a.m[9], y, fmaf(a.m[10], z, a.m[11])));
    return r;
}

// get_BSplineBasisValues<float> (reg-lib/cpu/_reg_splineBasis.cpp:17-26)
__device__ __forceinline__ void bspline_basis(float t, float v[4]) {
#pragma clang fp contract(off)
    const float ff = t * t;
    const float fff = ff * t;
    const float mf = (float)(1.0 - (double)t);
    v[0] = (float)((double)(mf * mf * mf) / 6.0);
    v[1] = (float)((3.0 * (double)fff - 6.0 * (double)ff + 4.0) / 6.0);
    v[2] = (float)((-3.0 * (double)fff + 3.0 * (double)ff + 3.0 * (double)t + 1.0) / 6.0);
    v[3] = (float)((double)fff / 6.0);
}

// Control-point cell and fractional position of voxel coordinate c along one axis:
// the 5-voxel look-up-table rule (reg-lib/cpu/_reg_localTrans.cpp:1127-1159) or the general
// rule (reg-lib/cpu/_reg_localTrans.cpp:1331-1364).
__device__ __forceinline__ void cell_of(int c, float gvs, bool lut, int& pre, float& t) {
#pragma clang fp contract(off)
    if (lut) {
        pre = c / 5;
        t = (float)(c - 5 * pre) / 5.f;
    } else {
        pre = (int)((float)c / gvs);
        t = (float)c / gvs - (float)pre;
        if (t < 0) t = 0;
    }
}

// Trilinear sample of a float volume at floating voxel position p, with the padding value for
// taps outside the volume (ResampleImage3D, reg-lib/cpu/_reg_resampling.cpp:316-464 with
// interpLinearKernel, :79-84).
__device__ __forceinline__ float sample_linear(const float *__restrict__ img, int nx, int ny, int nz, float3 p,
                                               float pad) {
#pragma clang fp contract(off)
    const int px = floor_int(p.x), py = floor_int(p.y), pz = floor_int(p.z);
    float rx = p.x - (float)px, ry = p.y - (float)py, rz = p.z - (float)pz;
    if (rx < 0) rx = 0;
    if (ry < 0) ry = 0;
    if (rz < 0) rz = 0;
    const float xb[2] = { 1.f - rx, rx }, yb[2] = { 1.f - ry, ry }, zb[2] = { 1.f - rz, rz };
    const bool inside = px > -1 && px + 1 < nx && py > -1 && py + 1 < ny && pz > -1 && pz + 1 < nz;
    float intensity = 0.f;
    for (int c = 0; c < 2; ++c) {
        const int Z = pz + c;
        float yT = 0.f;
        for (int b = 0; b < 2; ++b) {
            const int Y = py + b;
            float xT = 0.f;
            for (int a = 0; a < 2; ++a) {
                const int X = px + a;
                float v;
                if (inside || (X > -1 && X < nx && Y > -1 && Y < ny && Z > -1 && Z < nz))
                    v = img[((size_t)Z * ny + Y) * nx + X];
                else v = pad;
                xT += v * xb[a];
            }
            yT += xT * yb[b];
        }
        intensity += yT * zb[c];
    }
    return intensity;
}

// GetBasisSplineValue<double> (reg-lib/cpu/_reg_nmi.h:116-129); 2.f/3.f is a float constant there.
__device__ __forceinline__ double nmi_basis(double x) {
    x = x < 0 ? -x : x;
    double v = 0;
    if (x < 2.0) {
        if (x < 1.0) v = (double)(2.f / 3.f) + (0.5 * x - 1.0) * x * x;
        else {
            x -= 2.0;
            v = -x * x * x / 6.0;
        }
    }
    return v;
}

// GetBasisSplineDerivativeValue<double> (reg-lib/cpu/_reg_nmi.h:131-145)
__device__ __forceinline__ double nmi_basis_deriv(double ox) {
    double x = ox < 0 ? -ox : ox;
    double v = 0;
    if (x < 2.0) {
        if (x < 1.0) v = (1.5 * x - 2.0) * ox;
        else {
            x -= 2.0;
            v = -0.5 * x * x;
            if (ox < 0) v = -v;
        }
    }
    return v;
}

// Float versions of the Parzen window and its derivative for the fused gradient.
__device__ __forceinline__ float nmi_basis_f(float x) {
    x = fabsf(x);
    if (x >= 2.f) return 0.f;
    if (x < 1.f) return 0.66666669f + (0.5f * x - 1.f) * x * x;
    x -= 2.f;
    return -x * x * x * (1.f / 6.f);
}

__device__ __forceinline__ float nmi_basis_deriv_f(float ox) {
    float x = fabsf(ox);
    if (x >= 2.f) return 0.f;
    if (x < 1.f) return (1.5f * x - 2.f) * ox;
    x -= 2.f;
    const float v = -0.5f * x * x;
    return ox < 0 ? -v : v;
}

// Block-wide sum of doubles (blockDim.x a multiple of 64, at most 1024).
__device__ __forceinline__ double block_sum_double(double v, double *smem) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) smem[wid] = v;
    __syncthreads();
    double r = 0;
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int i = 0; i < nw; ++i) r += smem[i];
    }
    return r;  // valid on thread 0
}

__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

}  // namespace nrh
