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

// One voxel of reg_cubic_spline_getDeformationField3D (non-composition, reg-lib/cpu/
// _reg_localTrans.cpp:1110-1451): the 64 control points of the cell starting at node (px, py, pz)
// (z outer, y, x inner as get_GridValues reads them, reg-lib/cpu/_reg_splineBasis.cpp:603-665),
// weighted by x_a * (y_b * z_c) in float and summed in the order of the selected build
// (NRH_FIELD_ORDER_*): sequentially (USE_SSE off), or in four x lanes whose partial sums are then
// added pairwise (the 5-voxel table path, SSE3 horizontal adds) or left to right (general path).
__device__ __forceinline__ float3 ref_spline_point(const float *__restrict__ cpp, size_t G, int gnx, int gny, int px,
                                                   int py, int pz, const float xb[4], const float yb[4],
                                                   const float zb[4], int order, bool lut) {
#pragma clang fp contract(off)
    // the 48 control-point values of each z slab are loaded before any is summed (one round of
    // memory latency per slab); the sums run in the reference's order
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
    float l0[4] = { 0.f, 0.f, 0.f, 0.f }, l1[4] = { 0.f, 0.f, 0.f, 0.f }, l2[4] = { 0.f, 0.f, 0.f, 0.f };
    for (int c = 0; c < 4; ++c) {
        float v[4][4][3];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const size_t row = ((size_t)(pz + c) * gny + (py + b)) * gnx + px;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                v[b][a][0] = cpp[row + a];
                v[b][a][1] = cpp[row + a + G];
                v[b][a][2] = cpp[row + a + 2 * G];
            }
        }
        if (order == NRH_FIELD_ORDER_SCALAR) {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const float yzw = yb[b] * zb[c];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const float w = xb[a] * yzw;
                    r0 += v[b][a][0] * w;
                    r1 += v[b][a][1] * w;
                    r2 += v[b][a][2] * w;
                }
            }
        } else {
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                const float yzw = yb[b] * zb[c];
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const float w = xb[a] * yzw;
                    l0[a] = w * v[b][a][0] + l0[a];
                    l1[a] = w * v[b][a][1] + l1[a];
                    l2[a] = w * v[b][a][2] + l2[a];
                }
            }
        }
    }
    float3 r;
    if (order == NRH_FIELD_ORDER_SCALAR) {
        r.x = r0; r.y = r1; r.z = r2;
    } else if (lut) {
        r.x = (l0[0] + l0[1]) + (l0[2] + l0[3]);
        r.y = (l1[0] + l1[1]) + (l1[2] + l1[3]);
        r.z = (l2[0] + l2[1]) + (l2[2] + l2[3]);
    } else {
        r.x = ((l0[0] + l0[1]) + l0[2]) + l0[3];
        r.y = ((l1[0] + l1[1]) + l1[2]) + l1[3];
        r.z = ((l2[0] + l2[1]) + l2[2]) + l2[3];
    }
    return r;
}

// TrilinearImageGradient at floating voxel position p (reg-lib/cpu/_reg_resampling.cpp:1797-1915):
// derivative weights (-1, 1) in floating voxel units, padding for taps outside the image, NaN
// components replaced by 0.
__device__ __forceinline__ float3 image_gradient_at(const float *__restrict__ flo, int fnx, int fny, int fnz, float3 p,
                                                    float pad) {
#pragma clang fp contract(off)
    float g0 = 0.f, g1 = 0.f, g2 = 0.f;
    const int px = floor_int(p.x), py = floor_int(p.y), pz = floor_int(p.z);
    float r = p.x - (float)px;
    const float xb[2] = { (float)(1.0 - (double)r), r };
    r = p.y - (float)py;
    const float yb[2] = { (float)(1.0 - (double)r), r };
    r = p.z - (float)pz;
    const float zb[2] = { (float)(1.0 - (double)r), r };
    const float deriv[2] = { -1.f, 1.f };
    for (int c = 0; c < 2; ++c) {
        const int Z = pz + c;
        const bool zIn = Z > -1 && Z < fnz;
        float xx = 0.f, yy = 0.f, zz = 0.f;
        for (int b = 0; b < 2; ++b) {
            const int Y = py + b;
            const bool yzIn = zIn && Y > -1 && Y < fny;
            float xT = 0.f, yT = 0.f;
            for (int a = 0; a < 2; ++a) {
                const int X = px + a;
                const float coeff = (yzIn && X > -1 && X < fnx) ? flo[((size_t)Z * fny + Y) * fnx + X] : pad;
                xT += coeff * deriv[a];
                yT += coeff * xb[a];
            }
            xx += xT * yb[b];
            yy += yT * deriv[b];
            zz += yT * yb[b];
        }
        g0 += xx * zb[c];
        g1 += yy * zb[c];
        g2 += zz * deriv[c];
    }
    if (g0 != g0) g0 = 0.f;
    if (g1 != g1) g1 = 0.f;
    if (g2 != g2) g2 = 0.f;
    return make_float3(g0, g1, g2);
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

// The voxel term of reg_getVoxelBasedNmiGradient3d (reg-lib/cpu/_reg_nmi.cpp:501-593): the FP64
// sums over the reference's bin loops r in [(int)(rv - 1), (int)(rv + 3)), w likewise, of
// B(rv - r) B'(wv - w) g log p, returned as weight (rd + wd - NMI jd) / (H_rf total) per axis.
// logh: the reference's jointHistogramLog ([rb fb] joint, then rb + fb marginals); ent: H_r, H_f,
// H_rf, total.  The caller has checked that rv and wv are numbers.
__device__ __forceinline__ void nmi_voxel_term(float rv, float wv, const float g[3], const double *__restrict__ logh,
                                               const double *__restrict__ ent, int rb, int fb, double weight,
                                               float out[3]) {
#pragma clang fp contract(off)
    const double nmi = (ent[0] + ent[1]) / ent[2];
    const size_t refOff = (size_t)rb * fb, floOff = refOff + rb;
    double jd[3] = { 0, 0, 0 }, rd[3] = { 0, 0, 0 }, wd[3] = { 0, 0, 0 };
    // the Parzen values of each axis are evaluated once instead of once per pair
    const int r0 = (int)(rv - 1.f), r1 = (int)(rv + 3.f), w0 = (int)(wv - 1.f), w1 = (int)(wv + 3.f);
    double bR[4], dW[4], rl[4], wl[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int r = r0 + k, w = w0 + k;
        bR[k] = nmi_basis((double)(rv - (float)r));
        dW[k] = nmi_basis_deriv((double)(wv - (float)w));
        rl[k] = (-1 < r && r < rb) ? logh[r + refOff] : 0.0;
        wl[k] = (-1 < w && w < fb) ? logh[w + floOff] : 0.0;
    }
    // the joint log values are loaded together before the sums
    double jlog[4][4];
    // every lane's 4 x 4 bins in range (the loop bounds r1, w1 give exactly four) and no NaN
    // gradient component: the same sums without the per-pair checks
    const bool full = r1 == r0 + 4 && r0 >= 0 && r0 + 3 < rb && w1 == w0 + 4 && w0 >= 0 && w0 + 3 < fb &&
                      g[0] == g[0] && g[1] == g[1] && g[2] == g[2];
    if (__all(full)) {
#pragma unroll
        for (int kr = 0; kr < 4; ++kr)
#pragma unroll
            for (int kw = 0; kw < 4; ++kw) jlog[kr][kw] = logh[(r0 + kr) + (w0 + kw) * rb];
#pragma unroll
        for (int kr = 0; kr < 4; ++kr)
#pragma unroll
            for (int kw = 0; kw < 4; ++kw) {
                const double common = bR[kr] * dW[kw];
#pragma unroll
                for (int d = 0; d < 3; ++d) {
                    rd[d] += common * (double)g[d] * rl[kr];
                    wd[d] += common * (double)g[d] * wl[kw];
                    jd[d] += common * (double)g[d] * jlog[kr][kw];
                }
            }
    } else {
#pragma unroll
    for (int kr = 0; kr < 4; ++kr)
#pragma unroll
        for (int kw = 0; kw < 4; ++kw) {
            const int r = r0 + kr, w = w0 + kw;
            jlog[kr][kw] = (-1 < r && r < rb && -1 < w && w < fb) ? logh[r + w * rb] : 0.0;
        }
#pragma unroll
    for (int kr = 0; kr < 4; ++kr) {
        const int r = r0 + kr;
        if (!(r < r1) || !(-1 < r && r < rb)) continue;
#pragma unroll
        for (int kw = 0; kw < 4; ++kw) {
            const int w = w0 + kw;
            if (!(w < w1) || !(-1 < w && w < fb)) continue;
            const double common = bR[kr] * dW[kw];
            const double jl = jlog[kr][kw];
            for (int d = 0; d < 3; ++d) {
                if (g[d] == g[d]) {
                    rd[d] += common * (double)g[d] * rl[kr];
                    wd[d] += common * (double)g[d] * wl[kw];
                    jd[d] += common * (double)g[d] * jl;
                }
            }
        }
    }
    }
    const double den = ent[2] * ent[3];
    for (int d = 0; d < 3; ++d) out[d] = (float)(weight * (rd[d] + wd[d] - nmi * jd[d]) / den);
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

// Three block sums at once, each in the same order as block_sum_double (smem: 48 doubles); the
// results are valid on thread 0.
__device__ __forceinline__ void block_sum_double3(double& a, double& b, double& c, double *smem) {
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        c += __shfl_down(c, off, 64);
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) {
        smem[wid] = a;
        smem[16 + wid] = b;
        smem[32 + wid] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        double ra = 0, rb = 0, rc = 0;
        for (int i = 0; i < nw; ++i) {
            ra += smem[i];
            rb += smem[16 + i];
            rc += smem[32 + i];
        }
        a = ra;
        b = rb;
        c = rc;
    }
}

__device__ __forceinline__ float wave_sum(float v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Coordinates of linear index i on an nx x ny x (any) grid, x fastest: 32-bit divisions while the
// index fits (a 64-bit division by a run-time value is a long emulated sequence on the GPU).
__device__ __forceinline__ void index_xyz(size_t i, int nx, int ny, int& x, int& y, int& z) {
    if (i <= 0xffffffffull) {
        const uint32_t j = (uint32_t)i, row = j / (uint32_t)nx;
        x = (int)(j - row * (uint32_t)nx);
        y = (int)(row % (uint32_t)ny);
        z = (int)(row / (uint32_t)ny);
    } else {
        const size_t row = i / (size_t)nx;
        x = (int)(i - row * (size_t)nx);
        y = (int)(row % (size_t)ny);
        z = (int)(row / (size_t)ny);
    }
}

// Node position in float, reg_getDisplacementFromDeformation_3D (reg-lib/cpu/_reg_tools.cpp:2004-2019)
__device__ __forceinline__ float node_pos(const Affine& m, int r, int x, int y, int z) {
#pragma clang fp contract(off)
    return m.m[r * 4 + 0] * (float)x + m.m[r * 4 + 1] * (float)y + m.m[r * 4 + 2] * (float)z + m.m[r * 4 + 3];
}

}  // namespace nrh
