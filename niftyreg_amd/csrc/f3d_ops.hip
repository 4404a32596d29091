// Per-operation F3D kernels: one launch per reference Compute / reg_nmi virtual.
//
// These kernels keep the reference's rounding (float/double split, summation order, no fused
// multiply-adds) so that each stage can be compared element by element with the CPU path.
// The hot loop of a registration uses the fused kernels of f3d_fused.hip instead.

#include "device.hpp"
#include "be_blocks.hpp"
#include "m33.hpp"
#include "m44.hpp"
#include "host.hpp"
#include "f3d_handle.hpp"
#include "f3d_batch.hpp"
#include <cmath>

namespace nrh {

namespace {

constexpr int kBlock = 256;

inline unsigned blocks_for(size_t n, int block = kBlock) {
    size_t b = (n + block - 1) / block;
    return (unsigned)(b > 0 ? b : 1);
}

// ------------------------------------------------------------------ deformation field
// reg_cubic_spline_getDeformationField3D, non-composition (reg-lib/cpu/_reg_localTrans.cpp:1080-1453)
__global__ void __launch_bounds__(kBlock)
k_spline_field(int nx, int ny, int nz, int gnx, int gny, size_t G, const float *__restrict__ cpp,
               const int *__restrict__ mask, float gx, float gy, float gz, int lut, int order,
               float *__restrict__ field, size_t N) {
#pragma clang fp contract(off)
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    const int x = (int)(idx % nx);
    const int y = (int)((idx / nx) % ny);
    const int z = (int)(idx / ((size_t)nx * ny));
    if (mask && mask[idx] < 0) {
        if (!lut) { field[idx] = 0.f; field[idx + N] = 0.f; field[idx + 2 * N] = 0.f; }
        return;
    }
    int px, py, pz;
    float tx, ty, tz, xb[4], yb[4], zb[4];
    cell_of(x, gx, lut, px, tx);
    cell_of(y, gy, lut, py, ty);
    cell_of(z, gz, lut, pz, tz);
    bspline_basis(tx, xb);
    bspline_basis(ty, yb);
    bspline_basis(tz, zb);
    const float3 r = ref_spline_point(cpp, G, gnx, gny, px, py, pz, xb, yb, zb, order, lut != 0);
    field[idx] = r.x;
    field[idx + N] = r.y;
    field[idx + 2 * N] = r.z;
}

// ------------------------------------------------------------------ resampling
// ResampleImage3D, linear kernel (reg-lib/cpu/_reg_resampling.cpp:316-464)
__global__ void __launch_bounds__(kBlock)
k_resample_linear(const float *__restrict__ field, size_t N, const int *__restrict__ mask,
                  const float *__restrict__ flo, int fnx, int fny, int fnz, Affine ijk, float pad,
                  float *__restrict__ warped) {
#pragma clang fp contract(off)
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    if (mask && mask[idx] < 0) return;
    const float3 p = affine_apply_d(ijk, field[idx], field[idx + N], field[idx + 2 * N]);
    warped[idx] = sample_linear(flo, fnx, fny, fnz, p, pad);
}

// TrilinearImageGradient (reg-lib/cpu/_reg_resampling.cpp:1797-1915)
__global__ void __launch_bounds__(kBlock)
k_image_gradient_linear(const float *__restrict__ field, size_t N, const int *__restrict__ mask,
                        const float *__restrict__ flo, int fnx, int fny, int fnz, Affine ijk, float pad,
                        float *__restrict__ grad) {
#pragma clang fp contract(off)
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    float3 g = make_float3(0.f, 0.f, 0.f);
    if (!mask || mask[idx] > -1)
        g = image_gradient_at(flo, fnx, fny, fnz, affine_apply_d(ijk, field[idx], field[idx + N], field[idx + 2 * N]),
                              pad);
    const float g0 = g.x, g1 = g.y, g2 = g.z;
    grad[idx] = g0;
    grad[idx + N] = g1;
    grad[idx + 2 * N] = g2;
}

// ------------------------------------------------------------------ joint histogram
// Approximate-Parzen count fill of reg_getNmiValue (reg-lib/cpu/_reg_nmi.cpp:211-248):
// integer counts, privatised per workgroup in LDS, merged with integer atomics (exact in
// any order).
__global__ void __launch_bounds__(kBlock)
k_joint_histogram(const float *__restrict__ ref, const float *__restrict__ war, const int *__restrict__ mask,
                  size_t N, int rb, int fb, uint32_t *__restrict__ counts) {
    extern __shared__ uint32_t lh[];
    const int nb = rb * fb;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) lh[i] = 0;
    __syncthreads();
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < N; idx += (size_t)gridDim.x * blockDim.x) {
        if (mask && mask[idx] < 0) continue;
        const float r = ref[idx], w = war[idx];
        if (r == r && w == w && 0.f <= r && r < (float)rb && 0.f <= w && w < (float)fb)
            atomicAdd(&lh[(int)r + (int)w * rb], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nb; i += blockDim.x)
        if (lh[i]) atomicAdd(&counts[i], lh[i]);
}

// Smoothing, normalisation, marginals and entropies of reg_getNmiValue (reg-lib/cpu/_reg_nmi.cpp:249-344)
// in FP64, one workgroup.  ent: H_r, H_f, H_rf, total, nmi.
__global__ void __launch_bounds__(1024)
k_nmi_finalize(uint32_t *__restrict__ counts, int rb, int fb, double *__restrict__ pro,
               double *__restrict__ logh, float *__restrict__ logf, double *__restrict__ ent) {
    __shared__ double red[16];
    const int nb = rb * fb, total = nb + rb + fb;
    const double k0 = 1.0 / 6.0, k1 = (double)(2.f / 3.f);  // GetBasisSplineValue(-1), (0)
    // the counts are consumed: zeroed as they are read, ready for the next histogram
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        pro[i] = i < nb ? (double)counts[i] : 0.0;
        if (i < nb) counts[i] = 0u;
    }
    __syncthreads();
    // along the reference axis into logh
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        const int r = i % rb, f = i / rb;
        double v = 0;
        if (r - 1 > -1) v += pro[i - 1] * k0;
        v += pro[i] * k1;
        if (r + 1 < rb) v += pro[i + 1] * k0;
        logh[r + rb * f] = v;
    }
    __syncthreads();
    // along the floating axis back into pro
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        const int f = i / rb;
        double v = 0;
        if (f - 1 > -1) v += logh[i - rb] * k0;
        v += logh[i] * k1;
        if (f + 1 < fb) v += logh[i + rb] * k0;
        pro[i] = v;
    }
    __syncthreads();
    double s = 0;
    for (int i = threadIdx.x; i < total; i += blockDim.x) s += pro[i];
    s = block_sum_double(s, red);
    __shared__ double totalS;
    if (threadIdx.x == 0) totalS = s;
    __syncthreads();
    const double act = totalS;
    for (int i = threadIdx.x; i < total; i += blockDim.x) pro[i] /= act;
    __syncthreads();
    for (int i = threadIdx.x; i < rb + fb; i += blockDim.x) {
        double m = 0;
        if (i < rb) {
            for (int f = 0; f < fb; ++f) m += pro[i + rb * f];
        } else {
            const int f = i - rb;
            for (int r = 0; r < rb; ++r) m += pro[rb * f + r];
        }
        pro[nb + i] = m;
    }
    __syncthreads();
    double hr = 0, hf = 0, hj = 0;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const double p = pro[i];
        double l = 0;
        if (p > 0) {
            l = log(p);
            if (i < nb) hj -= p * l;
            else if (i < nb + rb) hr -= p * l;
            else hf -= p * l;
        }
        logh[i] = l;
        logf[i] = (float)l;
    }
    hr = block_sum_double(hr, red);
    __syncthreads();
    hf = block_sum_double(hf, red);
    __syncthreads();
    hj = block_sum_double(hj, red);
    if (threadIdx.x == 0) {
        ent[0] = hr;
        ent[1] = hf;
        ent[2] = hj;
        ent[3] = act;
        ent[4] = (hr + hf) / hj;
    }
}

// K[i] = log p_r + log p_w - NMI log p_rw for joint bin i = r + w rb, from the logs L (joint
// [0, nb), reference marginal [nb, nb + rb), floating marginal after), rounded once to float.
__device__ __forceinline__ float nmi_table_entry(const double *L, int nb, int rb, double nmi, int i) {
#pragma clang fp contract(off)
    return (float)((L[nb + i % rb] + L[nb + rb + i / rb]) - nmi * L[i]);
}

// The same finalisation with the working array in LDS (histograms of up to kFinLds entries):
// identical arithmetic and thread partition, no global-memory round trips between phases.
constexpr int kFinLds = 7936;
constexpr int kFinPer = (kFinLds + 1023) / 1024;

// Phase 1: P (LDS) = the smoothed joint histogram normalised by its total, then its marginals;
// consume: the counts are zeroed as they are read, ready for the next histogram.
__device__ __forceinline__ void fin_probabilities(uint32_t *__restrict__ counts, bool consume, int rb, int fb,
                                                  double *P, double *red, double& totalS) {
    const int nb = rb * fb, total = nb + rb + fb;
    const double k0 = 1.0 / 6.0, k1 = (double)(2.f / 3.f);
    // the counts are consumed: zeroed as they are read, ready for the next histogram
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        P[i] = i < nb ? (double)counts[i] : 0.0;
        if (consume && i < nb) counts[i] = 0u;
    }
    __syncthreads();
    double v[kFinPer];
    // along the reference axis
#pragma unroll
    for (int q = 0; q < kFinPer; ++q) {
        const int i = threadIdx.x + q * 1024;
        v[q] = 0;
        if (i < nb) {
            const int r = i % rb;
            double t = 0;
            if (r - 1 > -1) t += P[i - 1] * k0;
            t += P[i] * k1;
            if (r + 1 < rb) t += P[i + 1] * k0;
            v[q] = t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kFinPer; ++q) {
        const int i = threadIdx.x + q * 1024;
        if (i < nb) P[i] = v[q];
    }
    __syncthreads();
    // along the floating axis
#pragma unroll
    for (int q = 0; q < kFinPer; ++q) {
        const int i = threadIdx.x + q * 1024;
        v[q] = 0;
        if (i < nb) {
            const int f = i / rb;
            double t = 0;
            if (f - 1 > -1) t += P[i - rb] * k0;
            t += P[i] * k1;
            if (f + 1 < fb) t += P[i + rb] * k0;
            v[q] = t;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kFinPer; ++q) {
        const int i = threadIdx.x + q * 1024;
        if (i < nb) P[i] = v[q];
    }
    __syncthreads();
    double s = 0;
    for (int i = threadIdx.x; i < total; i += blockDim.x) s += P[i];
    s = block_sum_double(s, red);
    if (threadIdx.x == 0) totalS = s;
    __syncthreads();
    const double act = totalS;
    for (int i = threadIdx.x; i < total; i += blockDim.x) P[i] /= act;
    __syncthreads();
    for (int i = threadIdx.x; i < rb + fb; i += blockDim.x) {
        double m = 0;
        if (i < rb) {
            for (int f = 0; f < fb; ++f) m += P[i + rb * f];
        } else {
            const int f = i - rb;
            for (int r = 0; r < rb; ++r) m += P[rb * f + r];
        }
        P[nb + i] = m;
    }
    __syncthreads();
}

// Phase 2: the logarithms (written with the probabilities), the entropies (each thread's strided
// terms, then block sums: H_r, H_f, H_rf), ent[] and the table K.  P holds the logs after.
__device__ __forceinline__ void fin_entropies(int rb, int fb, double *P, double *red /* 48 */, double& totalS,
                                              double *__restrict__ pro, double *__restrict__ logh,
                                              float *__restrict__ logf, double *__restrict__ ent,
                                              float *__restrict__ kTab) {
    const int nb = rb * fb, total = nb + rb + fb;
    const double act = totalS;
    double hr = 0, hf = 0, hj = 0;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        const double p = P[i];
        double l = 0;
        if (p > 0) {
            l = log(p);
            if (i < nb) hj -= p * l;
            else if (i < nb + rb) hr -= p * l;
            else hf -= p * l;
        }
        pro[i] = p;
        logh[i] = l;
        logf[i] = (float)l;
        P[i] = l;  // the logs stay in LDS for the table K
    }
    block_sum_double3(hr, hf, hj, red);
    if (threadIdx.x == 0) {
        ent[0] = hr;
        ent[1] = hf;
        ent[2] = hj;
        ent[3] = act;
        ent[4] = (hr + hf) / hj;
        totalS = (hr + hf) / hj;  // the NMI, for K
    }
    if (kTab) {
        // the fused gradient sweep's table K[r + w rb] = log p_r + log p_w - NMI log p_rw (as
        // k_nmi_table computes it, from the same FP64 logs)
        __syncthreads();
        const double nmi = totalS;
        for (int i = threadIdx.x; i < nb; i += blockDim.x) kTab[i] = nmi_table_entry(P, nb, rb, nmi, i);
    }
}

// The whole finalisation in one workgroup.
__device__ __forceinline__ void
nmi_finalize_lds_block(uint32_t *__restrict__ counts, int rb, int fb, double *__restrict__ pro,
                   double *__restrict__ logh, float *__restrict__ logf, double *__restrict__ ent,
                   float *__restrict__ kTab = nullptr) {
    __shared__ double P[kFinLds];
    __shared__ double red[48];
    __shared__ double totalS;
    fin_probabilities(counts, true, rb, fb, P, red, totalS);
    fin_entropies(rb, fb, P, red, totalS, pro, logh, logf, ent, kTab);
}

// reg_getVoxelBasedNmiGradient3d<float> (reg-lib/cpu/_reg_nmi.cpp:501-593); grad +=
__global__ void __launch_bounds__(kBlock)
k_nmi_voxel_gradient(const float *__restrict__ ref, const float *__restrict__ war, const float *__restrict__ wg,
                     const int *__restrict__ mask, size_t N, int rb, int fb, const double *__restrict__ logh,
                     const double *__restrict__ ent, double weight, float *__restrict__ grad) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    if (mask && mask[i] < 0) return;
    const float rv = ref[i], wv = war[i];
    if (!(rv == rv && wv == wv)) return;
    const float g[3] = { wg[i], wg[i + N], wg[i + 2 * N] };
    float d[3];
    nmi_voxel_term(rv, wv, g, logh, ent, rb, fb, weight, d);
    for (int k = 0; k < 3; ++k) grad[i + k * N] += d[k];
}

// The same voxel term with the log histogram staged in LDS once per workgroup (its 16 joint and 8
// marginal look-ups per voxel then hit LDS): a persistent grid over the voxels, identical
// arithmetic.  logLds: rb * fb + rb + fb doubles of dynamic LDS.
__global__ void __launch_bounds__(kBlock)
k_nmi_voxel_gradient_lds(const float *__restrict__ ref, const float *__restrict__ war, const float *__restrict__ wg,
                         const int *__restrict__ mask, size_t N, int rb, int fb, const double *__restrict__ logh,
                         const double *__restrict__ ent, double weight, float *__restrict__ grad) {
#pragma clang fp contract(off)
    extern __shared__ double logLds[];
    const int total = rb * fb + rb + fb;
    for (int i = threadIdx.x; i < total; i += blockDim.x) logLds[i] = logh[i];
    __syncthreads();
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
        if (mask && mask[i] < 0) continue;
        const float rv = ref[i], wv = war[i];
        if (!(rv == rv && wv == wv)) continue;
        const float g[3] = { wg[i], wg[i + N], wg[i + 2 * N] };
        float d[3];
        nmi_voxel_term(rv, wv, g, logLds, ent, rb, fb, weight, d);
        for (int k = 0; k < 3; ++k) grad[i + k * N] += d[k];
    }
}

// ------------------------------------------------------------------ convolution
// Density preparation of reg_tools_kernelConvolution (reg-lib/cpu/_reg_tools.cpp:860-866).
__global__ void __launch_bounds__(kBlock)
k_conv_prepare(float *__restrict__ data, const int *__restrict__ mask, size_t N, float *__restrict__ dens,
               uint8_t *__restrict__ isnan_) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float v = data[i];
    const float d = ((!mask || mask[i] >= 0) && v == v) ? 1.f : 0.f;
    dens[i] = d;
    isnan_[i] = d == 0.f;
    if (d == 0.f) data[i] = 0.f;
}

// One axis of the density-normalised line convolution (reg-lib/cpu/_reg_tools.cpp:977-1093,
// scalar path: kernel-weighted sums in double over the in-range taps).
__global__ void __launch_bounds__(kBlock)
k_conv_axis(const float *__restrict__ data, const float *__restrict__ dens, int nx, int ny, int nz, int axis,
            const float *__restrict__ kernel, int radius, float *__restrict__ outData, float *__restrict__ outDens,
            int order) {
#pragma clang fp contract(off)
    const size_t N = (size_t)nx * ny * nz;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int x, y, z;
    index_xyz(i, nx, ny, x, y, z);
    const int L = axis == 0 ? nx : axis == 1 ? ny : nz;
    const int l = axis == 0 ? x : axis == 1 ? y : z;
    const size_t stride = axis == 0 ? 1 : axis == 1 ? (size_t)nx : (size_t)nx * ny;
    const size_t base = i - (size_t)l * stride;
    int pre = l - radius, pst = l + radius + 1;
    int k0 = 0;
    if (pre < 0) { k0 = -pre; pre = 0; }
    if (pst > L) pst = L;
    double is = 0, ds = 0;
    int k = pre, kk = k0;
    if (order == NRH_FIELD_ORDER_SSE) {
        // the SSE build (reg-lib/cpu/_reg_tools.cpp:1018-1049): blocks of four taps summed in four
        // float lanes (lane j takes the block's tap 3 - j), the lanes reduced as SSE3's two
        // horizontal adds do, (l0 + l1) + (l2 + l3); the last < 4 taps added in double
        float li[4] = { 0.f, 0.f, 0.f, 0.f }, ld[4] = { 0.f, 0.f, 0.f, 0.f };
        for (; k < pst - 3; k += 4, kk += 4)
            for (int j = 0; j < 4; ++j) {
                const float kv = kernel[kk + 3 - j];
                const size_t o = base + (size_t)(k + 3 - j) * stride;
                li[j] = kv * data[o] + li[j];
                ld[j] = kv * dens[o] + ld[j];
            }
        is = (double)((li[0] + li[1]) + (li[2] + li[3]));
        ds = (double)((ld[0] + ld[1]) + (ld[2] + ld[3]));
    }
    for (; k < pst; ++k, ++kk) {
        const float kv = kernel[kk];
        is += (double)(kv * data[base + (size_t)k * stride]);
        ds += (double)(kv * dens[base + (size_t)k * stride]);
    }
    outData[i] = (float)is;
    outDens[i] = (float)ds;
}

__global__ void __launch_bounds__(kBlock)
k_conv_normalise(float *__restrict__ data, const float *__restrict__ dens, const uint8_t *__restrict__ isnan_, size_t N) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    data[i] = isnan_[i] ? __builtin_nanf("") : data[i] / dens[i];
}

// ------------------------------------------------------------------ voxel -> node
// reg_voxelCentricToNodeCentric, update = false (reg-lib/cpu/_reg_localTrans.cpp:1540-1684)
__global__ void __launch_bounds__(kBlock)
k_voxel_to_node(const float *__restrict__ vg, int nx, int ny, int nz, int gnx, int gny, int gnz, Affine T,
                Affine reorient, float weight, float *__restrict__ ng) {
#pragma clang fp contract(off)
    const size_t G = (size_t)gnx * gny * gnz, N = (size_t)nx * ny * nz;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= G) return;
    int x, y, z;
    index_xyz(n, gnx, gny, x, y, z);
    const float3 vc = affine_apply_d(T, (float)x, (float)y, (float)z);
    const int pre[3] = { floor_int(vc.x), floor_int(vc.y), floor_int(vc.z) };
    float bx[2], by[2], bz[2];
    bx[1] = vc.x - (float)pre[0]; bx[0] = 1.f - bx[1];
    by[1] = vc.y - (float)pre[1]; by[0] = 1.f - by[1];
    bz[1] = vc.z - (float)pre[2]; bz[0] = 1.f - bz[1];
    float iv0 = 0.f, iv1 = 0.f, iv2 = 0.f;
    for (int c = 0; c < 2; ++c) {
        const int Z = pre[2] + c;
        if (Z < 0 || Z >= nz) continue;
        for (int b = 0; b < 2; ++b) {
            const int Y = pre[1] + b;
            if (Y < 0 || Y >= ny) continue;
            for (int a = 0; a < 2; ++a) {
                const int X = pre[0] + a;
                if (X < 0 || X >= nx) continue;
                const size_t idx = ((size_t)Z * ny + Y) * nx + X;
                float w = bx[a] * by[b];
                w *= bz[c];
                iv0 += w * vg[idx];
                iv1 += w * vg[idx + N];
                iv2 += w * vg[idx + 2 * N];
            }
        }
    }
    const float *R = reorient.m;
    const float r0 = R[0] * iv0 + R[4] * iv1 + R[8] * iv2;
    const float r1 = R[1] * iv0 + R[5] * iv1 + R[9] * iv2;
    const float r2 = R[2] * iv0 + R[6] * iv1 + R[10] * iv2;
    ng[n] = r0 * weight;
    ng[n + G] = r1 * weight;
    ng[n + 2 * G] = r2 * weight;
}

// ------------------------------------------------------------------ bending energy
// (the basis, brick grid and value block body are in be_blocks.hpp)

// reg_spline_approxBendingEnergyGradient3D (regul.cpp:287-451) on a brick: the displacement
// (cpp - node) of the brick with a two-node halo in LDS, the 18 second derivatives of every node
// of the brick with a one-node halo (regul.cpp:317-379, the cross terms doubled in double), then
// the adjoint 27-point gather grad += weight / G * value (regul.cpp:381-448) for the brick's
// nodes.  Optionally tracks the max length of the final gradient.  The reference's in-place
// (p - node) + node round trip of the points is k_round_trip, run after this kernel.  RAW: the
// gathered sums alone are written to grad (planar [3][G]) for a later grad += weight / G * value.
template<bool RAW = false>
__device__ __forceinline__ void
be_gradient_block(const float *__restrict__ cpp, int gnx, int gny, int gnz, BrickGrid bg, Affine xyz, BeBasis B,
              float ratio, float *__restrict__ grad, uint32_t *__restrict__ maxBits, int bid) {
#pragma clang fp contract(off)
    constexpr int CX = BBX + 4, CY = BBY + 4, CZ = BBZ + 4, CN = CX * CY * CZ;  // points: two-node halo
    constexpr int DX = BBX + 2, DY = BBY + 2, DZ = BBZ + 2, DN = DX * DY * DZ;  // derivatives: one-node halo
    __shared__ float sC[3][CN];
    __shared__ float sD[18][DN];
    __shared__ float sB[27][8];  // basis values per neighbour (xx, yy, zz, xy, yz, xz), read as broadcasts
    const size_t G = (size_t)gnx * gny * gnz;
    int x0, y0, z0;
    brick_origin(bg, bid, x0, y0, z0);
    if (threadIdx.x < 27 * 6) sB[threadIdx.x / 6][threadIdx.x % 6] = B.b[threadIdx.x % 6][threadIdx.x / 6];
    {
        constexpr int kIter = (3 * CN + 255) / 256;
        float v[kIter];
#pragma unroll
        for (int u = 0; u < kIter; ++u) {
            // clamped, unconditional loads keep v in registers (out-of-grid entries are never read)
            const int i = min((int)threadIdx.x + u * 256, 3 * CN - 1);
            const int c = i / CN, r = i % CN, hx = r % CX, hy = (r / CX) % CY, hz = r / (CX * CY);
            const int X = min(max(x0 - 2 + hx, 0), gnx - 1), Y = min(max(y0 - 2 + hy, 0), gny - 1),
                      Z = min(max(z0 - 2 + hz, 0), gnz - 1);
            v[u] = cpp[c * G + ((size_t)Z * gny + Y) * gnx + X] - node_pos(xyz, c, X, Y, Z);
        }
#pragma unroll
        for (int u = 0; u < kIter; ++u) {
            const int i = threadIdx.x + u * 256;
            if (i < 3 * CN) sC[i / CN][i % CN] = v[u];
        }
    }
    __syncthreads();
    for (int r = threadIdx.x; r < DN; r += 256) {
        const int hx = r % DX, hy = (r / DX) % DY, hz = r / (DX * DY);
        const int x = x0 - 1 + hx, y = y0 - 1 + hy, z = z0 - 1 + hz;
        float d[3][6] = {};
        if (x >= 0 && y >= 0 && z >= 0 && x < gnx && y < gny && z < gnz) {
            int i = 0;
#pragma unroll
            for (int c = -1; c < 2; ++c)
#pragma unroll
                for (int b = -1; b < 2; ++b)
#pragma unroll
                    for (int a = -1; a < 2; ++a, ++i) {
                        const int X = x + a, Y = y + b, Z = z + c;
                        if (X < 0 || Y < 0 || Z < 0 || X >= gnx || Y >= gny || Z >= gnz) continue;
                        const int h = ((hz + 1 + c) * CY + hy + 1 + b) * CX + hx + 1 + a;
                        float bt[6];
#pragma unroll
                        for (int t = 0; t < 6; ++t) bt[t] = sB[i][t];
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            const float cv = sC[k][h];
#pragma unroll
                            for (int t = 0; t < 6; ++t) d[k][t] += bt[t] * cv;
                        }
                    }
        }
        for (int t = 0; t < 3; ++t)
            for (int k = 0; k < 3; ++k) sD[t * 3 + k][r] = d[k][t];
        for (int t = 3; t < 6; ++t)
            for (int k = 0; k < 3; ++k) sD[t * 3 + k][r] = (float)(2.0 * (double)d[k][t]);
    }
    __syncthreads();
    const int tx = threadIdx.x % BBX, ty = (threadIdx.x / BBX) % BBY, tz = threadIdx.x / (BBX * BBY);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    const bool valid = x < gnx && y < gny && z < gnz;
    float g0 = 0.f, g1 = 0.f, g2 = 0.f;
    if (valid) {
        const size_t n = ((size_t)z * gny + y) * gnx + x;
        float gv[3] = { 0.f, 0.f, 0.f };
        int a = 0;
#pragma unroll
        for (int c = -1; c < 2; ++c)
#pragma unroll
            for (int b = -1; b < 2; ++b)
#pragma unroll
                for (int q = -1; q < 2; ++q, ++a) {
                    const int X = x + q, Y = y + b, Z = z + c;
                    if (X < 0 || Y < 0 || Z < 0 || X >= gnx || Y >= gny || Z >= gnz) continue;
                    const int r = ((tz + 1 + c) * DY + ty + 1 + b) * DX + tx + 1 + q;
                    float bt[6];
#pragma unroll
                    for (int t = 0; t < 6; ++t) bt[t] = sB[a][t];
#pragma unroll
                    for (int t = 0; t < 6; ++t)
#pragma unroll
                        for (int k = 0; k < 3; ++k) gv[k] += sD[t * 3 + k][r] * bt[t];
                }
        if (RAW) {
            grad[n] = gv[0];
            grad[n + G] = gv[1];
            grad[n + 2 * G] = gv[2];
        } else {
            g0 = grad[n] + ratio * gv[0];
            g1 = grad[n + G] + ratio * gv[1];
            g2 = grad[n + 2 * G] + ratio * gv[2];
            grad[n] = g0;
            grad[n + G] = g1;
            grad[n + 2 * G] = g2;
        }
    }
    if (!RAW && maxBits) {
        // reg_getMaximalLength of the final gradient (lengths >= 0 order like their bit patterns)
        uint32_t bm = valid ? __float_as_uint(sqrtf(g0 * g0 + g1 * g1 + g2 * g2)) : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = __shfl_down(bm, off, 64);
            bm = o > bm ? o : bm;
        }
        if ((threadIdx.x & 63) == 0) atomicMax(maxBits, bm);
    }
}

// The in-place reg_getDisplacementFromDeformation / reg_getDeformationFromDisplacement pair the
// reference wraps around the bending-energy gradient (regul.cpp:312, 450): p <- (p - node) + node.
__device__ __forceinline__ void
round_trip_block(float *__restrict__ cpp, int gnx, int gny, int gnz, Affine xyz, int bid) {
#pragma clang fp contract(off)
    const size_t G = (size_t)gnx * gny * gnz;
    const size_t n = (size_t)bid * blockDim.x + threadIdx.x;
    if (n >= G) return;
    int x, y, z;
    index_xyz(n, gnx, gny, x, y, z);
    for (int k = 0; k < 3; ++k) {
        const float node = node_pos(xyz, k, x, y, z);
        cpp[n + k * G] = (cpp[n + k * G] - node) + node;
    }
}

// ------------------------------------------------------------------ linear energy
struct LeBasis {
    float b[3][27];  // first-order tensor products along x, y, z
};

// Rotation-free displacement Jacobian at an interior node (reg-lib/cpu/_reg_localTrans_regul.cpp:596-634).
// B: the first-order basis in LDS (broadcast reads; as a kernel argument its 81 values would not
// fit the scalar registers).
__device__ M33 le_matrix(const float *__restrict__ cpp, size_t G, int gnx, int gny, int x, int y, int z,
                         const float (*B)[27], const M33& reo) {
#pragma clang fp contract(off)
    M33 M = {};
    int i = 0;
#pragma unroll
    for (int c = -1; c < 2; ++c)
#pragma unroll
        for (int b = -1; b < 2; ++b)
#pragma unroll
            for (int a = -1; a < 2; ++a, ++i) {
                const size_t idx = ((size_t)(z + c) * gny + y + b) * gnx + x + a;
                const float cv[3] = { cpp[idx], cpp[idx + G], cpp[idx + 2 * G] };
                const float b0 = B[0][i], b1 = B[1][i], b2 = B[2][i];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    M.m[0][k] += b0 * cv[k];
                    M.m[1][k] += b1 * cv[k];
                    M.m[2][k] += b2 * cv[k];
                }
            }
    M = m33_mul(reo, M);
    const M33 R = m33_inverse(m33_polar(M));
    M = m33_mul(R, M);
    M.m[0][0] -= 1.f;
    M.m[1][1] -= 1.f;
    M.m[2][2] -= 1.f;
    return M;
}

// Per-node terms of reg_spline_approxLinearEnergyValue3D (:559-647) and the diagonal the gradient
// uses; the last block to finish sums the block partials in block order into *out (/ 3G).
__global__ void __launch_bounds__(kBlock)
k_le_terms(const float *__restrict__ cpp, int gnx, int gny, int gnz, LeBasis B, M33 reo, float *__restrict__ diag,
           double *__restrict__ partials, uint32_t *__restrict__ ticket, double scale, double *__restrict__ out) {
#pragma clang fp contract(off)
    __shared__ double red[16];
    __shared__ bool last;
    __shared__ float sB[3][27];
    if (threadIdx.x < 81) sB[threadIdx.x / 27][threadIdx.x % 27] = B.b[threadIdx.x / 27][threadIdx.x % 27];
    __syncthreads();
    const size_t G = (size_t)gnx * gny * gnz;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0;
    if (n < G) {
        int x, y, z;
    index_xyz(n, gnx, gny, x, y, z);
        float d0 = 0.f, d1 = 0.f, d2 = 0.f;
        if (x > 0 && y > 0 && z > 0 && x < gnx - 1 && y < gny - 1 && z < gnz - 1) {
            const M33 M = le_matrix(cpp, G, gnx, gny, x, y, z, sB, reo);
            for (int b = 0; b < 3; ++b)
                for (int a = 0; a < 3; ++a) {
                    const double v = 0.5 * (double)(M.m[a][b] + M.m[b][a]);
                    s += v * v;
                }
            d0 = M.m[0][0];
            d1 = M.m[1][1];
            d2 = M.m[2][2];
        }
        diag[n] = d0;
        diag[n + G] = d1;
        diag[n + 2 * G] = d2;
    }
    s = block_sum_double(s, red);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = s;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t = 0;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x) t += partials[i];
    t = block_sum_double(t, red);
    if (threadIdx.x == 0) {
        *out = t * scale;
        *ticket = 0u;
    }
}

// reg_spline_approxLinearEnergyGradient3D (:739-824) as a gather: each node adds the
// contributions of the interior nodes around it in the reference's source order, so every sum
// runs in the same order as the reference's scatter.  Optionally tracks the final max length.
__global__ void __launch_bounds__(kBlock)
k_le_grad(const float *__restrict__ diag, int gnx, int gny, int gnz, LeBasis B, M33 inv, float ratio,
          float *__restrict__ grad, uint32_t *__restrict__ maxBits) {
#pragma clang fp contract(off)
    __shared__ float sB[3][27];
    if (threadIdx.x < 81) sB[threadIdx.x / 27][threadIdx.x % 27] = B.b[threadIdx.x / 27][threadIdx.x % 27];
    __syncthreads();
    const size_t G = (size_t)gnx * gny * gnz;
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = n < G;
    const size_t m = valid ? n : 0;
    int x, y, z;
    index_xyz(m, gnx, gny, x, y, z);
    float g0 = grad[m], g1 = grad[m + G], g2 = grad[m + 2 * G];
    for (int c = 1; c > -2; --c)
        for (int b = 1; b > -2; --b)
            for (int a = 1; a > -2; --a) {
                const int X = x - a, Y = y - b, Z = z - c;  // the source node
                if (X < 1 || Y < 1 || Z < 1 || X > gnx - 2 || Y > gny - 2 || Z > gnz - 2) continue;
                const size_t s = ((size_t)Z * gny + Y) * gnx + X;
                const int i = 26 - ((c + 1) * 9 + (b + 1) * 3 + (a + 1));
                const float v0 = -2.f * diag[s] * sB[0][i];
                const float v1 = -2.f * diag[s + G] * sB[1][i];
                const float v2 = -2.f * diag[s + 2 * G] * sB[2][i];
                g0 += ratio * (inv.m[0][0] * v0 + inv.m[0][1] * v1 + inv.m[0][2] * v2);
                g1 += ratio * (inv.m[1][0] * v0 + inv.m[1][1] * v1 + inv.m[1][2] * v2);
                g2 += ratio * (inv.m[2][0] * v0 + inv.m[2][1] * v1 + inv.m[2][2] * v2);
            }
    if (valid) {
        grad[n] = g0;
        grad[n + G] = g1;
        grad[n + 2 * G] = g2;
    }
    if (maxBits) {
        uint32_t bits = valid ? __float_as_uint(sqrtf(g0 * g0 + g1 * g1 + g2 * g2)) : 0u;
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = __shfl_down(bits, off, 64);
            bits = o > bits ? o : bits;
        }
        if ((threadIdx.x & 63) == 0) atomicMax(maxBits, bits);
    }
}

// ------------------------------------------------------------------ optimiser helpers
// reg_getMaximalLength<float> (reg-lib/cpu/_reg_tools.cpp:335-353): lengths are >= 0, so their
// IEEE bit patterns order like unsigned integers.
__global__ void __launch_bounds__(kBlock)
k_max_length(const float *__restrict__ g, size_t G, uint32_t *__restrict__ out) {
#pragma clang fp contract(off)
    uint32_t best = 0;
    for (size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x; n < G; n += (size_t)gridDim.x * blockDim.x) {
        const float x = g[n], y = g[n + G], z = g[2 * G + n];
        const float l = sqrtf(x * x + y * y + z * z);
        const uint32_t b = __float_as_uint(l);
        best = b > best ? b : best;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t o = __shfl_down(best, off, 64);
        best = o > best ? o : best;
    }
    if ((threadIdx.x & 63) == 0) atomicMax(out, best);
}

__global__ void k_max_to_scalar(const uint32_t *__restrict__ bits, double *__restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = (double)__uint_as_float(*bits);
}

// Compute::NormaliseGradient (reg-lib/Compute.cpp:152-183)
__global__ void __launch_bounds__(kBlock)
k_normalise(float *__restrict__ g, size_t n, double inv) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) g[i] = (float)((double)g[i] * inv);
}

// Normalisation by the max length held as float bits; block 0 also publishes it as a double.
__device__ __forceinline__ void
normalise_bits_block(float *__restrict__ g, size_t n, const uint32_t *__restrict__ bits, double *__restrict__ maxLen, int bid) {
    const size_t i = (size_t)bid * blockDim.x + threadIdx.x;
    const double m = (double)__uint_as_float(*bits);
    if (i == 0) *maxLen = m;
    if (i < n && m != 0.0) g[i] = (float)((double)g[i] * (1.0 / m));
}

// Normalisation by a device-resident max length (fused gradient path): skips when it is 0.
__global__ void __launch_bounds__(kBlock)
k_normalise_dev(float *__restrict__ g, size_t n, const double *__restrict__ maxLen) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double m = *maxLen;
    if (i < n && m != 0.0) g[i] = (float)((double)g[i] * (1.0 / m));
}

// ConjugateGradient<float>::UpdateGradientValues (reg-lib/Optimiser.cpp:231-317)
__device__ __forceinline__ void
cg_init_block(float *__restrict__ grad, float *__restrict__ G_, float *__restrict__ H, size_t n, int bid) {
    const size_t i = (size_t)bid * blockDim.x + threadIdx.x;
    if (i < n) H[i] = G_[i] = -grad[i];
}

__device__ __forceinline__ void
cg_dots_block(const float *__restrict__ grad, const float *__restrict__ G_, const float *__restrict__ H, size_t n,
          double *__restrict__ partials, int bid, int nbk) {
#pragma clang fp contract(off)
    __shared__ double red[16];
    double gg = 0, dgg = 0;
    for (size_t i = (size_t)bid * blockDim.x + threadIdx.x; i < n; i += (size_t)nbk * blockDim.x) {
        gg += (double)(H[i] * G_[i]);
        dgg += (double)((grad[i] + G_[i]) * grad[i]);
    }
    gg = block_sum_double(gg, red);
    __syncthreads();
    dgg = block_sum_double(dgg, red);
    if (threadIdx.x == 0) {
        partials[2 * bid] = gg;
        partials[2 * bid + 1] = dgg;
    }
}

__device__ __forceinline__ void
cg_gamma_block(const double *__restrict__ partials, int nb, double *__restrict__ gam) {
    __shared__ double red[16];
    double gg = 0, dgg = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        gg += partials[2 * i];
        dgg += partials[2 * i + 1];
    }
    gg = block_sum_double(gg, red);
    __syncthreads();
    dgg = block_sum_double(dgg, red);
    if (threadIdx.x == 0) *gam = dgg / gg;
}

__device__ __forceinline__ void
cg_apply_block(float *__restrict__ grad, float *__restrict__ G_, float *__restrict__ H, size_t n, const double *__restrict__ gam, int bid) {
#pragma clang fp contract(off)
    const size_t i = (size_t)bid * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double gm = *gam;
    G_[i] = -grad[i];
    H[i] = (float)((double)G_[i] + gm * (double)H[i]);
    grad[i] = -H[i];
}

// Compute::UpdateControlPointPosition (reg-lib/Compute.cpp:91-127)
__device__ __forceinline__ void
update_dof_block(float *__restrict__ cur, const float *__restrict__ best, const float *__restrict__ grad, float scale, size_t n, int bid) {
#pragma clang fp contract(off)
    const size_t i = (size_t)bid * blockDim.x + threadIdx.x;
    if (i < n) cur[i] = best[i] + scale * grad[i];
}

// ------------------------------------------------------------------ intensity rescale
// reg_intensityRescale_core<float> (reg-lib/cpu/_reg_tools.cpp:93-183): min/max over finite values.
__global__ void __launch_bounds__(kBlock)
k_minmax(const float *__restrict__ d, size_t n, uint32_t *__restrict__ mm) {
    float lo = INFINITY, hi = -INFINITY;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float v = d[i];
        if (v == v) { lo = fminf(lo, v); hi = fmaxf(hi, v); }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_down(lo, off, 64));
        hi = fmaxf(hi, __shfl_down(hi, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        // order-preserving integer keys of the floats
        auto key = [](float f) {
            uint32_t u = __float_as_uint(f);
            return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
        };
        atomicMin(&mm[0], key(lo));
        atomicMax(&mm[1], key(hi));
    }
}

__device__ __forceinline__ float unkey(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    return __uint_as_float(u);
}

__global__ void __launch_bounds__(kBlock)
k_rescale(float *__restrict__ d, size_t n, const uint32_t *__restrict__ mm, float newMin, float newMax) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float lo = unkey(mm[0]), hi = unkey(mm[1]);
    const double diff = (double)(hi - lo), newDiff = (double)(newMax - newMin);
    double v = (double)d[i];
    if (v == v) {
        v = (v - (double)lo) / diff;
        v = v * newDiff + (double)newMin;
    }
    d[i] = (float)v;
}

__global__ void k_fill_u32(uint32_t *p, size_t n, uint32_t v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_f32(float *p, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------ single-pair kernels
// Each wraps the block body above with the launch's own block index and count.
// The single-pair finalisation (one workgroup).  pub (may be null): it then publishes ent[0..4]
// and s1[0..n1) to coherent host memory followed by the sequence number seq (k_publish's
// layout), so the objective needs no publishing kernel of its own.
__global__ void __launch_bounds__(1024)
k_nmi_finalize_lds(uint32_t *__restrict__ counts, int rb, int fb, double *__restrict__ pro,
                   double *__restrict__ logh, float *__restrict__ logf, double *__restrict__ ent,
                   float *__restrict__ kTab, double *__restrict__ pub, const double *__restrict__ s1, int n1,
                   uint64_t seq) {
    nmi_finalize_lds_block(counts, rb, fb, pro, logh, logf, ent, kTab);
    if (pub) {
        __syncthreads();
        const int i = threadIdx.x;
        if (i < 5) pub[i] = ent[i];
        else if (i < 5 + n1) pub[i] = s1[i - 5];
        __threadfence_system();
        __syncthreads();
        if (i == 0) __hip_atomic_store((uint64_t *)(pub + 15), seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
    }
}

__global__ void __launch_bounds__(256)
k_be_value(const float *__restrict__ cpp, int gnx, int gny, int gnz, BrickGrid bg, BeBasis B,
           double *__restrict__ partials, uint32_t *__restrict__ ticket, double scale, double *__restrict__ out) {
    be_value_block(cpp, gnx, gny, gnz, bg, B, partials, ticket, scale, out, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(256)
k_be_gradient(const float *__restrict__ cpp, int gnx, int gny, int gnz, BrickGrid bg, Affine xyz, BeBasis B,
              float ratio, float *__restrict__ grad, uint32_t *__restrict__ maxBits) {
    be_gradient_block(cpp, gnx, gny, gnz, bg, xyz, B, ratio, grad, maxBits, blockIdx.x);
}

__global__ void __launch_bounds__(256)
k_be_gradient_raw(const float *__restrict__ cpp, int gnx, int gny, int gnz, BrickGrid bg, Affine xyz, BeBasis B,
                  float *__restrict__ out) {
    be_gradient_block<true>(cpp, gnx, gny, gnz, bg, xyz, B, 0.f, out, nullptr, blockIdx.x);
}

__global__ void __launch_bounds__(256) k_round_trip(float *__restrict__ cpp, int gnx, int gny, int gnz, Affine xyz) {
    round_trip_block(cpp, gnx, gny, gnz, xyz, blockIdx.x);
}

__global__ void __launch_bounds__(kBlock)
k_normalise_bits(float *__restrict__ g, size_t n, const uint32_t *__restrict__ bits, double *__restrict__ maxLen) {
    normalise_bits_block(g, n, bits, maxLen, blockIdx.x);
}

__global__ void __launch_bounds__(kBlock) k_cg_init(float *__restrict__ grad, float *__restrict__ G_, float *__restrict__ H,
                                                  size_t n) {
    cg_init_block(grad, G_, H, n, blockIdx.x);
}

__global__ void __launch_bounds__(kBlock)
k_cg_dots(const float *__restrict__ grad, const float *__restrict__ G_, const float *__restrict__ H, size_t n,
          double *__restrict__ partials) {
    cg_dots_block(grad, G_, H, n, partials, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(1024) k_cg_gamma(const double *__restrict__ partials, int nb, double *__restrict__ gam) {
    cg_gamma_block(partials, nb, gam);
}

__global__ void __launch_bounds__(kBlock)
k_cg_apply(float *__restrict__ grad, float *__restrict__ G_, float *__restrict__ H, size_t n, const double *__restrict__ gam) {
    cg_apply_block(grad, G_, H, n, gam, blockIdx.x);
}

__global__ void __launch_bounds__(kBlock)
k_update_dof(float *__restrict__ cur, const float *__restrict__ best, const float *__restrict__ grad, float scale, size_t n) {
    update_dof_block(cur, best, grad, scale, n, blockIdx.x);
}

// ------------------------------------------------------------------ batched kernels
// A batch of pairs runs every small kernel of an evaluation in one launch: the blocks of all
// pairs are listed in an item table (pair, block of that pair), and each block runs the same body
// as the single-pair kernel with that pair's arguments, its own block index and count, so every
// pair's arithmetic and reduction order are those of the pair alone.  Pairs whose flag in
// `active` is 0 are skipped.
struct PairOps {
    float *cpp, *best, *grad, *cgG, *cgH;
    int gnx, gny, gnz;
    size_t G;
    Affine xyz;
    BrickGrid bg;
    int beBlocks, cgBlocks;
    float beRatio;         // bending-energy weight / G
    double beScale;        // 1 / (3 G)
    double *partials;      // bending-energy value and CG dot partials
    uint32_t *beTicket;
    double *beOut;         // bending energy (scalars[0])
    uint32_t *maxBits;     // max gradient length as float bits
    double *maxLen;        // scalars[1]
    double *gam;           // scalars[2]
    uint32_t *counts;
    double *pro, *logh;
    float *logf;
    double *ent;
    int rb, fb;
};

__device__ __forceinline__ bool skip_pair(const int *active, int pair) { return active && !active[pair]; }

__global__ void __launch_bounds__(1024) k_nmi_finalize_batch(const PairOps *__restrict__ P, const int *active) {
    if (skip_pair(active, blockIdx.x)) return;
    const PairOps& q = P[blockIdx.x];
    nmi_finalize_lds_block(q.counts, q.rb, q.fb, q.pro, q.logh, q.logf, q.ent);
}

__global__ void __launch_bounds__(256)
k_be_value_batch(const PairOps *__restrict__ P, const int2 *__restrict__ items, BeBasis B, const int *active) {
    const int2 it = items[blockIdx.x];
    if (skip_pair(active, it.x)) return;
    const PairOps& q = P[it.x];
    be_value_block(q.cpp, q.gnx, q.gny, q.gnz, q.bg, B, q.partials, q.beTicket, q.beScale, q.beOut, it.y, q.beBlocks);
}

__global__ void __launch_bounds__(256)
k_be_gradient_batch(const PairOps *__restrict__ P, const int2 *__restrict__ items, BeBasis B, const int *active) {
    const int2 it = items[blockIdx.x];
    if (skip_pair(active, it.x)) return;
    const PairOps& q = P[it.x];
    be_gradient_block(q.cpp, q.gnx, q.gny, q.gnz, q.bg, q.xyz, B, q.beRatio, q.grad, q.maxBits, it.y);
}

__global__ void __launch_bounds__(256)
k_round_trip_batch(const PairOps *__restrict__ P, const int2 *__restrict__ items, const int *active) {
    const int2 it = items[blockIdx.x];
    if (skip_pair(active, it.x)) return;
    const PairOps& q = P[it.x];
    round_trip_block(q.cpp, q.gnx, q.gny, q.gnz, q.xyz, it.y);
}

// items over 3 G: element-wise kernels of the control-point arrays
__global__ void __launch_bounds__(kBlock)
k_normalise_batch(const PairOps *__restrict__ P, const int2 *__restrict__ items, const int *active) {
    const int2 it = items[blockIdx.x];
    if (skip_pair(active, it.x)) return;
    const PairOps& q = P[it.x];
    normalise_bits_block(q.grad, 3 * q.G, q.maxBits, q.maxLen, it.y);
}

// first[pair]: 1 = steepest descent (reset G, H), 0 = conjugate update
__global__ void __launch_bounds__(kBlock)
k_cg_init_batch(const PairOps *__restrict__ P, const int2 *__restrict__ items, const int *active,
                const int *__restrict__ first) {
    const int2 it = items[blockIdx.x];
    if (skip_pair(active, it.x) || !first[it.x]) return;
    const PairOps& q = P[it.x];
    cg_init_block(q.grad, q.cgG, q.cgH, 3 * q.G, it.y);
}

__global__ void __launch_bounds__(kBlock)
k_cg_dots_batch(const PairOps *__restrict__ P, const int2 *__restrict__ items, const int *active,
                const int *__restrict__ first) {
    const int2 it = items[blockIdx.x];
    if (skip_pair(active, it.x) || first[it.x]) return;
    const PairOps& q = P[it.x];
    cg_dots_block(q.grad, q.cgG, q.cgH, 3 * q.G, q.partials, it.y, q.cgBlocks);
}

__global__ void __launch_bounds__(1024)
k_cg_gamma_batch(const PairOps *__restrict__ P, const int *active, const int *__restrict__ first) {
    if (skip_pair(active, blockIdx.x) || first[blockIdx.x]) return;
    const PairOps& q = P[blockIdx.x];
    cg_gamma_block(q.partials, q.cgBlocks, q.gam);
}

__global__ void __launch_bounds__(kBlock)
k_cg_apply_batch(const PairOps *__restrict__ P, const int2 *__restrict__ items, const int *active,
                 const int *__restrict__ first) {
    const int2 it = items[blockIdx.x];
    if (skip_pair(active, it.x) || first[it.x]) return;
    const PairOps& q = P[it.x];
    cg_apply_block(q.grad, q.cgG, q.cgH, 3 * q.G, q.gam, it.y);
}

__global__ void __launch_bounds__(kBlock)
k_update_dof_batch(const PairOps *__restrict__ P, const int2 *__restrict__ items, const int *active,
                   const float *__restrict__ scales) {
    const int2 it = items[blockIdx.x];
    if (skip_pair(active, it.x)) return;
    const PairOps& q = P[it.x];
    update_dof_block(q.cpp, q.best, q.grad, scales[it.x], 3 * q.G, it.y);
}

// toBest: best <- cpp (Optimiser::StoreCurrentDof), else cpp <- best (RestoreBestDof)
__global__ void __launch_bounds__(kBlock)
k_copy_dof_batch(const PairOps *__restrict__ P, const int2 *__restrict__ items, const int *active, int toBest) {
    const int2 it = items[blockIdx.x];
    if (skip_pair(active, it.x)) return;
    const PairOps& q = P[it.x];
    const size_t i = (size_t)it.y * blockDim.x + threadIdx.x;
    if (i >= 3 * q.G) return;
    if (toBest) q.best[i] = q.cpp[i];
    else q.cpp[i] = q.best[i];
}

__global__ void k_zero_max_bits_batch(const PairOps *__restrict__ P, int n, const int *active) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !skip_pair(active, i)) *P[i].maxBits = 0u;
}

// Round to the 6-decimal literals the reference tabulates (reg-lib/cpu/_reg_splineBasis.cpp:292-457).
float round6(double v) {
    const double r = std::floor(std::fabs(v) * 1e6 + 0.5) / 1e6;
    return (float)(v < 0 ? -r : r);
}

BeBasis make_be_basis() {
    // Tensor products of the cubic B-spline at a node: value {1/6, 2/3, 1/6}, first
    // derivative {-1/2, 0, 1/2}, second derivative {1, -2, 1}.
    const double v[3] = { 1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0 }, d1[3] = { -0.5, 0.0, 0.5 }, d2[3] = { 1.0, -2.0, 1.0 };
    BeBasis B;
    int i = 0;
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a, ++i) {
                B.b[0][i] = round6(d2[a] * v[b] * v[c]);
                B.b[1][i] = round6(v[a] * d2[b] * v[c]);
                B.b[2][i] = round6(v[a] * v[b] * d2[c]);
                B.b[3][i] = round6(d1[a] * d1[b] * v[c]);
                B.b[4][i] = round6(v[a] * d1[b] * d1[c]);
                B.b[5][i] = round6(d1[a] * v[b] * d1[c]);
            }
    return B;
}

}  // namespace

BeBasis be_basis() {
    static const BeBasis B = make_be_basis();
    return B;
}

// ------------------------------------------------------------------ launches
void op_fill_u32(uint32_t *p, size_t n, uint32_t v, hipStream_t s) {
    const unsigned b = std::min(blocks_for(n), 4096u);
    hipLaunchKernelGGL(k_fill_u32, dim3(b), dim3(kBlock), 0, s, p, n, v);
}

void op_fill_f32(float *p, size_t n, float v, hipStream_t s) {
    const unsigned b = std::min(blocks_for(n), 4096u);
    hipLaunchKernelGGL(k_fill_f32, dim3(b), dim3(kBlock), 0, s, p, n, v);
}

static bool lut_rule(const nrh_f3d *h, float g[3]) {
    g[0] = h->cpp.dx / h->ref.dx;
    g[1] = h->cpp.dy / h->ref.dy;
    g[2] = h->cpp.dz / h->ref.dz;
    return g[0] == 5.f && g[1] == 5.f && g[2] == 5.f;
}

void op_spline_field(nrh_f3d *h) {
    ensure_field(h);
    float g[3];
    const int lut = lut_rule(h, g) ? 1 : 0;
    // every voxel's 4x4x4 support must lie inside the grid (the reference would extrapolate
    // through get_SlidedValues for a grid that does not cover the image)
    const int dims[3] = { h->ref.nx, h->ref.ny, h->ref.nz }, gd[3] = { h->cpp.nx, h->cpp.ny, h->cpp.nz };
    for (int i = 0; i < 3; ++i) {
        const int c = dims[i] - 1;
        const int pre = lut ? c / 5 : (int)((float)c / g[i]);
        NRH_REQUIRE(pre + 3 < gd[i], NRH_ERR_UNSUPPORTED, "control point grid does not cover the reference image");
    }
    h->velFieldVersion = 0;  // the field no longer holds a velocity exponentiation
    h->pairStateVersion = 0;
    hipLaunchKernelGGL(k_spline_field, dim3(blocks_for(h->N)), dim3(kBlock), 0, h->stream, h->ref.nx, h->ref.ny,
                       h->ref.nz, h->cpp.nx, h->cpp.ny, h->G, h->cppPos.p, h->mask.p, g[0], g[1], g[2], lut,
                       h->params.field_order, h->field.p, h->N);
    NRH_HIP(hipGetLastError());
}

void op_resample(nrh_f3d *h) {
    h->pairStateVersion = 0;
    ensure_warped(h);
    NRH_REQUIRE(!h->field.empty(), NRH_ERR_STATE, "resample before the deformation field");
    NRH_REQUIRE(!h->floImg.empty(), NRH_ERR_UNSUPPORTED, "per-operation resampling needs float32 image storage");
    hipLaunchKernelGGL(k_resample_linear, dim3(blocks_for(h->N)), dim3(kBlock), 0, h->stream, h->field.p, h->N,
                       h->mask.p, h->floImg.p, h->flo.nx, h->flo.ny, h->flo.nz, to_affine(h->flo.ijk),
                       h->params.padding, h->warped.p);
    NRH_HIP(hipGetLastError());
}

void op_image_gradient(nrh_f3d *h) {
    ensure_warped_grad(h);
    NRH_REQUIRE(!h->field.empty(), NRH_ERR_STATE, "image gradient before the deformation field");
    NRH_REQUIRE(!h->floImg.empty(), NRH_ERR_UNSUPPORTED, "per-operation gradient needs float32 image storage");
    hipLaunchKernelGGL(k_image_gradient_linear, dim3(blocks_for(h->N)), dim3(kBlock), 0, h->stream, h->field.p,
                       h->N, h->mask.p, h->floImg.p, h->flo.nx, h->flo.ny, h->flo.nz, to_affine(h->flo.ijk),
                       h->params.padding, h->warpedGrad.p);
    NRH_HIP(hipGetLastError());
}

void op_joint_histogram(nrh_f3d *h, const float *warped) {
    h->pairStateVersion = 0;
    const int rb = h->params.ref_bins, fb = h->params.flo_bins;
    // histCount is zero here: allocation zeroes it and every finalisation consumes it
    const unsigned b = std::min(blocks_for(h->N), 2048u);
    hipLaunchKernelGGL(k_joint_histogram, dim3(b), dim3(kBlock), (size_t)rb * fb * sizeof(uint32_t), h->stream,
                       ref_ptr(h), warped, h->mask.p, h->N, rb, fb, h->histCount.p);
    NRH_HIP(hipGetLastError());
}

void op_nmi_finalize(nrh_f3d *h, float *kTab, const double *pubS1, int pubN1) {
    h->kTabCurrent = kTab != nullptr && h->totalBins() <= kFinLds;
    NRH_REQUIRE(pubN1 < 0 || (h->totalBins() <= kFinLds && pubN1 <= 10), NRH_ERR_ARG, "finalisation cannot publish");
    if (h->totalBins() <= kFinLds) {
        const bool pub = pubN1 >= 0;
        hipLaunchKernelGGL(k_nmi_finalize_lds, dim3(1), dim3(1024), 0, h->stream, h->histCount.p, h->params.ref_bins,
                           h->params.flo_bins, h->histPro.p, h->histLog.p, h->histLogF.p, h->ent.p,
                           h->kTabCurrent ? kTab : nullptr, pub ? h->publishedDev : nullptr, pubS1, pub ? pubN1 : 0,
                           pub ? ++h->publishSeq : (uint64_t)0);
    } else {
        hipLaunchKernelGGL(k_nmi_finalize, dim3(1), dim3(1024), 0, h->stream, h->histCount.p, h->params.ref_bins,
                           h->params.flo_bins, h->histPro.p, h->histLog.p, h->histLogF.p, h->ent.p);
    }
    NRH_HIP(hipGetLastError());
}

void op_nmi_voxel_gradient(nrh_f3d *h, double weight) {
    ensure_voxel_grad(h);
    NRH_REQUIRE(!h->warped.empty() && !h->warpedGrad.empty(), NRH_ERR_STATE,
                "NMI voxel gradient needs the warped image and its gradient");
    const size_t logBytes = (size_t)h->totalBins() * sizeof(double);
    if (logBytes <= 64 * 1024) {
        // four workgroups per CU on 256 CUs, each staging the log histogram once
        const unsigned grid = std::min(blocks_for(h->N), 1024u);
        hipLaunchKernelGGL(k_nmi_voxel_gradient_lds, dim3(grid), dim3(kBlock), logBytes, h->stream, ref_ptr(h),
                           h->warped.p, h->warpedGrad.p, h->mask.p, h->N, h->params.ref_bins, h->params.flo_bins,
                           h->histLog.p, h->ent.p, weight, h->voxelGrad.p);
    } else {
        hipLaunchKernelGGL(k_nmi_voxel_gradient, dim3(blocks_for(h->N)), dim3(kBlock), 0, h->stream, ref_ptr(h),
                           h->warped.p, h->warpedGrad.p, h->mask.p, h->N, h->params.ref_bins, h->params.flo_bins,
                           h->histLog.p, h->ent.p, weight, h->voxelGrad.p);
    }
    NRH_HIP(hipGetLastError());
}

void op_convolve_voxel_gradient(nrh_f3d *h) {
    NRH_REQUIRE(!h->voxelGrad.empty(), NRH_ERR_STATE, "no voxel-based gradient to convolve");
    const size_t N = h->N;
    h->convTmp.alloc(N);
    h->convDens.alloc(N);
    h->convDensTmp.alloc(N);
    h->convNan.alloc(N);
    const int dims[3] = { h->ref.nx, h->ref.ny, h->ref.nz };
    const float pix[3] = { h->ref.dx, h->ref.dy, h->ref.dz };
    const float sp[3] = { h->cpp.dx, h->cpp.dy, h->cpp.dz };
    for (int axis = 0; axis < 3; ++axis) {
        // Cubic B-spline kernel of reg_tools_kernelConvolution (reg-lib/cpu/_reg_tools.cpp:868-904)
        const double temp = (double)(sp[axis] / pix[axis]);
        const int radius = (int)(temp * 2.0f);
        std::vector<float> kern(2 * radius + 1);
        for (int i = -radius; i <= radius; ++i) {
            const double rel = std::fabs(i / temp);
            float kv = 0.f;
            if (rel < 1.0) kv = (float)(2.0 / 3.0 - rel * rel + 0.5 * rel * rel * rel);
            else if (rel < 2.0) kv = (float)(-((rel - 2.0) * (rel - 2.0) * (rel - 2.0)) / 6.0);
            kern[i + radius] = kv;
        }
        if (radius > 0) {
            h->convKernel.alloc(std::max<size_t>(kern.size(), 64));
            NRH_HIP(hipMemcpyAsync(h->convKernel.p, kern.data(), kern.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
        }
        for (int comp = 0; comp < 3; ++comp) {
            float *d = h->voxelGrad.p + comp * N;
            hipLaunchKernelGGL(k_conv_prepare, dim3(blocks_for(N)), dim3(kBlock), 0, h->stream, d, (const int *)nullptr,
                               N, h->convDens.p, h->convNan.p);
            if (radius > 0 && dims[axis] > 1) {
                hipLaunchKernelGGL(k_conv_axis, dim3(blocks_for(N)), dim3(kBlock), 0, h->stream, d, h->convDens.p,
                                   dims[0], dims[1], dims[2], axis, h->convKernel.p, radius, h->convTmp.p, h->convDensTmp.p,
                                   h->params.field_order);
                NRH_HIP(hipMemcpyAsync(d, h->convTmp.p, N * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
                hipLaunchKernelGGL(k_conv_normalise, dim3(blocks_for(N)), dim3(kBlock), 0, h->stream, d, h->convDensTmp.p,
                                   h->convNan.p, N);
            } else {
                hipLaunchKernelGGL(k_conv_normalise, dim3(blocks_for(N)), dim3(kBlock), 0, h->stream, d, h->convDens.p,
                                   h->convNan.p, N);
            }
        }
    }
    NRH_HIP(hipGetLastError());
}

static Affine mat_mul_affine(const float *a, const float *b) {
    // mat44 product in double stored in float (reg-lib/cpu/Maths.hpp:235-245)
    Affine r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 4; ++j)
            r.m[i * 4 + j] = (float)((double)a[i * 4 + 0] * (double)b[0 * 4 + j] + (double)a[i * 4 + 1] * (double)b[1 * 4 + j] +
                                     (double)a[i * 4 + 2] * (double)b[2 * 4 + j] + (double)a[i * 4 + 3] * (double)b[3 * 4 + j]);
    return r;
}

float node_weight_ratio(const nrh_f3d *h, float weight) {
    // weight *= node spacing / voxel spacing per axis (reg-lib/cpu/_reg_localTrans.cpp:1595-1604)
    const float pd[3] = { h->ref.dx, h->ref.dy, h->ref.dz };
    const float gd[3] = { h->cpp.dx, h->cpp.dy, h->cpp.dz };
    for (int i = 0; i < 3; ++i) {
        float r = gd[i];
        if (h->cpp.sform) {
            const float *m = h->cpp.xyz;
            r = std::sqrt(m[i * 4 + 0] * m[i * 4 + 0] + m[i * 4 + 1] * m[i * 4 + 1] + m[i * 4 + 2] * m[i * 4 + 2]);
        }
        r /= pd[i];
        weight *= r;
    }
    return weight;
}

void op_voxel_to_node(nrh_f3d *h, float weight) {
    NRH_REQUIRE(!h->voxelGrad.empty(), NRH_ERR_STATE, "no voxel-based gradient");
    Affine T = mat_mul_affine(h->ref.ijk, h->cpp.xyz);
    Affine R = to_affine(h->flo.ijk);
    if (h->velocity && h->velExt0) {
        // the grid's extension 0 maps the grid into the reference space (reg-lib/cpu/
        // _reg_localTrans.cpp:1568-1592): inverse(A) between the grid and the image, and
        // inverse(A)'s 3x3 times the floating orientation in the reorientation
        float inv[16], gridToRef[16];
        nifti_inverse(h->velExt0M, inv);
        mat44_product(inv, h->cpp.xyz, gridToRef);
        T = mat_mul_affine(h->ref.ijk, gridToRef);
        M33 a, o;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                a.m[i][j] = h->velExt0M[i * 4 + j];
                o.m[i][j] = h->flo.ijk[i * 4 + j];
            }
        const M33 r = m33_mul(m33_inverse(a), o);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) R.m[i * 4 + j] = r.m[i][j];
    }
    const float w = node_weight_ratio(h, weight);
    hipLaunchKernelGGL(k_voxel_to_node, dim3(blocks_for(h->G)), dim3(kBlock), 0, h->stream, h->voxelGrad.p, h->ref.nx,
                       h->ref.ny, h->ref.nz, h->cpp.nx, h->cpp.ny, h->cpp.nz, T, R, w, h->nodeGrad.p);
    NRH_HIP(hipGetLastError());
}

static BrickGrid brick_grid(const nrh_f3d *h) {
    BrickGrid g;
    g.bx = (h->cpp.nx + BBX - 1) / BBX;
    g.by = (h->cpp.ny + BBY - 1) / BBY;
    g.bz = (h->cpp.nz + BBZ - 1) / BBZ;
    return g;
}

void op_bending_energy(nrh_f3d *h, double *out_device, hipStream_t stream) {
    static const BeBasis B = make_be_basis();
    const BrickGrid bg = brick_grid(h);
    const unsigned nb = (unsigned)(bg.bx * bg.by * bg.bz);
    h->partials.alloc(std::max<size_t>(h->partials.n, 2 * (size_t)nb + 64));
    hipLaunchKernelGGL(k_be_value, dim3(nb), dim3(256), 0, stream ? stream : h->stream, h->cppPos.p, h->cpp.nx,
                       h->cpp.ny, h->cpp.nz, bg, B, h->partials.p, h->maxBits.p + 6, 1.0 / (double)(3 * h->G),
                       out_device);
    NRH_HIP(hipGetLastError());
}

void fork_side(nrh_f3d *h) {
    if (!h->side) {
        NRH_HIP(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
        NRH_HIP(hipEventCreateWithFlags(&h->evFork, hipEventDisableTiming));
        NRH_HIP(hipEventCreateWithFlags(&h->evJoin, hipEventDisableTiming));
    }
    NRH_HIP(hipEventRecord(h->evFork, h->stream));
    NRH_HIP(hipStreamWaitEvent(h->side, h->evFork, 0));
}

void join_side(nrh_f3d *h) {
    NRH_HIP(hipEventRecord(h->evJoin, h->side));
    NRH_HIP(hipStreamWaitEvent(h->stream, h->evJoin, 0));
}

void op_bending_energy_gradient_raw(nrh_f3d *h, hipStream_t stream) {
    static const BeBasis B = make_be_basis();
    const BrickGrid bg = brick_grid(h);
    h->beGrad.alloc(3 * h->G);
    hipLaunchKernelGGL(k_be_gradient_raw, dim3((unsigned)(bg.bx * bg.by * bg.bz)), dim3(256), 0, stream,
                       h->cppPos.p, h->cpp.nx, h->cpp.ny, h->cpp.nz, bg, to_affine(h->cpp.xyz), B, h->beGrad.p);
    NRH_HIP(hipGetLastError());
}

void op_bending_energy_gradient(nrh_f3d *h, float weight, bool trackMax) {
    static const BeBasis B = make_be_basis();
    const BrickGrid bg = brick_grid(h);
    uint32_t *mb = trackMax ? h->maxBits.p : nullptr;
    if (trackMax) op_fill_u32(h->maxBits.p, 1, 0u, h->stream);
    const float ratio = weight / (float)h->G;
    hipLaunchKernelGGL(k_be_gradient, dim3((unsigned)(bg.bx * bg.by * bg.bz)), dim3(256), 0, h->stream, h->cppPos.p,
                       h->cpp.nx, h->cpp.ny, h->cpp.nz, bg, to_affine(h->cpp.xyz), B, ratio, h->nodeGrad.p, mb);
    hipLaunchKernelGGL(k_round_trip, dim3(blocks_for(h->G)), dim3(256), 0, h->stream, h->cppPos.p, h->cpp.nx,
                       h->cpp.ny, h->cpp.nz, to_affine(h->cpp.xyz));
    NRH_HIP(hipGetLastError());
    h->cppVersion++;  // the round trip may have changed the last bit of some points
}

namespace {

float sig6(double v) {
    if (v == 0) return 0.f;
    const double a = std::fabs(v);
    const int e = (int)std::floor(std::log10(a));
    const double scale = std::pow(10.0, 5 - e);
    const double r = std::floor(a * scale + 0.5) / scale;
    return (float)(v < 0 ? -r : r);
}

// set_first_order_basis_values 3-D (reg-lib/cpu/_reg_splineBasis.cpp:182-290): the reference writes
// the tensor products of the B-spline first derivative {-1/2, 0, 1/2} and value {1/6, 2/3, 1/6} as
// 6-significant-digit literals; they are regenerated here.
LeBasis make_le_basis() {
    const double v[3] = { 1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0 }, d1[3] = { -0.5, 0.0, 0.5 };
    LeBasis B;
    int i = 0;
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a, ++i) {
                B.b[0][i] = sig6(d1[a] * v[b] * v[c]);
                B.b[1][i] = sig6(v[a] * d1[b] * v[c]);
                B.b[2][i] = sig6(v[a] * v[b] * d1[c]);
            }
    return B;
}

M33 grid_reorientation(const nrh_f3d *h) {
    M33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = h->cpp.ijk[i * 4 + j];
    return r;
}

// nifti_mat33_inverse on the host, with the long double reciprocal of niftilib.
M33 host_m33_inverse(const M33& R) {
    const double r11 = R.m[0][0], r12 = R.m[0][1], r13 = R.m[0][2];
    const double r21 = R.m[1][0], r22 = R.m[1][1], r23 = R.m[1][2];
    const double r31 = R.m[2][0], r32 = R.m[2][1], r33 = R.m[2][2];
    double deti = r11 * r22 * r33 - r11 * r32 * r23 - r21 * r12 * r33 + r21 * r32 * r13 + r31 * r12 * r23 -
                  r31 * r22 * r13;
    if (deti != 0.0L) deti = (double)(1.0L / (long double)deti);
    M33 Q;
    Q.m[0][0] = (float)(deti * (r22 * r33 - r32 * r23));
    Q.m[0][1] = (float)(deti * (-r12 * r33 + r32 * r13));
    Q.m[0][2] = (float)(deti * (r12 * r23 - r22 * r13));
    Q.m[1][0] = (float)(deti * (-r21 * r33 + r31 * r23));
    Q.m[1][1] = (float)(deti * (r11 * r33 - r31 * r13));
    Q.m[1][2] = (float)(deti * (-r11 * r23 + r21 * r13));
    Q.m[2][0] = (float)(deti * (r21 * r32 - r31 * r22));
    Q.m[2][1] = (float)(deti * (-r11 * r32 + r31 * r12));
    Q.m[2][2] = (float)(deti * (r11 * r22 - r21 * r12));
    return Q;
}

}  // namespace

void op_linear_energy(nrh_f3d *h, double *out_device, bool keepDiag) {
    static const LeBasis B = make_le_basis();
    const unsigned nb = blocks_for(h->G);
    h->leDiag.alloc(3 * h->G);
    h->partials.alloc(std::max<size_t>(h->partials.n, 2 * (size_t)nb + 64));
    hipLaunchKernelGGL(k_le_terms, dim3(nb), dim3(kBlock), 0, h->stream, h->cppPos.p, h->cpp.nx, h->cpp.ny, h->cpp.nz,
                       B, grid_reorientation(h), h->leDiag.p, h->partials.p + nb + 32, h->maxBits.p + 7,
                       1.0 / (double)(3 * h->G), out_device);
    NRH_HIP(hipGetLastError());
    h->leDiagVersion = keepDiag ? h->cppVersion : 0;
}

void op_linear_energy_gradient(nrh_f3d *h, float weight, bool trackMax) {
    static const LeBasis B = make_le_basis();
    if (h->leDiagVersion != h->cppVersion || h->leDiag.empty())
        op_linear_energy(h, h->scalars.p + 7, true);  // the per-node diagonal of the current points
    if (trackMax) op_fill_u32(h->maxBits.p, 1, 0u, h->stream);
    const float ratio = weight / (float)h->G;
    hipLaunchKernelGGL(k_le_grad, dim3(blocks_for(h->G)), dim3(kBlock), 0, h->stream, h->leDiag.p, h->cpp.nx,
                       h->cpp.ny, h->cpp.nz, B, host_m33_inverse(grid_reorientation(h)), ratio, h->nodeGrad.p,
                       trackMax ? h->maxBits.p : nullptr);
    NRH_HIP(hipGetLastError());
}

void op_max_length(nrh_f3d *h) {
    op_fill_u32(h->maxBits.p, 1, 0u, h->stream);
    const unsigned b = std::min(blocks_for(h->G), 1024u);
    hipLaunchKernelGGL(k_max_length, dim3(b), dim3(kBlock), 0, h->stream, h->nodeGrad.p, h->G, h->maxBits.p);
    hipLaunchKernelGGL(k_max_to_scalar, dim3(1), dim3(64), 0, h->stream, h->maxBits.p, h->scalars.p + 1);
    NRH_HIP(hipGetLastError());
}

void op_normalise(nrh_f3d *h, double max_length) {
    if (max_length == 0) return;
    hipLaunchKernelGGL(k_normalise, dim3(blocks_for(3 * h->G)), dim3(kBlock), 0, h->stream, h->nodeGrad.p, 3 * h->G,
                       1.0 / max_length);
    NRH_HIP(hipGetLastError());
}

void op_normalise_from_bits(nrh_f3d *h) {
    hipLaunchKernelGGL(k_normalise_bits, dim3(blocks_for(3 * h->G)), dim3(kBlock), 0, h->stream, h->nodeGrad.p,
                       3 * h->G, (const uint32_t *)h->maxBits.p, h->scalars.p + 1);
    NRH_HIP(hipGetLastError());
}

void op_max_length_bits(nrh_f3d *h) {
    op_fill_u32(h->maxBits.p, 1, 0u, h->stream);
    const unsigned b = std::min(blocks_for(h->G), 1024u);
    hipLaunchKernelGGL(k_max_length, dim3(b), dim3(kBlock), 0, h->stream, h->nodeGrad.p, h->G, h->maxBits.p);
    NRH_HIP(hipGetLastError());
}

void op_normalise_device(nrh_f3d *h) {
    hipLaunchKernelGGL(k_normalise_dev, dim3(blocks_for(3 * h->G)), dim3(kBlock), 0, h->stream, h->nodeGrad.p,
                       3 * h->G, (const double *)(h->scalars.p + 1));
    NRH_HIP(hipGetLastError());
}

void op_rescale(nrh_f3d *h, float *data, size_t n, float lo, float hi) {
    const uint32_t init[2] = { 0xffffffffu, 0u };
    NRH_HIP(hipMemcpyAsync(h->maxBits.p + 2, init, sizeof init, hipMemcpyHostToDevice, h->stream));
    const unsigned b = std::min(blocks_for(n), 2048u);
    hipLaunchKernelGGL(k_minmax, dim3(b), dim3(kBlock), 0, h->stream, data, n, h->maxBits.p + 2);
    hipLaunchKernelGGL(k_rescale, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, data, n, h->maxBits.p + 2, lo, hi);
    NRH_HIP(hipGetLastError());
}

void op_cg(nrh_f3d *h, int first) {
    const size_t n = 3 * h->G;
    if (first) {
        hipLaunchKernelGGL(k_cg_init, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, h->nodeGrad.p, h->cgG.p, h->cgH.p, n);
    } else {
        const unsigned nb = std::min(blocks_for(n), 512u);
        h->partials.alloc(std::max<size_t>(h->partials.n, 2 * (size_t)nb + 64));
        hipLaunchKernelGGL(k_cg_dots, dim3(nb), dim3(kBlock), 0, h->stream, h->nodeGrad.p, h->cgG.p, h->cgH.p, n, h->partials.p);
        hipLaunchKernelGGL(k_cg_gamma, dim3(1), dim3(1024), 0, h->stream, h->partials.p, (int)nb, h->scalars.p + 2);
        hipLaunchKernelGGL(k_cg_apply, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, h->nodeGrad.p, h->cgG.p, h->cgH.p,
                           n, (const double *)(h->scalars.p + 2));
    }
    NRH_HIP(hipGetLastError());
}

// The symmetric conjugate-gradient update (reg-lib/Optimiser.cpp:231-310): both directions'
// dot products summed before one gamma updates both; launched on fw's stream after bw's work.
__global__ void __launch_bounds__(1024)
k_cg_gamma_pair(const double *__restrict__ pf, int nf, const double *__restrict__ pb, int nb, double *__restrict__ gf,
                double *__restrict__ gb) {
    __shared__ double red[16];
    double g[2][2];
    const double *parts[2] = { pf, pb };
    const int counts[2] = { nf, nb };
    for (int d = 0; d < 2; ++d) {
        double gg = 0, dgg = 0;
        for (int i = threadIdx.x; i < counts[d]; i += blockDim.x) {
            gg += parts[d][2 * i];
            dgg += parts[d][2 * i + 1];
        }
        gg = block_sum_double(gg, red);
        __syncthreads();
        dgg = block_sum_double(dgg, red);
        __syncthreads();
        g[d][0] = gg;
        g[d][1] = dgg;
    }
    if (threadIdx.x == 0) *gf = *gb = (g[0][1] + g[1][1]) / (g[0][0] + g[1][0]);
}

void op_cg_pair(nrh_f3d *fw, nrh_f3d *bw, int first) {
    // bw's arrays are updated on fw's stream: wait for bw's queued work first
    NRH_HIP(hipStreamSynchronize(bw->stream));
    hipStream_t s = fw->stream;
    const size_t nf = 3 * fw->G, nb = 3 * bw->G;
    if (first) {
        hipLaunchKernelGGL(k_cg_init, dim3(blocks_for(nf)), dim3(kBlock), 0, s, fw->nodeGrad.p, fw->cgG.p, fw->cgH.p, nf);
        hipLaunchKernelGGL(k_cg_init, dim3(blocks_for(nb)), dim3(kBlock), 0, s, bw->nodeGrad.p, bw->cgG.p, bw->cgH.p, nb);
    } else {
        const unsigned bf = std::min(blocks_for(nf), 512u), bb = std::min(blocks_for(nb), 512u);
        fw->partials.alloc(std::max<size_t>(fw->partials.n, 2 * (size_t)bf + 64));
        bw->partials.alloc(std::max<size_t>(bw->partials.n, 2 * (size_t)bb + 64));
        hipLaunchKernelGGL(k_cg_dots, dim3(bf), dim3(kBlock), 0, s, fw->nodeGrad.p, fw->cgG.p, fw->cgH.p, nf,
                           fw->partials.p);
        hipLaunchKernelGGL(k_cg_dots, dim3(bb), dim3(kBlock), 0, s, bw->nodeGrad.p, bw->cgG.p, bw->cgH.p, nb,
                           bw->partials.p);
        hipLaunchKernelGGL(k_cg_gamma_pair, dim3(1), dim3(1024), 0, s, fw->partials.p, (int)bf, bw->partials.p, (int)bb,
                           fw->scalars.p + 2, bw->scalars.p + 2);
        hipLaunchKernelGGL(k_cg_apply, dim3(blocks_for(nf)), dim3(kBlock), 0, s, fw->nodeGrad.p, fw->cgG.p, fw->cgH.p,
                           nf, (const double *)(fw->scalars.p + 2));
        hipLaunchKernelGGL(k_cg_apply, dim3(blocks_for(nb)), dim3(kBlock), 0, s, bw->nodeGrad.p, bw->cgG.p, bw->cgH.p,
                           nb, (const double *)(bw->scalars.p + 2));
    }
    NRH_HIP(hipGetLastError());
    NRH_HIP(hipStreamSynchronize(s));
}

void op_update_dof(nrh_f3d *h, float scale) {
    const size_t n = 3 * h->G;
    hipLaunchKernelGGL(k_update_dof, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, h->cppPos.p, h->best.p,
                       h->nodeGrad.p, scale, n);
    NRH_HIP(hipGetLastError());
    h->cppVersion++;
}

namespace {
// The values, then (after a system-scope fence) the sequence number in slot 15, which the host
// polls: a reader that sees the number sees the values.
__global__ void k_publish(double *__restrict__ dst, const double *__restrict__ s0, int n0,
                          const double *__restrict__ s1, int n1, const double *__restrict__ s2, int n2,
                          uint64_t seq) {
    const int i = threadIdx.x;
    if (i < n0) dst[i] = s0[i];
    else if (i < n0 + n1) dst[i] = s1[i - n0];
    else if (i < n0 + n1 + n2) dst[i] = s2[i - n0 - n1];
    __threadfence_system();
    __syncthreads();
    if (i == 0) __hip_atomic_store((uint64_t *)(dst + 15), seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}
}  // namespace

void op_publish(nrh_f3d *h, const double *s0, int n0, const double *s1, int n1, const double *s2, int n2) {
    NRH_REQUIRE(n0 + n1 + n2 <= 15, NRH_ERR_ARG, "too many scalars to publish");
    hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, h->stream, h->publishedDev, s0, n0, s1, n1, s2, n2,
                       ++h->publishSeq);
    NRH_HIP(hipGetLastError());
}

void wait_published(nrh_f3d *h) {
    // poll the sequence number the publish kernel writes last (a driver synchronisation costs
    // tens of microseconds of wake-up); every 256 polls ask the stream for an error, and once it
    // reports completion the number must be there
    const volatile uint64_t *seq = (const volatile uint64_t *)(h->published + 15);
    for (unsigned spins = 1;; ++spins) {
        if (__atomic_load_n((const uint64_t *)seq, __ATOMIC_ACQUIRE) == h->publishSeq) return;
        if (spins % 256 == 0) {
            const hipError_t e = hipStreamQuery(h->stream);
            if (e == hipSuccess) {
                NRH_REQUIRE(__atomic_load_n((const uint64_t *)seq, __ATOMIC_ACQUIRE) == h->publishSeq, NRH_ERR_HIP,
                            "the stream completed without publishing its results");
                return;
            }
            if (e != hipErrorNotReady) NRH_HIP(e);
        }
    }
}

double read_scalar(nrh_f3d *h, const double *dev) {
    op_publish(h, dev, 1);
    wait_published(h);
    return h->published[0];
}

// ------------------------------------------------------------------ batches of pairs
namespace {

// The objective terms of each active pair into results[8 q ..]: NMI, H_r, H_f, H_rf, total, the
// bending energy, the weighted value (reg_f3d::GetObjectiveFunctionValue with the linear energy
// off), and the current max gradient length; mirrored into coherent host memory.
__global__ void k_batch_values(const PairOps *__restrict__ P, int n, const int *active, const float *__restrict__ w,
                               double *__restrict__ res, double *__restrict__ host) {
#pragma clang fp contract(off)
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n || skip_pair(active, q)) return;
    const PairOps& o = P[q];
    const double simW = (double)w[2 * q], beW = (double)w[2 * q + 1];
    double r[kBatchResults];
    r[0] = o.ent[4];
    for (int k = 0; k < 4; ++k) r[1 + k] = o.ent[k];
    r[5] = beW > 0 ? *o.beOut : 0.0;
    const double wMeasure = simW > 0 ? simW * r[0] : 0.0;
    const double wBE = beW > 0 ? beW * r[5] : 0.0;
    r[6] = wMeasure - wBE;
    r[7] = *o.maxLen;
    for (int k = 0; k < kBatchResults; ++k) {
        res[q * kBatchResults + k] = r[k];
        host[q * kBatchResults + k] = r[k];
    }
}

}  // namespace

void batch_ops_setup(nrh_batch *b) {
    std::vector<PairOps> ops(b->P);
    std::vector<int2> itBE, itG, itG3, itCg;
    std::vector<float> w(2 * b->P);
    for (int q = 0; q < b->P; ++q) {
        nrh_f3d *h = b->pairs[q];
        NRH_REQUIRE(h->params.measure == NRH_MEASURE_NMI && h->params.le_weight == 0.f && h->mask.empty(),
                    NRH_ERR_UNSUPPORTED, "batched pairs use NMI, no linear energy and no mask");
        NRH_REQUIRE(h->totalBins() <= kFinLds, NRH_ERR_UNSUPPORTED, "too many histogram bins for a batch");
        const BrickGrid bg = brick_grid(h);
        const int nb = bg.bx * bg.by * bg.bz;
        const size_t n3 = 3 * h->G;
        const int cgBlocks = (int)std::min(blocks_for(n3), 512u);
        h->partials.alloc(std::max<size_t>(h->partials.n, 2 * (size_t)std::max(nb, cgBlocks) + 64));
        PairOps o{};
        o.cpp = h->cppPos.p;
        o.best = h->best.p;
        o.grad = h->nodeGrad.p;
        o.cgG = h->cgG.p;
        o.cgH = h->cgH.p;
        o.gnx = h->cpp.nx;
        o.gny = h->cpp.ny;
        o.gnz = h->cpp.nz;
        o.G = h->G;
        o.xyz = to_affine(h->cpp.xyz);
        o.bg = bg;
        o.beBlocks = nb;
        o.cgBlocks = cgBlocks;
        o.beRatio = h->params.be_weight / (float)h->G;
        o.beScale = 1.0 / (double)(3 * h->G);
        o.partials = h->partials.p;
        o.beTicket = h->maxBits.p + 6;
        o.beOut = h->scalars.p + 0;
        o.maxBits = h->maxBits.p;
        o.maxLen = h->scalars.p + 1;
        o.gam = h->scalars.p + 2;
        o.counts = h->histCount.p;
        o.pro = h->histPro.p;
        o.logh = h->histLog.p;
        o.logf = h->histLogF.p;
        o.ent = h->ent.p;
        o.rb = h->params.ref_bins;
        o.fb = h->params.flo_bins;
        ops[q] = o;
        w[2 * q] = h->params.similarity_weight;
        w[2 * q + 1] = h->params.be_weight;
        for (int k = 0; k < nb; ++k) itBE.push_back(make_int2(q, k));
        for (unsigned k = 0; k < blocks_for(h->G); ++k) itG.push_back(make_int2(q, (int)k));
        for (unsigned k = 0; k < blocks_for(n3); ++k) itG3.push_back(make_int2(q, (int)k));
        for (int k = 0; k < cgBlocks; ++k) itCg.push_back(make_int2(q, k));
    }
    b->pairOps.alloc(ops.size() * sizeof(PairOps));
    NRH_HIP(hipMemcpyAsync(b->pairOps.p, ops.data(), ops.size() * sizeof(PairOps), hipMemcpyHostToDevice, b->stream));
    b->weights.alloc(w.size());
    NRH_HIP(hipMemcpyAsync(b->weights.p, w.data(), w.size() * sizeof(float), hipMemcpyHostToDevice, b->stream));
    auto up = [&](DevBuf<int2>& dst, const std::vector<int2>& v, int& n) {
        dst.alloc(v.size());
        NRH_HIP(hipMemcpyAsync(dst.p, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream));
        n = (int)v.size();
    };
    up(b->itBE, itBE, b->nBE);
    up(b->itG, itG, b->nG);
    up(b->itG3, itG3, b->nG3);
    up(b->itCg, itCg, b->nCg);
    b->results.alloc((size_t)kBatchResults * b->P);
    NRH_HIP(hipMemsetAsync(b->results.p, 0, b->results.bytes(), b->stream));
    NRH_HIP(hipStreamSynchronize(b->stream));
}

void batch_finalize(nrh_batch *b, const int *active) {
    hipLaunchKernelGGL(k_nmi_finalize_batch, dim3(b->P), dim3(1024), 0, b->stream, (const PairOps *)b->pairOps.p,
                       active);
    NRH_HIP(hipGetLastError());
}

void batch_be_value(nrh_batch *b, const int *active) {
    static const BeBasis B = make_be_basis();
    hipLaunchKernelGGL(k_be_value_batch, dim3(b->nBE), dim3(256), 0, b->stream, (const PairOps *)b->pairOps.p,
                       b->itBE.p, B, active);
    NRH_HIP(hipGetLastError());
}

void batch_be_gradient(nrh_batch *b, const int *active) {
    static const BeBasis B = make_be_basis();
    const PairOps *P = (const PairOps *)b->pairOps.p;
    hipLaunchKernelGGL(k_zero_max_bits_batch, dim3((b->P + 255) / 256), dim3(256), 0, b->stream, P, b->P, active);
    hipLaunchKernelGGL(k_be_gradient_batch, dim3(b->nBE), dim3(256), 0, b->stream, P, b->itBE.p, B, active);
    hipLaunchKernelGGL(k_round_trip_batch, dim3(b->nG), dim3(256), 0, b->stream, P, b->itG.p, active);
    NRH_HIP(hipGetLastError());
}

void batch_normalise(nrh_batch *b, const int *active) {
    hipLaunchKernelGGL(k_normalise_batch, dim3(b->nG3), dim3(kBlock), 0, b->stream, (const PairOps *)b->pairOps.p,
                       b->itG3.p, active);
    NRH_HIP(hipGetLastError());
}

void batch_cg(nrh_batch *b, const int *active, const int *first) {
    const PairOps *P = (const PairOps *)b->pairOps.p;
    hipLaunchKernelGGL(k_cg_init_batch, dim3(b->nG3), dim3(kBlock), 0, b->stream, P, b->itG3.p, active, first);
    hipLaunchKernelGGL(k_cg_dots_batch, dim3(b->nCg), dim3(kBlock), 0, b->stream, P, b->itCg.p, active, first);
    hipLaunchKernelGGL(k_cg_gamma_batch, dim3(b->P), dim3(1024), 0, b->stream, P, active, first);
    hipLaunchKernelGGL(k_cg_apply_batch, dim3(b->nG3), dim3(kBlock), 0, b->stream, P, b->itG3.p, active, first);
    NRH_HIP(hipGetLastError());
}

void batch_update(nrh_batch *b, const int *active, const float *scales) {
    hipLaunchKernelGGL(k_update_dof_batch, dim3(b->nG3), dim3(kBlock), 0, b->stream, (const PairOps *)b->pairOps.p,
                       b->itG3.p, active, scales);
    NRH_HIP(hipGetLastError());
}

void batch_copy_dof(nrh_batch *b, const int *active, bool toBest) {
    hipLaunchKernelGGL(k_copy_dof_batch, dim3(b->nG3), dim3(kBlock), 0, b->stream, (const PairOps *)b->pairOps.p,
                       b->itG3.p, active, toBest ? 1 : 0);
    NRH_HIP(hipGetLastError());
}

void batch_values(nrh_batch *b, const int *active) {
    hipLaunchKernelGGL(k_batch_values, dim3((b->P + 63) / 64), dim3(64), 0, b->stream, (const PairOps *)b->pairOps.p,
                       b->P, active, b->weights.p, b->results.p, b->publishedDev);
    NRH_HIP(hipGetLastError());
}

}  // namespace nrh
