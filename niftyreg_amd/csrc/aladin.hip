// Aladin (block-matching affine registration) on gfx950.
//
// One handle holds one pyramid level of reg_aladin (the HIP counterpart of AladinContent,
// reg-lib/AladinContent.cpp:6-29).  The per-iteration loop of reg_aladin::ResolveMatrix
// (reg-lib/_reg_aladin.cpp:302-313) is three launches on the handle's stream with the current
// matrix kept in device memory:
//   k_affine_warp   field = (A * ref.sto_xyz) * voxel in FP64, trilinear warp of the floating image
//   k_block_match   one workgroup per active 4x4x4 block, one lane per capture offset (343)
//   k_lts           one 1024-lane workgroup: compaction, FP64 moment fits, radix-select trimming
// Block variances for the active-block choice are computed on the device; their heap-sort
// ranking runs on the host because its tie order is part of the reference's result.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <memory>
#include <vector>

#include "common.hpp"
#include <array>
#include <chrono>
#include <thread>
#include <map>
#include <mutex>
#include "device.hpp"
#include "recip_x87.hpp"
#include "m44.hpp"
#include "host.hpp"

struct nrh_aladin {
    nrh_geom ref{}, flo{};
    nrh_aladin_params params{};
    hipStream_t stream = nullptr;
    bool ownsStream = true;           // false when the stream is another handle's (symmetric pair)
    int device = 0;
    size_t N = 0, Nf = 0;
    nrh::DevBuf<float> refImg, floImg, warped;
    nrh::DevBuf<int> mask;  // empty: every voxel active
    int nb[3] = { 0, 0, 0 };
    int totalBlocks = 0, activeBlocks = 0;
    std::vector<int> totalBlock;      // _reg_blockMatchingParam::totalBlock
    nrh::DevBuf<int> activeList;      // rank -> block index
    nrh::DevBuf<float> refPos, warpPos;
    nrh::DevBuf<float> matrix;        // current affine, 16 floats row-major
    nrh::DevBuf<int> counters;        // [0] defined blocks of the last matching, [1] LTS status,
                                      // [2] symmetric-update status
    nrh::DevBuf<float> ltsP, ltsQ;    // compacted correspondences
    nrh::DevBuf<uint32_t> ltsRes;     // residual bit patterns
    int *pinned = nullptr;            // 4 ints of host-visible status
};

namespace nrh {
namespace {

constexpr int kBlock = 256;
constexpr int kMatchThreads = 384;  // 343 capture offsets rounded up to whole waves
constexpr int kLtsThreads = 1024;
constexpr int kLtsUnroll = 4;  // points per thread whose loads the LTS passes issue together

struct Mat4 {
    float m[16];
};

__device__ __forceinline__ float3 mat44_mul_f(const float *m, float x, float y, float z) {
#pragma clang fp contract(off)
    float3 r;
    r.x = (float)((double)m[0] * (double)x + (double)m[1] * (double)y + (double)m[2] * (double)z + (double)m[3]);
    r.y = (float)((double)m[4] * (double)x + (double)m[5] * (double)y + (double)m[6] * (double)z + (double)m[7]);
    r.z = (float)((double)m[8] * (double)x + (double)m[9] * (double)y + (double)m[10] * (double)z + (double)m[11]);
    return r;
}

// ------------------------------------------------------------------ active blocks
// Variance of each 4x4x4 block (reg-lib/cpu/_reg_blockMatching.cpp:70-128): the mean over finite
// voxels with mask > -1, the variance over every finite voxel, float sums in z, y, x order.
__global__ void __launch_bounds__(kBlock)
k_block_variance(const float *__restrict__ img, const int *__restrict__ mask, int nx, int ny, int nz, int nbx,
                 int nby, int nbz, float *__restrict__ var) {
#pragma clang fp contract(off)
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbx * nby * nbz) return;
    const int bi = b % nbx, bj = (b / nbx) % nby, bk = b / (nbx * nby);
    const int x1 = min(4 * bi + 4, nx), y1 = min(4 * bj + 4, ny), z1 = min(4 * bk + 4, nz);
    float mean = 0.f, count = 0.f;
    for (int z = 4 * bk; z < z1; ++z)
        for (int y = 4 * bj; y < y1; ++y)
            for (int x = 4 * bi; x < x1; ++x) {
                const size_t i = ((size_t)z * ny + y) * nx + x;
                const float v = img[i];
                if (v == v && (!mask || mask[i] > -1)) {
                    mean += v;
                    count += 1.f;
                }
            }
    mean /= count;
    float variance = 0.f;
    for (int z = 4 * bk; z < z1; ++z)
        for (int y = 4 * bj; y < y1; ++y)
            for (int x = 4 * bi; x < x1; ++x) {
                const float v = img[((size_t)z * ny + y) * nx + x];
                if (v == v) variance += (mean - v) * (mean - v);
            }
    variance /= count;
    var[b] = (count > 32.f && variance > 0.f) ? variance : -1.f;
}

// Heap sort of values with a companion index, ascending: the Numerical-Recipes sift-down
// sequence of HeapSort(float*, int*, int) (reg-lib/cpu/Maths.cpp:114-153), whose order of equal
// values decides which of several equal-variance blocks are kept.
void heap_sort_indexed(std::vector<float>& v, std::vector<int>& ix) {
    const int n = (int)v.size();
    if (n < 2) return;
    float *a = v.data() - 1;
    int *id = ix.data() - 1;
    int l = (n >> 1) + 1, ir = n;
    for (;;) {
        float val;
        int iv;
        if (l > 1) {
            --l;
            val = a[l];
            iv = id[l];
        } else {
            val = a[ir];
            iv = id[ir];
            a[ir] = a[1];
            id[ir] = id[1];
            if (--ir == 1) {
                a[1] = val;
                id[1] = iv;
                return;
            }
        }
        int i = l, j = l + l;
        while (j <= ir) {
            if (j < ir && a[j] < a[j + 1]) ++j;
            if (!(val < a[j])) break;
            a[i] = a[j];
            id[i] = id[j];
            i = j;
            j <<= 1;
        }
        a[i] = val;
        id[i] = iv;
    }
}

// ------------------------------------------------------------------ affine warp
// reg_affine_deformationField3D (reg-lib/cpu/_reg_globalTrans.cpp:74-127): T = A * ref.sto_xyz
// (mat44 product with double sums, float entries), position = T * voxel in double, stored as
// float; then the trilinear resampling of that position (reg-lib/cpu/_reg_resampling.cpp:316-464).
__global__ void __launch_bounds__(kBlock)
k_affine_warp(const float *__restrict__ flo, int fnx, int fny, int fnz, Affine floIjk,
              const float *__restrict__ A, Mat4 refXyz, int nx, int ny, int nz, const int *__restrict__ mask,
              float pad, float *__restrict__ warped) {
#pragma clang fp contract(off)
    __shared__ double T[12];
    if (threadIdx.x < 12) {
        const int r = threadIdx.x >> 2, c = threadIdx.x & 3;
        const float t = (float)((double)A[r * 4 + 0] * (double)refXyz.m[c] +
                                (double)A[r * 4 + 1] * (double)refXyz.m[4 + c] +
                                (double)A[r * 4 + 2] * (double)refXyz.m[8 + c] +
                                (double)A[r * 4 + 3] * (double)refXyz.m[12 + c]);
        T[threadIdx.x] = (double)t;
    }
    __syncthreads();
    const size_t N = (size_t)nx * ny * nz;
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= N) return;
    if (mask && mask[idx] < 0) return;
    const double x = (double)(idx % nx), y = (double)((idx / nx) % ny), z = (double)(idx / ((size_t)nx * ny));
    const float fx = (float)(T[0] * x + T[1] * y + T[2] * z + T[3]);
    const float fy = (float)(T[4] * x + T[5] * y + T[6] * z + T[7]);
    const float fz = (float)(T[8] * x + T[9] * y + T[10] * z + T[11]);
    const float3 p = affine_apply_d(floIjk, fx, fy, fz);
    warped[idx] = sample_linear(flo, fnx, fny, fnz, p, pad);
}

// ------------------------------------------------------------------ block matching
// block_matching_method3D<float> (reg-lib/cpu/_reg_blockMatching.cpp:455-690).  Workgroup =
// one active block (grid = active block count, in rank order); lane t < 343 = capture offset
// (l, m, n) = (t % 7, t / 7 % 7, t / 49) - 3, the reference's loop order.  Each lane keeps its 64
// warped values in registers and runs the reference's two float passes (means, then variances
// and covariance) in voxel order; wave 0 then replays the sequential "cc > best -> best =
// cc + 1e-7" rule with ballots, which gives the same winner as the reference's scan.
__global__ void __launch_bounds__(kMatchThreads)
k_block_match(const float *__restrict__ ref, const float *__restrict__ warped, const int *__restrict__ mask,
              int nx, int ny, int nz, int nbx, int nby, const int *__restrict__ activeList, Affine xyz,
              float *__restrict__ refPos, float *__restrict__ warpPos, int *__restrict__ definedCount) {
#pragma clang fp contract(off)
    __shared__ float sR[64];
    __shared__ float sW[1000];
    __shared__ float sCC[kMatchThreads];
    const int rank = blockIdx.x;
    const int b = activeList[rank];
    const int bi = b % nbx, bj = (b / nbx) % nby, bk = b / (nbx * nby);
    const int x0 = 4 * bi, y0 = 4 * bj, z0 = 4 * bk;
    const int t = threadIdx.x;
    const float qnan = __builtin_nanf("");
    if (t < 64) {
        const int x = x0 + (t & 3), y = y0 + ((t >> 2) & 3), z = z0 + (t >> 4);
        float v = qnan;
        if (x < nx && y < ny && z < nz) {
            const size_t i = ((size_t)z * ny + y) * nx + x;
            const float r = ref[i];
            if (r == r && (!mask || mask[i] > -1)) v = r;
        }
        sR[t] = v;
    }
    for (int q = t; q < 1000; q += kMatchThreads) {
        const int x = x0 - 3 + q % 10, y = y0 - 3 + (q / 10) % 10, z = z0 - 3 + q / 100;
        float v = qnan;
        if (x > -1 && x < nx && y > -1 && y < ny && z > -1 && z < nz) v = warped[((size_t)z * ny + y) * nx + x];
        sW[q] = v;
    }
    __syncthreads();
    float cc = -1.f;
    if (t < 343) {
        const int l = t % 7, m = (t / 7) % 7, n = t / 49;
        // the lane's 64 warped values are read from LDS in each pass (constant offsets from one
        // base) rather than held in registers: the kernel then fits several workgroups per CU
        const float *wb = sW + n * 100 + m * 10 + l;
        auto W = [&](int a) { return wb[(a >> 4) * 100 + ((a >> 2) & 3) * 10 + (a & 3)]; };
        float rm = 0.f, wm = 0.f, cnt = 0.f;
#pragma unroll
        for (int a = 0; a < 64; ++a) {
            const float r = sR[a];
            if (r == r && W(a) == W(a)) {
                rm += r;
                wm += W(a);
                cnt += 1.f;
            }
        }
        if (cnt > 32.f) {
            asm volatile("" ::: "memory");  // the second pass reads LDS again
            rm /= cnt;
            wm /= cnt;
            float rv = 0.f, wv = 0.f, c = 0.f;
#pragma unroll
            for (int a = 0; a < 64; ++a) {
                const float r = sR[a];
                if (r == r && W(a) == W(a)) {
                    const float rt = r - rm, wt = W(a) - wm;
                    rv += rt * rt;
                    wv += wt * wt;
                    c += rt * wt;
                }
            }
            cc = (rv * wv) > 0.f ? fabsf(c / sqrtf(rv * wv)) : 0.f;
        }
    }
    sCC[t] = cc;
    __syncthreads();
    if (t >= 64) return;
    float best = 0.f;
    int win = -1;
    for (int c = 0; c < kMatchThreads / 64; ++c) {
        const int o = c * 64 + t;
        const float v = o < 343 ? sCC[o] : -1.f;
        unsigned long long mk = __ballot(v > best);
        while (mk) {
            const int first = __ffsll((long long)mk) - 1;
            best = __shfl(v, first, 64) + 1.0e-7f;
            win = c * 64 + first;
            mk = __ballot(v > best && t > first);
        }
    }
    if (t != 0) return;
    const float cx = (float)x0, cy = (float)y0, cz = (float)z0;
    float dx = qnan, dy = 0.f, dz = 0.f;
    if (win >= 0) {
        dx = (float)(win % 7 - 3);
        dy = (float)((win / 7) % 7 - 3);
        dz = (float)(win / 49 - 3);
    }
    dx += cx;
    dy += cy;
    dz += cz;
    const float3 rp = affine_apply_d(xyz, cx, cy, cz);
    const float3 wp = affine_apply_d(xyz, dx, dy, dz);
    refPos[3 * rank + 0] = rp.x;
    refPos[3 * rank + 1] = rp.y;
    refPos[3 * rank + 2] = rp.z;
    warpPos[3 * rank + 0] = wp.x;
    warpPos[3 * rank + 1] = wp.y;
    warpPos[3 * rank + 2] = wp.z;
    if (win >= 0) atomicAdd(definedCount, 1);
}

// ------------------------------------------------------------------ LTS
struct alignas(16) LtsShared {
    double red[kLtsThreads / 64][24];
    double out[24];
    double fit[21];  // the fit's centred moments and means
    uint32_t hist[2048];
    int wave[kLtsThreads / 64];
    float M[16], last[16];
    int scal[4];
    double dist;
};

// Sum of K doubles over the workgroup in a fixed order; the result lands in s.out[0..K).
// Inside a wave the values are halved at each exchange: the lanes whose exchange bit is clear keep
// the lower half of the (zero-padded, power-of-two) set and send the upper half, the others the
// reverse, so the wave makes W/2 + W/4 + ... + 1 exchanges instead of K per lane bit; after log2 W
// of them lane l holds value (l >> (6 - log2 W)) summed over its group, and the remaining lane bits
// are summed by plain exchanges.
template<int K>
__device__ void wg_sum(double (&v)[K], LtsShared& s) {
    static_assert(K <= 32, "at most 32 values");
    constexpr int W = K <= 8 ? 8 : K <= 16 ? 16 : 32;
    constexpr int LW = W == 8 ? 3 : W == 16 ? 4 : 5;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double a[W];
#pragma unroll
    for (int k = 0; k < W; ++k) a[k] = k < K ? v[k] : 0.0;
#pragma unroll
    for (int st = 0; st < LW; ++st) {
        const int off = 32 >> st, half = W >> (st + 1);
        const bool upper = (lane & off) != 0;
#pragma unroll
        for (int j = 0; j < half; ++j) {
            const double send = upper ? a[j] : a[j + half];
            const double keep = upper ? a[j + half] : a[j];
            a[j] = keep + __shfl_xor(send, off, 64);
        }
    }
#pragma unroll
    for (int off = 32 >> LW; off > 0; off >>= 1) a[0] += __shfl_xor(a[0], off, 64);
    const int idx = (lane >> (6 - LW)) & (W - 1);
    if ((lane & ((1 << (6 - LW)) - 1)) == 0 && idx < K) s.red[wv][idx] = a[0];
    __syncthreads();
    if (threadIdx.x < K) {
        double r = 0;
        for (int w = 0; w < kLtsThreads / 64; ++w) r += s.red[w][threadIdx.x];
        s.out[threadIdx.x] = r;
    }
    __syncthreads();
}

// Requires every value of a batch of loads in registers at this point, so that the batch's loads
// are all issued before the first is used (the compiler otherwise sinks each load into the
// conditional code that uses it, and the loads wait one after another).
template<int U>
__device__ __forceinline__ void batch_loaded(const float (&v)[U][6]) {
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
        for (int c = 0; c < 6; ++c) asm volatile("" : : "v"(v[u][c]));
}

// Exclusive prefix sum over the 64 lanes of a wave (all lanes active), by data-parallel
// primitives (DPP): shifts right by 1, 2, 4 and 8 lanes inside each row of 16 (lanes with no source
// add 0), then lane 15 of each row added to the next row and lane 31 to the upper half.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v) {
    int incl = (int)v;
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x111, 0xf, 0xf, false);  // row_shr:1
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x112, 0xf, 0xf, false);  // row_shr:2
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x114, 0xf, 0xf, false);  // row_shr:4
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x118, 0xf, 0xf, false);  // row_shr:8
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x142, 0xa, 0xf, false);  // row_bcast:15
    incl += __builtin_amdgcn_update_dpp(0, incl, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return (uint32_t)incl - v;
}

// Exclusive prefix of a 0/1 flag over the workgroup (lane order = index order); returns the
// position of this lane and writes the chunk total to *total (uniform).
__device__ int wg_scan(bool flag, LtsShared& s, int *total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const unsigned long long bal = __ballot(flag);
    const int inWave = __popcll(bal & ((1ull << lane) - 1ull));
    if (lane == 0) s.wave[wv] = __popcll(bal);
    __syncthreads();
    int before = 0, all = 0;
    for (int w = 0; w < kLtsThreads / 64; ++w) {
        if (w < wv) before += s.wave[w];
        all += s.wave[w];
    }
    __syncthreads();
    *total = all;
    return before + inWave;
}

// H = U S V^T of a 3x3 matrix by one-sided Jacobi rotations; singular values descending.
__device__ void svd3(const double H[3][3], double U[3][3], double V[3][3]) {
    double A[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            A[i][j] = H[i][j];
            V[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0;
        // unrolled: every index is a compile-time constant, so A and V stay in registers
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = p + 1; q < 3; ++q) {
                const double a = A[0][p] * A[0][p] + A[1][p] * A[1][p] + A[2][p] * A[2][p];
                const double b = A[0][q] * A[0][q] + A[1][q] * A[1][q] + A[2][q] * A[2][q];
                const double c = A[0][p] * A[0][q] + A[1][p] * A[1][q] + A[2][p] * A[2][q];
                if (fabs(c) <= 1e-300 || fabs(c) <= 1e-15 * sqrt(a * b)) continue;
                off = fmax(off, fabs(c) / sqrt(a * b));
                const double zeta = (b - a) / (2 * c);
                const double tn = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
                const double cs = 1 / sqrt(1 + tn * tn), sn = cs * tn;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const double x = A[i][p], y = A[i][q];
                    A[i][p] = cs * x - sn * y;
                    A[i][q] = sn * x + cs * y;
                    const double vx = V[i][p], vy = V[i][q];
                    V[i][p] = cs * vx - sn * vy;
                    V[i][q] = sn * vx + cs * vy;
                }
            }
        if (off < 1e-15) break;
    }
    double sv[3];
    int o[3] = { 0, 1, 2 };
    for (int j = 0; j < 3; ++j) sv[j] = sqrt(A[0][j] * A[0][j] + A[1][j] * A[1][j] + A[2][j] * A[2][j]);
    for (int i = 0; i < 3; ++i)  // sort column order by descending singular value
        for (int j = i + 1; j < 3; ++j)
            if (sv[o[j]] > sv[o[i]]) {
                const int tmp = o[i];
                o[i] = o[j];
                o[j] = tmp;
            }
    double V2[3][3];
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) {
            U[i][j] = sv[o[j]] > 0 ? A[i][o[j]] / sv[o[j]] : 0.0;
            V2[i][j] = V[i][o[j]];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = V2[i][j];
}

// Thread 0's half of the fit: the 3x3 solve from s.fit (the centred moments Spp 00 01 02 11 12 22,
// Spq row-major, then the means of p and q) into s.M and s.scal[0].  Out of line, so that its
// registers do not add to those of the passes over the points.
__device__ __noinline__ void lts_solve(LtsShared& s, bool affine) {
    const double *o = s.fit, *mp = s.fit + 15, *mq = s.fit + 18;
    // Spq[a][b] = sum (p_a - mp_a)(q_b - mq_b)
    const double Spq[3][3] = { { o[6], o[7], o[8] }, { o[9], o[10], o[11] }, { o[12], o[13], o[14] } };
    double L[3][3];
    int ok = 1;
    if (affine) {
        // Spp X = Spq by Gaussian elimination with partial pivoting; L = X^T.
        double a[3][6] = { { o[0], o[1], o[2], Spq[0][0], Spq[0][1], Spq[0][2] },
                           { o[1], o[3], o[4], Spq[1][0], Spq[1][1], Spq[1][2] },
                           { o[2], o[4], o[5], Spq[2][0], Spq[2][1], Spq[2][2] } };
        const double scale = fmax(fabs(o[0]), fmax(fabs(o[3]), fabs(o[5])));
        // fully unrolled, the pivot row swapped in by selects: a stays in registers
#pragma unroll
        for (int col = 0; col < 3; ++col) {
            int piv = col;
            double best = fabs(a[col][col]);
#pragma unroll
            for (int r = col + 1; r < 3; ++r)
                if (fabs(a[r][col]) > best) {
                    best = fabs(a[r][col]);
                    piv = r;
                }
            if (!(best > 1e-12 * scale)) ok = 0;
#pragma unroll
            for (int r = col + 1; r < 3; ++r)
                if (piv == r)
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        const double tmp = a[col][k];
                        a[col][k] = a[r][k];
                        a[r][k] = tmp;
                    }
#pragma unroll
            for (int r = 0; r < 3; ++r) {
                if (r == col) continue;
                const double f = a[r][col] / a[col][col];
#pragma unroll
                for (int k = col; k < 6; ++k) a[r][k] -= f * a[col][k];
            }
        }
        if (ok)
            for (int r = 0; r < 3; ++r)
                for (int k = 0; k < 3; ++k) L[k][r] = a[r][3 + k] / a[r][r];
    } else {
        double U[3][3], V[3][3];
        svd3(Spq, U, V);
        auto vut = [&] {
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) L[a][b] = V[a][0] * U[b][0] + V[a][1] * U[b][1] + V[a][2] * U[b][2];
        };
        vut();
        const double det = L[0][0] * (L[1][1] * L[2][2] - L[1][2] * L[2][1]) -
                           L[0][1] * (L[1][0] * L[2][2] - L[1][2] * L[2][0]) +
                           L[0][2] * (L[1][0] * L[2][1] - L[1][1] * L[2][0]);
        if (det < 0) {
            for (int i = 0; i < 3; ++i) V[i][2] = -V[i][2];
            vut();
        }
    }
    if (ok) {
        for (int r = 0; r < 3; ++r) {
            for (int k = 0; k < 3; ++k) s.M[r * 4 + k] = (float)L[r][k];
            s.M[r * 4 + 3] = (float)(mq[r] - (L[r][0] * mp[0] + L[r][1] * mp[1] + L[r][2] * mp[2]));
        }
        s.M[12] = s.M[13] = s.M[14] = 0.f;
        s.M[15] = 1.f;
    }
    s.scal[0] = ok;
}

// Least-squares fit of the kept correspondences (keep == nullptr: all n; fromSums:
// from the sums the caller's keep pass left in s.out) into s.M.
// Affine: the FP64 solution of [p 1] P = q through centred normal equations (the same minimiser
// as EstimateAffineLeastSquares' QR, reg-lib/cpu/Maths.cpp:325-349).  Rigid: Kabsch on the centred
// cross-covariance with the reflection guard (reg-lib/cpu/Maths.cpp:351-390).  Returns false
// (uniformly) when the normal matrix is singular.
__device__ bool lts_fit(const float *__restrict__ P, const float *__restrict__ Q, const uint8_t *keep, int n,
                        bool affine, LtsShared& s, bool fromSums = false) {
    if (fromSums) {
        // the kept points' sums from the caller's pass, in s.out[1..22]: p, q, count, then the
        // products p_a p_b (00 01 02 11 12 22) and p_a q_b; centred moments S = sum - count m m^T
        // in FP64
        if (threadIdx.x == 0) {
            const double *sums = s.out + 1;
            double *oc = s.fit, *mp = s.fit + 15, *mq = s.fit + 18;
            const double cnt = sums[6];
            for (int k = 0; k < 3; ++k) {
                mp[k] = sums[k] / cnt;
                mq[k] = sums[3 + k] / cnt;
            }
            const int pa[6] = { 0, 0, 0, 1, 1, 2 }, pb[6] = { 0, 1, 2, 1, 2, 2 };
            for (int k = 0; k < 6; ++k) oc[k] = sums[7 + k] - cnt * mp[pa[k]] * mp[pb[k]];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) oc[6 + a * 3 + b] = sums[13 + a * 3 + b] - cnt * mp[a] * mq[b];
        }
    } else {
        double mp[3], mq[3];
        double m[7] = { 0, 0, 0, 0, 0, 0, 0 };
        for (int i = threadIdx.x; i < n; i += kLtsThreads) {
            if (keep && !keep[i]) continue;
            for (int k = 0; k < 3; ++k) {
                m[k] += (double)P[3 * i + k];
                m[3 + k] += (double)Q[3 * i + k];
            }
            m[6] += 1.0;
        }
        wg_sum<7>(m, s);
        const double cnt = s.out[6];
        for (int k = 0; k < 3; ++k) {
            mp[k] = s.out[k] / cnt;
            mq[k] = s.out[3 + k] / cnt;
        }
        __syncthreads();
        double c[15] = { 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0 };
        for (int i = threadIdx.x; i < n; i += kLtsThreads) {
            if (keep && !keep[i]) continue;
            const double p0 = (double)P[3 * i] - mp[0], p1 = (double)P[3 * i + 1] - mp[1], p2 = (double)P[3 * i + 2] - mp[2];
            const double q0 = (double)Q[3 * i] - mq[0], q1 = (double)Q[3 * i + 1] - mq[1], q2 = (double)Q[3 * i + 2] - mq[2];
            c[0] += p0 * p0; c[1] += p0 * p1; c[2] += p0 * p2; c[3] += p1 * p1; c[4] += p1 * p2; c[5] += p2 * p2;
            c[6] += p0 * q0; c[7] += p0 * q1; c[8] += p0 * q2;
            c[9] += p1 * q0; c[10] += p1 * q1; c[11] += p1 * q2;
            c[12] += p2 * q0; c[13] += p2 * q1; c[14] += p2 * q2;
        }
        wg_sum<15>(c, s);
        if (threadIdx.x == 0) {
            for (int k = 0; k < 15; ++k) s.fit[k] = s.out[k];
            for (int k = 0; k < 3; ++k) {
                s.fit[15 + k] = mp[k];
                s.fit[18 + k] = mq[k];
            }
        }
    }
    if (threadIdx.x == 0) lts_solve(s, affine);
    __syncthreads();
    return s.scal[0] != 0;
}

// optimize + optimize_3D (reg-lib/cpu/_reg_blockMatching.cpp:722-821,
// reg-lib/cpu/_reg_globalTrans.cpp:328-394).  status: 0 ok, 1 too few correspondences,
// 2 singular fit (the matrix is then left unchanged).  The trimmed subset is the keep smallest
// residuals (float bit patterns, found by a three-digit radix select); ties at the threshold
// keep the lowest indices.
__device__ __forceinline__ void
lts_body(const float *__restrict__ refPos, const float *__restrict__ warpPos, int nActive, float *__restrict__ matrix,
         float *__restrict__ P, float *__restrict__ Q, uint32_t *__restrict__ res, int percent, int affine,
         int *__restrict__ status, LtsShared& s, uint8_t *keepIn = nullptr) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
    if (tid < 16) s.M[tid] = matrix[tid];
    __syncthreads();
    // defined correspondences in rank order, warped side mapped through the current matrix
    int n = 0;
    for (int base = 0; base < nActive; base += kLtsThreads) {
        const int i = base + tid;
        const bool flag = i < nActive && warpPos[3 * i] == warpPos[3 * i];
        int total;
        const int pos = n + wg_scan(flag, s, &total);
        if (flag) {
            const float3 q = mat44_mul_f(s.M, warpPos[3 * i], warpPos[3 * i + 1], warpPos[3 * i + 2]);
            P[3 * pos] = refPos[3 * i];
            P[3 * pos + 1] = refPos[3 * i + 1];
            P[3 * pos + 2] = refPos[3 * i + 2];
            Q[3 * pos] = q.x;
            Q[3 * pos + 1] = q.y;
            Q[3 * pos + 2] = q.z;
        }
        n += total;
    }
    const unsigned keep = (unsigned)((float)n * ((float)percent / 100.0f));
    if (n < (affine ? 8 : 4) || keep < 4) {
        if (tid == 0) *status = 1;
        return;
    }
    __threadfence_block();
    __syncthreads();
    if (!lts_fit(P, Q, nullptr, n, affine != 0, s)) {
        if (tid == 0) *status = 2;
        return;
    }
    // keep flags: after the residuals in global memory, or a separate array (LDS)
    uint8_t *keepFlag = keepIn ? keepIn : reinterpret_cast<uint8_t *>(res + n);
    double lastDistance = DBL_MAX;
    if (tid < 16) s.last[tid] = 0.f;
    __syncthreads();
    for (int it = 0; it < 30; ++it) {
        // residuals, counted into the first digit's histogram as they are formed
        for (int b = tid; b < 2048; b += kLtsThreads) s.hist[b] = 0;
        __syncthreads();
        // kLtsUnroll points per thread per step: their loads are issued together (indexes
        // clamped to n - 1, so every load stays in bounds), then the residuals formed
        for (int i0 = tid; i0 < n; i0 += kLtsUnroll * kLtsThreads) {
            float pq[kLtsUnroll][6];
#pragma unroll
            for (int u = 0; u < kLtsUnroll; ++u) {
                const int i = min(i0 + u * kLtsThreads, n - 1);
                for (int c = 0; c < 3; ++c) {
                    pq[u][c] = P[3 * i + c];
                    pq[u][3 + c] = Q[3 * i + c];
                }
            }
            batch_loaded(pq);
#pragma unroll
            for (int u = 0; u < kLtsUnroll; ++u) {
                const int i = i0 + u * kLtsThreads;
                if (i >= n) break;
                const float3 t = mat44_mul_f(s.M, pq[u][0], pq[u][1], pq[u][2]);
                const float dx = t.x - pq[u][3], dy = t.y - pq[u][4], dz = t.z - pq[u][5];
                const uint32_t key = __float_as_uint(sqrtf(dx * dx + dy * dy + dz * dz));
                res[i] = key;
                atomicAdd(&s.hist[key >> 21], 1u);
            }
        }
        __syncthreads();
        // radix select of the keep-th smallest residual: digits of 11, 11 and 10 bits
        uint32_t prefix = 0, prefMask = 0, eqCount = 0;
        int k = (int)keep;
        const int shifts[3] = { 21, 10, 0 }, widths[3] = { 11, 11, 10 };
        for (int d = 0; d < 3; ++d) {
            const int sh = shifts[d];
            const uint32_t dm = (1u << widths[d]) - 1u;
            if (d > 0) {
                for (int b = tid; b < 2048; b += kLtsThreads) s.hist[b] = 0;
                __syncthreads();
                for (int i = tid; i < n; i += kLtsThreads) {
                    const uint32_t key = res[i];
                    if ((key & prefMask) == prefix) atomicAdd(&s.hist[(key >> sh) & dm], 1u);
                }
                __syncthreads();
            }
            if (tid < 64) {
                // the group of 32 bins holding the k-th key (prefix sums over the 64 groups),
                // then the bin inside it (prefix sums over the group's 32 bins), both found
                // by a ballot: the first bin whose inclusive count reaches k
                const uint4 *grp = reinterpret_cast<const uint4 *>(&s.hist[tid * 32]);
                uint32_t local = 0;
#pragma unroll
                for (int b = 0; b < 8; ++b) {
                    const uint4 v = grp[b];
                    local += v.x + v.y + v.z + v.w;
                }
                const uint32_t excl = wave_excl_scan(local);
                const unsigned long long inGroup = __ballot(excl < (uint32_t)k && (uint32_t)k <= excl + local);
                const int g = inGroup ? __ffsll((long long)inGroup) - 1 : 0;
                const uint32_t base = __shfl(excl, g, 64);
                const uint32_t h = tid < 32 ? s.hist[g * 32 + tid] : 0u;
                const uint32_t before = base + wave_excl_scan(h);
                const unsigned long long inBin = __ballot(tid < 32 && before + h >= (uint32_t)k);
                if (tid == (inBin ? __ffsll((long long)inBin) - 1 : 0)) {
                    s.scal[1] = g * 32 + tid;
                    s.scal[2] = k - (int)before;
                }
            }
            __syncthreads();
            if (d == 2) eqCount = s.hist[s.scal[1]];  // keys equal to the whole threshold
            prefix |= (uint32_t)s.scal[1] << sh;
            prefMask |= dm << sh;
            k = s.scal[2];
            __syncthreads();
        }
        // keep: residual below the threshold, or equal to it within the first k in index order
        // (when every key equal to the threshold is kept, the order plays no part: one pass)
        // the distance and the kept points' sums for the next fit (p, q, count, p p^T, p q^T): the
        // fit then needs no pass over the points of its own
        double dv[23];
        for (int c = 0; c < 23; ++c) dv[c] = 0.0;
        auto add_point = [&](const float *pf, const float *qf, uint32_t key) {
            const double p[3] = { (double)pf[0], (double)pf[1], (double)pf[2] };
            const double q[3] = { (double)qf[0], (double)qf[1], (double)qf[2] };
            dv[0] += (double)__uint_as_float(key);
            for (int c = 0; c < 3; ++c) {
                dv[1 + c] += p[c];
                dv[4 + c] += q[c];
            }
            dv[7] += 1.0;
            dv[8] += p[0] * p[0]; dv[9] += p[0] * p[1]; dv[10] += p[0] * p[2];
            dv[11] += p[1] * p[1]; dv[12] += p[1] * p[2]; dv[13] += p[2] * p[2];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) dv[14 + a * 3 + b] += p[a] * q[b];
        };
        auto add_kept = [&](int i, uint32_t key) { add_point(P + 3 * i, Q + 3 * i, key); };
        if ((uint32_t)k == eqCount) {
            for (int i0 = tid; i0 < n; i0 += kLtsUnroll * kLtsThreads) {
                float pq[kLtsUnroll][6];
#pragma unroll
                for (int u = 0; u < kLtsUnroll; ++u) {
                    const int i = min(i0 + u * kLtsThreads, n - 1);
                    for (int c = 0; c < 3; ++c) {
                        pq[u][c] = P[3 * i + c];
                        pq[u][3 + c] = Q[3 * i + c];
                    }
                }
                batch_loaded(pq);
#pragma unroll
                for (int u = 0; u < kLtsUnroll; ++u) {
                    const int i = i0 + u * kLtsThreads;
                    if (i >= n) break;
                    const uint32_t key = res[i];
                    const bool kept = key <= prefix;
                    keepFlag[i] = kept ? 1 : 0;
                    if (kept) add_point(pq[u], pq[u] + 3, key);
                }
            }
        } else {
            int seenEqual = 0;
            for (int base = 0; base < n; base += kLtsThreads) {
                const int i = base + tid;
                const uint32_t key = i < n ? res[i] : 0xffffffffu;
                int total;
                const int eqPos = seenEqual + wg_scan(i < n && key == prefix, s, &total);
                const bool kept = i < n && (key < prefix || (key == prefix && eqPos < k));
                if (i < n) keepFlag[i] = kept ? 1 : 0;
                if (kept) add_kept(i, key);
                seenEqual += total;
            }
        }
        wg_sum<23>(dv, s);
        const double distance = s.out[0];
        if (distance > lastDistance || (lastDistance - distance) < 1e-3) {
            if (tid < 16) s.M[tid] = s.last[tid];
            __syncthreads();
            break;
        }
        lastDistance = distance;
        if (tid < 16) s.last[tid] = s.M[tid];
        __syncthreads();
        if (!lts_fit(P, Q, keepFlag, n, affine != 0, s, true)) {
            if (tid == 0) *status = 2;
            return;
        }
    }
    if (tid < 16) matrix[tid] = s.M[tid];
}

__global__ void __launch_bounds__(kLtsThreads)
k_lts(const float *__restrict__ refPos, const float *__restrict__ warpPos, int nActive, float *__restrict__ matrix,
      float *__restrict__ P, float *__restrict__ Q, uint32_t *__restrict__ res, int percent, int affine,
      int *__restrict__ status) {
    __shared__ LtsShared s;
    lts_body(refPos, warpPos, nActive, matrix, P, Q, res, percent, affine, status, s);
}

// The same fit with the residuals and keep flags of up to kLtsLdsN correspondences in LDS (the
// radix passes and the fits then read them from LDS instead of memory).
constexpr int kLtsLdsN = 16384;

__global__ void __launch_bounds__(kLtsThreads)
k_lts_lds(const float *__restrict__ refPos, const float *__restrict__ warpPos, int nActive, float *__restrict__ matrix,
          float *__restrict__ P, float *__restrict__ Q, int percent, int affine, int *__restrict__ status) {
    __shared__ LtsShared s;
    __shared__ uint32_t sRes[kLtsLdsN];
    __shared__ uint8_t sKeep[kLtsLdsN];
    lts_body(refPos, warpPos, nActive, matrix, P, Q, sRes, percent, affine, status, s, sKeep);
}

// One LTS fit of nActive correspondences on `stream`: in LDS when they fit, else with the
// residual scratch `res` (nActive + nActive / 4 words) in memory.
void launch_lts(hipStream_t stream, const float *refPos, const float *warpPos, int nActive, float *matrix, float *P,
                float *Q, uint32_t *res, int percent, int affine, int *status) {
    if (nActive <= kLtsLdsN)
        hipLaunchKernelGGL(k_lts_lds, dim3(1), dim3(kLtsThreads), 0, stream, refPos, warpPos, nActive, matrix, P, Q,
                           percent, affine, status);
    else
        hipLaunchKernelGGL(k_lts, dim3(1), dim3(kLtsThreads), 0, stream, refPos, warpPos, nActive, matrix, P, Q, res,
                           percent, affine, status);
}

// The LTS fits of a batch of pairs in one launch: one workgroup per pair.
struct LtsArgs {
    const float *refPos, *warpPos;
    int nActive;
    float *matrix, *P, *Q;
    uint32_t *res;
    int percent;
    int *status;
};

__global__ void __launch_bounds__(kLtsThreads) k_lts_batch(const LtsArgs *__restrict__ A, int affine) {
    __shared__ LtsShared s;
    const LtsArgs& a = A[blockIdx.x];
    lts_body(a.refPos, a.warpPos, a.nActive, a.matrix, a.P, a.Q, a.res, a.percent, affine, a.status, s);
}

// Empties the defined-block count before a block matching (a kernel, so that an iteration can be
// captured in a graph).
__global__ void k_zero_int(int *p) {
    if (threadIdx.x == 0) *p = 0;
}

// ------------------------------------------------------------------ symmetric update
// reg_aladin_sym::UpdateTransformationMatrix (reg-lib/_reg_aladin_sym.cpp:128-157) after the two
// LTS fits: forward <- avg(forward, inv(backward)), backward <- avg(inv(forward), backward), where
// avg = NiftyReg::Mat44Avg2 (reg-lib/cpu/Maths.cpp:479-491) = expm((logm A + logm B) / 2).  The
// reference evaluates logm / expm with Eigen in double; here they are computed in double by
// inverse scaling and squaring (Denman-Beavers square roots, then the [8/8] Pade approximant of
// log(I + X) in Gauss-Legendre form) and by Taylor scaling and squaring, with the reference's
// float roundings between the steps.  One thread: 4x4 matrices, once per iteration.
// Mat44Avg2: the logarithms rounded to float with zero bottom rows, summed in double and
// rounded to float, halved, exponentiated in double, rounded to float.
__device__ bool mat44_avg2(const float *A, const float *B, const GaussLegendre8& gl, float *out) {
    double a[16], b[16], la[16], lb[16], h[16], e[16];
    for (int i = 0; i < 16; ++i) {
        a[i] = A[i];
        b[i] = B[i];
    }
    if (!d4_logm(a, gl, la) || !d4_logm(b, gl, lb)) return false;
    for (int i = 0; i < 16; ++i) {
        float fa = (float)la[i], fb = (float)lb[i];
        if (i >= 12) fa = fb = 0.f;
        const float sum = (float)((double)fa + (double)fb);
        h[i] = (double)(sum * 0.5f);
    }
    d4_expm(h, e);
    for (int i = 0; i < 16; ++i) out[i] = (float)e[i];
    return true;
}

__global__ void __launch_bounds__(64)
k_symmetric_update(float *__restrict__ fw, float *__restrict__ bw, GaussLegendre8 gl, int *__restrict__ status) {
    if (threadIdx.x != 0) return;
    float F[16], B[16], fi[16], bi[16], nf[16], nb[16];
    for (int i = 0; i < 16; ++i) {
        F[i] = fw[i];
        B[i] = bw[i];
    }
    nifti_inverse(F, fi);
    nifti_inverse(B, bi);
    if (!mat44_avg2(F, bi, gl, nf) || !mat44_avg2(fi, B, gl, nb)) {
        *status = 3;  // matrices left unchanged
        return;
    }
    for (int i = 12; i < 15; ++i) nf[i] = nb[i] = 0.f;
    nf[15] = nb[15] = 1.f;
    for (int i = 0; i < 16; ++i) {
        fw[i] = nf[i];
        bw[i] = nb[i];
    }
}

// ------------------------------------------------------------------ host side
unsigned blocks_for(size_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

void check_aladin_params(const nrh_aladin_params *p) {
    NRH_REQUIRE(p, NRH_ERR_ARG, "null Aladin parameters");
    NRH_REQUIRE(p->block_percent > 0 && p->block_percent <= 100, NRH_ERR_ARG, "block percentage must be in (0, 100]");
    NRH_REQUIRE(p->inlier_lts > 0 && p->inlier_lts <= 100, NRH_ERR_ARG, "LTS inlier percentage must be in (0, 100]");
}

// initialise_block_matching_method (reg-lib/cpu/_reg_blockMatching.cpp:213-273) and
// _reg_set_active_blocks<float> (:49-206).
void select_blocks(nrh_aladin *h) {
    h->nb[0] = (h->ref.nx + 3) / 4;
    h->nb[1] = (h->ref.ny + 3) / 4;
    h->nb[2] = (h->ref.nz + 3) / 4;
    const int total = h->nb[0] * h->nb[1] * h->nb[2];
    h->totalBlocks = total;
    DevBuf<float> var;
    var.alloc(total);
    hipLaunchKernelGGL(k_block_variance, dim3(blocks_for(total)), dim3(kBlock), 0, h->stream, h->refImg.p,
                       h->mask.p, h->ref.nx, h->ref.ny, h->ref.nz, h->nb[0], h->nb[1], h->nb[2], var.p);
    NRH_HIP(hipGetLastError());
    std::vector<float> v(total);
    NRH_HIP(hipMemcpyAsync(v.data(), var.p, total * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    NRH_HIP(hipStreamSynchronize(h->stream));
    int unusable = 0;
    for (float x : v) unusable += x < 0.f;
    int active = (int)((double)total * ((double)h->params.block_percent / 100.0));
    active = std::min(active, total - unusable);
    std::vector<int> ix(total);
    for (int b = 0; b < total; ++b) ix[b] = b;
    heap_sort_indexed(v, ix);
    h->totalBlock.assign(total, -1);
    std::vector<int> list(std::max(active, 0));
    for (int r = 0; r < active; ++r) {
        const int b = ix[total - 1 - r];
        h->totalBlock[b] = r;
        list[r] = b;
    }
    NRH_REQUIRE(active >= 2, NRH_ERR_NUMERIC, "There are less than 2 active blocks");
    h->activeBlocks = active;
    h->activeList.alloc(active);
    NRH_HIP(hipMemcpyAsync(h->activeList.p, list.data(), active * sizeof(int), hipMemcpyHostToDevice, h->stream));
    h->refPos.alloc(3 * (size_t)active);
    h->warpPos.alloc(3 * (size_t)active);
    h->ltsP.alloc(3 * (size_t)active);
    h->ltsQ.alloc(3 * (size_t)active);
    h->ltsRes.alloc((size_t)active + (active + 3) / 4);  // residual bits, then one keep byte per point
    NRH_HIP(hipStreamSynchronize(h->stream));
}

Mat4 mat4_of(const float *m) {
    Mat4 r;
    for (int i = 0; i < 16; ++i) r.m[i] = m[i];
    return r;
}

void launch_warp(nrh_aladin *h) {
    hipLaunchKernelGGL(k_affine_warp, dim3(blocks_for(h->N)), dim3(kBlock), 0, h->stream, h->floImg.p, h->flo.nx,
                       h->flo.ny, h->flo.nz, to_affine(h->flo.ijk), h->matrix.p, mat4_of(h->ref.xyz), h->ref.nx,
                       h->ref.ny, h->ref.nz, h->mask.p, h->params.padding, h->warped.p);
    NRH_HIP(hipGetLastError());
}

void launch_block_matching(nrh_aladin *h) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, h->stream, h->counters.p);
    hipLaunchKernelGGL(k_block_match, dim3(h->activeBlocks), dim3(kMatchThreads), 0, h->stream, h->refImg.p,
                       h->warped.p, h->mask.p, h->ref.nx, h->ref.ny, h->ref.nz, h->nb[0], h->nb[1], h->activeList.p,
                       to_affine(h->ref.xyz), h->refPos.p, h->warpPos.p, h->counters.p);
    NRH_HIP(hipGetLastError());
}

void launch_lts(nrh_aladin *h, bool affine) {
    launch_lts(h->stream, h->refPos.p, h->warpPos.p, h->activeBlocks, h->matrix.p, h->ltsP.p, h->ltsQ.p,
               h->ltsRes.p, h->params.inlier_lts, affine ? 1 : 0, h->counters.p + 1);
    NRH_HIP(hipGetLastError());
}

void read_counters(nrh_aladin *h) {
    NRH_HIP(hipMemcpyAsync(h->pinned, h->counters.p, 3 * sizeof(int), hipMemcpyDeviceToHost, h->stream));
    NRH_HIP(hipStreamSynchronize(h->stream));
}

void raise_lts_status(int status) {
    NRH_REQUIRE(status != 1, NRH_ERR_NUMERIC,
                "Not enough correspondences were found - it is impossible to estimate the transformation");
    NRH_REQUIRE(status != 2, NRH_ERR_NUMERIC, "singular least-squares system in the LTS fit");
}

void raise_symmetric_status(int status) {
    NRH_REQUIRE(status == 0, NRH_ERR_NUMERIC,
                "the matrix logarithm of the symmetric update failed (a matrix with a non-positive determinant)");
}

// Both matrices live on the device; the kernel runs on the forward handle's stream after the
// backward handle's queued work.
void launch_symmetric_update(nrh_aladin *fw, nrh_aladin *bw) {
    static const GaussLegendre8 gl = gauss_legendre8();
    if (fw->stream != bw->stream) {
        hipEvent_t ev;
        NRH_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        NRH_HIP(hipEventRecord(ev, bw->stream));
        NRH_HIP(hipStreamWaitEvent(fw->stream, ev, 0));
        NRH_HIP(hipEventDestroy(ev));
    }
    hipLaunchKernelGGL(k_symmetric_update, dim3(1), dim3(64), 0, fw->stream, fw->matrix.p, bw->matrix.p, gl,
                       fw->counters.p + 2);
    NRH_HIP(hipGetLastError());
}

void destroy(nrh_aladin *h);

nrh_aladin *create(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                   const int *mask, const nrh_aladin_params *params, hipStream_t sharedStream = nullptr) {
    check_geom(ref, "reference");
    check_geom(flo, "floating");
    check_aladin_params(params);
    NRH_REQUIRE(ref_data && flo_data, NRH_ERR_ARG, "null image data");
    NRH_REQUIRE(ref->nz > 1, NRH_ERR_UNSUPPORTED, "2-D images are not supported");
    std::unique_ptr<nrh_aladin> h(new nrh_aladin);
    h->ref = *ref;
    h->flo = *flo;
    h->params = *params;
    h->N = voxels(*ref);
    h->Nf = voxels(*flo);
    NRH_HIP(hipGetDevice(&h->device));
    if (sharedStream) {
        h->stream = sharedStream;
        h->ownsStream = false;
    } else {
        NRH_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    }
    NRH_HIP(hipHostMalloc((void **)&h->pinned, 4 * sizeof(int), hipHostMallocDefault));
    h->refImg.alloc(h->N);
    h->floImg.alloc(h->Nf);
    h->warped.alloc(h->N);
    NRH_HIP(hipMemcpyAsync(h->refImg.p, ref_data, h->N * sizeof(float), hipMemcpyHostToDevice, h->stream));
    NRH_HIP(hipMemcpyAsync(h->floImg.p, flo_data, h->Nf * sizeof(float), hipMemcpyHostToDevice, h->stream));
    NRH_HIP(hipMemsetAsync(h->warped.p, 0, h->warped.bytes(), h->stream));
    if (mask) {
        bool any = false;
        for (size_t i = 0; i < h->N && !any; ++i) any = mask[i] < 0;
        if (any) {
            h->mask.alloc(h->N);
            NRH_HIP(hipMemcpyAsync(h->mask.p, mask, h->N * sizeof(int), hipMemcpyHostToDevice, h->stream));
        }
    }
    h->matrix.alloc(16);
    const float eye[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    NRH_HIP(hipMemcpyAsync(h->matrix.p, eye, sizeof(eye), hipMemcpyHostToDevice, h->stream));
    h->counters.alloc(4);
    NRH_HIP(hipMemsetAsync(h->counters.p, 0, h->counters.bytes(), h->stream));
    h->pinned[0] = -1;
    h->pinned[1] = 0;
    h->pinned[2] = 0;
    try {
        select_blocks(h.get());
    } catch (...) {
        destroy(h.release());
        throw;
    }
    return h.release();
}

void destroy(nrh_aladin *h) {
    if (!h) return;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->pinned) (void)hipHostFree(h->pinned);
    hipStream_t s = h->ownsStream ? h->stream : nullptr;
    delete h;
    if (s) (void)hipStreamDestroy(s);
}

void set_matrix(nrh_aladin *h, const float *m) {
    NRH_HIP(hipMemcpyAsync(h->matrix.p, m, 16 * sizeof(float), hipMemcpyHostToDevice, h->stream));
    NRH_HIP(hipStreamSynchronize(h->stream));
}

void get_matrix(nrh_aladin *h, float *m) {
    NRH_HIP(hipMemcpyAsync(m, h->matrix.p, 16 * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    NRH_HIP(hipStreamSynchronize(h->stream));
}

struct AladinHandle {
    nrh_aladin *h = nullptr;
    ~AladinHandle() { destroy(h); }
};

// reg_aladin<float>::Run (reg-lib/_reg_aladin.cpp:315-364) with InitialiseRegistration's
// centre alignment (:160-183); with cfg.symmetric, reg_aladin_sym<float> (reg-lib/_reg_aladin_sym.cpp):
// a backward handle (floating -> reference, every voxel active) shares the forward handle's
// stream, its matrix starts as the inverse of the forward one, and each iteration warps both
// ways, fits both ways and averages the two matrices on the device.
void aladin_register(const nrh_geom& refG, const float *refData, const nrh_geom& floG, const float *floData,
                     const nrh_aladin_config& cfg, float *out) {
    NRH_REQUIRE(cfg.levels >= 1 && cfg.max_iterations >= 0, NRH_ERR_ARG, "invalid Aladin configuration");
    NRH_REQUIRE(cfg.rigid || cfg.affine, NRH_ERR_ARG, "neither a rigid nor an affine stage requested");
    const int toPerform = cfg.levels_to_perform > 0 ? std::min(cfg.levels_to_perform, cfg.levels) : cfg.levels;
    float M[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
    if (cfg.align_centre) {
        auto centre = [](const nrh_geom& g, float c[3]) {
            const float v[3] = { (float)g.nx / 2.0f, (float)g.ny / 2.0f, (float)g.nz / 2.0f };
            for (int r = 0; r < 3; ++r)
                c[r] = (float)((double)g.xyz[r * 4] * v[0] + (double)g.xyz[r * 4 + 1] * v[1] +
                               (double)g.xyz[r * 4 + 2] * v[2] + (double)g.xyz[r * 4 + 3]);
        };
        float fc[3], rc[3];
        centre(floG, fc);
        centre(refG, rc);
        M[3] = fc[0] - rc[0];
        M[7] = fc[1] - rc[1];
        M[11] = fc[2] - rc[2];
    }
    float Mb[16];
    mat44_inverse(M, Mb);
    std::vector<Level> rp = pyramid(refG, refData, cfg.levels, toPerform);
    std::vector<Level> fp = pyramid(floG, floData, cfg.levels, toPerform);
    const nrh_aladin_params params{ cfg.block_percent, cfg.inlier_lts, std::numeric_limits<float>::quiet_NaN() };
    for (int level = 0; level < toPerform; ++level) {
        AladinHandle ah, bh;  // bh (sharing ah's stream) is destroyed first
        ah.h = create(&rp[level].geom, rp[level].data.data(), &fp[level].geom, fp[level].data.data(), nullptr, &params);
        nrh_aladin *h = ah.h, *b = nullptr;
        set_matrix(h, M);
        if (cfg.symmetric) {
            bh.h = create(&fp[level].geom, fp[level].data.data(), &rp[level].geom, rp[level].data.data(), nullptr,
                          &params, h->stream);
            b = bh.h;
            set_matrix(b, Mb);
        }
        const int maxIt = level == 0 ? 2 * cfg.max_iterations : cfg.max_iterations;
        auto resolve = [&](int iterations, bool affine) {
            for (int it = 0; it < iterations; ++it) {
                launch_warp(h);
                if (b) launch_warp(b);
                launch_block_matching(h);
                launch_lts(h, affine);
                if (b) {
                    launch_block_matching(b);
                    launch_lts(b, affine);
                    launch_symmetric_update(h, b);
                }
            }
        };
        if ((cfg.rigid && !cfg.affine) || (cfg.affine && cfg.rigid && level == 0))
            resolve(maxIt * ((cfg.affine && cfg.rigid && level == 0) ? 4 : 1), false);
        if (cfg.affine) resolve(maxIt, true);
        read_counters(h);
        raise_lts_status(h->pinned[1]);
        raise_symmetric_status(h->pinned[2]);
        if (b) {
            read_counters(b);
            raise_lts_status(b->pinned[1]);
            get_matrix(b, Mb);
        }
        get_matrix(h, M);
    }
    std::memcpy(out, M, sizeof(M));
}

// reg_aladin<float>::Run for many pairs at once (non-symmetric): every pair runs aladin_register's
// levels, stages and iterations, with the pairs' warps and block matchings launched in turn on one
// stream and all pairs' LTS fits in one launch (a workgroup each); an iteration of all pairs is
// captured once per stage as a HIP graph and replayed.  A pair whose block selection or fit fails
// (a flat image: fewer than 2 active blocks, or too few correspondences) stops with its error in
// status[q]; the other pairs go on.
void aladin_register_batch(int n, const nrh_geom *refs, const float *const *refData, const nrh_geom *flos,
                           const float *const *floData, const nrh_aladin_config& cfg, float *out, int *status,
                           nrh_batch_stats *stats) {
    NRH_REQUIRE(cfg.levels >= 1 && cfg.max_iterations >= 0, NRH_ERR_ARG, "invalid Aladin configuration");
    NRH_REQUIRE(cfg.rigid || cfg.affine, NRH_ERR_ARG, "neither a rigid nor an affine stage requested");
    NRH_REQUIRE(!cfg.symmetric, NRH_ERR_UNSUPPORTED, "batched Aladin runs the non-symmetric registration");
    const auto t0 = std::chrono::steady_clock::now();
    const int toPerform = cfg.levels_to_perform > 0 ? std::min(cfg.levels_to_perform, cfg.levels) : cfg.levels;
    std::vector<std::array<float, 16>> M(n);
    std::vector<std::vector<Level>> rp(n), fp(n);
    std::vector<std::string> errs(n);
    {
        std::vector<std::thread> pool;
        const int T = std::max(1, std::min(n, (int)std::thread::hardware_concurrency()));
        for (int t = 0; t < T; ++t)
            pool.emplace_back([&, t] {
                for (int q = t; q < n; q += T) {
                    try {
                        rp[q] = pyramid(refs[q], refData[q], cfg.levels, toPerform);
                        fp[q] = pyramid(flos[q], floData[q], cfg.levels, toPerform);
                    } catch (const std::exception& e) {
                        errs[q] = e.what();
                    }
                }
            });
        for (auto& th : pool) th.join();
    }
    for (int q = 0; q < n; ++q) {
        NRH_REQUIRE(errs[q].empty(), NRH_ERR_ARG, "pair " + std::to_string(q) + ": " + errs[q]);
        float m[16] = { 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1 };
        if (cfg.align_centre) {
            auto centre = [](const nrh_geom& g, float c[3]) {
                const float v[3] = { (float)g.nx / 2.0f, (float)g.ny / 2.0f, (float)g.nz / 2.0f };
                for (int r = 0; r < 3; ++r)
                    c[r] = (float)((double)g.xyz[r * 4] * v[0] + (double)g.xyz[r * 4 + 1] * v[1] +
                                   (double)g.xyz[r * 4 + 2] * v[2] + (double)g.xyz[r * 4 + 3]);
            };
            float fc[3], rc[3];
            centre(flos[q], fc);
            centre(refs[q], rc);
            m[3] = fc[0] - rc[0];
            m[7] = fc[1] - rc[1];
            m[11] = fc[2] - rc[2];
        }
        std::copy(m, m + 16, M[q].begin());
        status[q] = NRH_OK;
    }
    hipStream_t stream;
    NRH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    struct StreamOwner {
        hipStream_t s;
        ~StreamOwner() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
    } so{ stream };
    const nrh_aladin_params params{ cfg.block_percent, cfg.inlier_lts, std::numeric_limits<float>::quiet_NaN() };
    double deviceSeconds = 0.0;
    int replays = 0;
    for (int level = 0; level < toPerform; ++level) {
        std::vector<AladinHandle> hs(n);
        std::vector<nrh_aladin *> live;
        std::vector<int> liveIdx;
        for (int q = 0; q < n; ++q) {
            if (status[q] != NRH_OK) continue;
            try {
                hs[q].h = create(&rp[q][level].geom, rp[q][level].data.data(), &fp[q][level].geom,
                                 fp[q][level].data.data(), nullptr, &params, stream);
                set_matrix(hs[q].h, M[q].data());
                live.push_back(hs[q].h);
                liveIdx.push_back(q);
            } catch (const Error& e) {
                if (e.code != NRH_ERR_NUMERIC) throw;
                status[q] = e.code;  // this pair stops; the others go on
            }
        }
        if (live.empty()) break;
        const int L = (int)live.size();
        std::vector<LtsArgs> la(L);
        for (int k = 0; k < L; ++k) {
            nrh_aladin *h = live[k];
            la[k] = LtsArgs{ h->refPos.p, h->warpPos.p, h->activeBlocks, h->matrix.p, h->ltsP.p, h->ltsQ.p, h->ltsRes.p,
                             h->params.inlier_lts, h->counters.p + 1 };
        }
        DevBuf<LtsArgs> dla;
        dla.alloc(L);
        NRH_HIP(hipMemcpyAsync(dla.p, la.data(), L * sizeof(LtsArgs), hipMemcpyHostToDevice, stream));
        NRH_HIP(hipStreamSynchronize(stream));
        const int maxIt = level == 0 ? 2 * cfg.max_iterations : cfg.max_iterations;
        auto resolve = [&](int iterations, bool affine) {
            if (iterations <= 0) return;
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            NRH_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            for (nrh_aladin *h : live) {
                launch_warp(h);
                launch_block_matching(h);
            }
            hipLaunchKernelGGL(k_lts_batch, dim3(L), dim3(kLtsThreads), 0, stream, dla.p, affine ? 1 : 0);
            NRH_HIP(hipStreamEndCapture(stream, &graph));
            const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            NRH_HIP(ie);
            const auto d0 = std::chrono::steady_clock::now();
            hipError_t ge = hipSuccess;
            for (int it = 0; it < iterations && ge == hipSuccess; ++it, ++replays) ge = hipGraphLaunch(exec, stream);
            if (ge == hipSuccess) ge = hipStreamSynchronize(stream);
            (void)hipGraphExecDestroy(exec);
            NRH_HIP(ge);
            deviceSeconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - d0).count();
        };
        if ((cfg.rigid && !cfg.affine) || (cfg.affine && cfg.rigid && level == 0))
            resolve(maxIt * ((cfg.affine && cfg.rigid && level == 0) ? 4 : 1), false);
        if (cfg.affine) resolve(maxIt, true);
        for (int k = 0; k < L; ++k) {
            nrh_aladin *h = live[k];
            read_counters(h);
            const int q = liveIdx[k];
            if (h->pinned[1] != 0) {
                status[q] = NRH_ERR_NUMERIC;  // too few correspondences or a singular fit
                continue;
            }
            get_matrix(h, M[q].data());
        }
    }
    for (int q = 0; q < n; ++q) std::copy(M[q].begin(), M[q].end(), out + 16 * q);
    if (stats) {
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        stats->device_seconds = deviceSeconds;
        stats->graph_replays = replays;
    }
}

// ------------------------------------------------------------------ allocation-free LTS
// Scratch of nrh_lts_static: the compacted correspondences (3 n floats each) and the residual
// bits plus keep bytes (n + n / 4 words) of one fit of at most kLtsStaticMax correspondences
// (7.6 MB on every GPU the library runs on, held until the process ends).  Host code never uses
// the array's name as an address: each device's address comes from hipGetSymbolAddress.
constexpr int kLtsStaticMax = 262144;
constexpr size_t kLtsScratchWords = 6 * (size_t)kLtsStaticMax + kLtsStaticMax + (kLtsStaticMax + 3) / 4;
static __device__ __align__(256) uint32_t scratch[kLtsScratchWords];

}  // namespace

// The address of `scratch` on the current device, looked up once per device (nrh_init does it for
// the device it selects; nrh_lts_static for any other current device on its first call there).
uint32_t *lts_static_scratch() {
    static std::mutex mu;
    static std::map<int, uint32_t *> ptrs;
    int dev = 0;
    NRH_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end()) return it->second;
    uint32_t *ptr = nullptr;
    NRH_HIP(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    NRH_REQUIRE(ptr != nullptr, NRH_ERR_HIP, "no device address for the LTS scratch");
    ptrs[dev] = ptr;
    return ptr;
}

}  // namespace nrh

using namespace nrh;

extern "C" {

int nrh_aladin_create(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                      const int *mask, const nrh_aladin_params *params, nrh_aladin **out) {
    return guard([&] {
        NRH_REQUIRE(out, NRH_ERR_ARG, "null output handle");
        *out = nullptr;
        *out = create(ref, ref_data, flo, flo_data, mask, params);
    });
}

int nrh_aladin_register_batch(int n, const nrh_geom *refs, const float *const *ref_data, const nrh_geom *flos,
                              const float *const *flo_data, const nrh_aladin_config *cfg, float *matrices,
                              int *status, nrh_batch_stats *stats) {
    return guard([&] {
        NRH_REQUIRE(n >= 1 && refs && ref_data && flos && flo_data && cfg && matrices && status, NRH_ERR_ARG,
                    "null argument");
        for (int q = 0; q < n; ++q) {
            check_geom(&refs[q], "reference");
            check_geom(&flos[q], "floating");
            NRH_REQUIRE(ref_data[q] && flo_data[q], NRH_ERR_ARG, "null image");
            NRH_REQUIRE(refs[q].nz > 1, NRH_ERR_UNSUPPORTED, "2-D images are not supported");
        }
        if (stats) *stats = nrh_batch_stats{};
        aladin_register_batch(n, refs, ref_data, flos, flo_data, *cfg, matrices, status, stats);
    });
}

int nrh_aladin_destroy(nrh_aladin *h) {
    return guard([&] { destroy(h); });
}

int nrh_aladin_blocks(nrh_aladin *h, int counts[3], int *total_block) {
    return guard([&] {
        NRH_REQUIRE(h && counts, NRH_ERR_ARG, "null argument");
        read_counters(h);
        counts[0] = h->totalBlocks;
        counts[1] = h->activeBlocks;
        counts[2] = h->pinned[0];
        if (total_block) std::memcpy(total_block, h->totalBlock.data(), h->totalBlocks * sizeof(int));
    });
}

int nrh_aladin_set_matrix(nrh_aladin *h, const float m[16]) {
    return guard([&] {
        NRH_REQUIRE(h && m, NRH_ERR_ARG, "null argument");
        set_matrix(h, m);
    });
}

int nrh_aladin_get_matrix(nrh_aladin *h, float m[16]) {
    return guard([&] {
        NRH_REQUIRE(h && m, NRH_ERR_ARG, "null argument");
        get_matrix(h, m);
    });
}

int nrh_aladin_warp(nrh_aladin *h) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        launch_warp(h);
    });
}

int nrh_aladin_upload_warped(nrh_aladin *h, const float *warped) {
    return guard([&] {
        NRH_REQUIRE(h && warped, NRH_ERR_ARG, "null argument");
        NRH_HIP(hipMemcpyAsync(h->warped.p, warped, h->N * sizeof(float), hipMemcpyHostToDevice, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

int nrh_aladin_download_warped(nrh_aladin *h, float *warped) {
    return guard([&] {
        NRH_REQUIRE(h && warped, NRH_ERR_ARG, "null argument");
        NRH_HIP(hipMemcpyAsync(warped, h->warped.p, h->N * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
    });
}

int nrh_aladin_block_matching(nrh_aladin *h) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        launch_block_matching(h);
    });
}

int nrh_aladin_download_positions(nrh_aladin *h, float *ref_pos, float *warp_pos, int *defined) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        const size_t bytes = 3 * (size_t)h->activeBlocks * sizeof(float);
        if (ref_pos) NRH_HIP(hipMemcpyAsync(ref_pos, h->refPos.p, bytes, hipMemcpyDeviceToHost, h->stream));
        if (warp_pos) NRH_HIP(hipMemcpyAsync(warp_pos, h->warpPos.p, bytes, hipMemcpyDeviceToHost, h->stream));
        read_counters(h);
        if (defined) *defined = h->pinned[0];
    });
}

int nrh_aladin_lts(nrh_aladin *h, int affine) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        NRH_HIP(hipMemsetAsync(h->counters.p + 1, 0, sizeof(int), h->stream));
        launch_lts(h, affine != 0);
        read_counters(h);
        raise_lts_status(h->pinned[1]);
    });
}

int nrh_lts_static(const float *ref_pos, const float *warp_pos, int n, int inlier_lts, int affine, float *matrix,
                   int *status, void *stream) {
    return guard([&] {
        NRH_REQUIRE(ref_pos && warp_pos && matrix && status, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(n > 0, NRH_ERR_ARG, "no correspondences");
        NRH_REQUIRE(n <= kLtsStaticMax, NRH_ERR_ARG,
                    "nrh_lts_static takes at most 262144 correspondences (its static scratch); use nrh_lts");
        NRH_REQUIRE(inlier_lts > 0 && inlier_lts <= 100, NRH_ERR_ARG, "LTS inlier percentage must be in (0, 100]");
        uint32_t *base = lts_static_scratch();
        float *P = (float *)base, *Q = P + 3 * (size_t)n;
        uint32_t *res = (uint32_t *)(Q + 3 * (size_t)n);
        hipStream_t st = (hipStream_t)stream;
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, status);
        launch_lts(st, ref_pos, warp_pos, n, matrix, P, Q, res, inlier_lts, affine ? 1 : 0, status);
        NRH_HIP(hipGetLastError());
    });
}

int nrh_lts(const float *ref_pos, const float *warp_pos, int n, int inlier_lts, int affine, float m[16]) {
    return guard([&] {
        NRH_REQUIRE(ref_pos && warp_pos && m, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(n > 0, NRH_ERR_ARG, "no correspondences");
        NRH_REQUIRE(inlier_lts > 0 && inlier_lts <= 100, NRH_ERR_ARG, "LTS inlier percentage must be in (0, 100]");
        hipStream_t stream;
        NRH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        struct StreamGuard {
            hipStream_t s;
            ~StreamGuard() { (void)hipStreamSynchronize(s); (void)hipStreamDestroy(s); }
        } sg{ stream };
        DevBuf<float> rp, wp, mat, P, Q;
        DevBuf<uint32_t> res;
        DevBuf<int> status;
        const size_t bytes = 3 * (size_t)n * sizeof(float);
        rp.alloc(3 * (size_t)n);
        wp.alloc(3 * (size_t)n);
        P.alloc(3 * (size_t)n);
        Q.alloc(3 * (size_t)n);
        res.alloc((size_t)n + (n + 3) / 4);
        mat.alloc(16);
        status.alloc(1);
        NRH_HIP(hipMemcpyAsync(rp.p, ref_pos, bytes, hipMemcpyHostToDevice, stream));
        NRH_HIP(hipMemcpyAsync(wp.p, warp_pos, bytes, hipMemcpyHostToDevice, stream));
        NRH_HIP(hipMemcpyAsync(mat.p, m, 16 * sizeof(float), hipMemcpyHostToDevice, stream));
        NRH_HIP(hipMemsetAsync(status.p, 0, sizeof(int), stream));
        launch_lts(stream, rp.p, wp.p, n, mat.p, P.p, Q.p, res.p, inlier_lts, affine ? 1 : 0, status.p);
        NRH_HIP(hipGetLastError());
        int st = 0;
        float out[16];
        NRH_HIP(hipMemcpyAsync(&st, status.p, sizeof(int), hipMemcpyDeviceToHost, stream));
        NRH_HIP(hipMemcpyAsync(out, mat.p, sizeof(out), hipMemcpyDeviceToHost, stream));
        NRH_HIP(hipStreamSynchronize(stream));
        raise_lts_status(st);
        std::memcpy(m, out, sizeof(out));
    });
}

int nrh_aladin_symmetric_update(nrh_aladin *forward, nrh_aladin *backward) {
    return guard([&] {
        NRH_REQUIRE(forward && backward && forward != backward, NRH_ERR_ARG, "two distinct handles are required");
        NRH_HIP(hipMemsetAsync(forward->counters.p + 2, 0, sizeof(int), forward->stream));
        launch_symmetric_update(forward, backward);
        read_counters(forward);
        if (backward->stream != forward->stream) NRH_HIP(hipStreamSynchronize(backward->stream));
        raise_symmetric_status(forward->pinned[2]);
    });
}

int nrh_aladin_register(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                        const nrh_aladin_config *cfg, float matrix[16]) {
    return guard([&] {
        check_geom(ref, "reference");
        check_geom(flo, "floating");
        NRH_REQUIRE(ref_data && flo_data && cfg && matrix, NRH_ERR_ARG, "null argument");
        aladin_register(*ref, ref_data, *flo, flo_data, *cfg, matrix);
    });
}

}  // extern "C"
