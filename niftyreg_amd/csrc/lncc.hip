// LNCC similarity (reg_lncc) on gfx950.
//
// The measure needs Gaussian-weighted local statistics of both images over the combined mask
// (reference mask minus the NaN voxels of either image): five masked convolutions per value
// (ref, ref^2, warped, warped^2, ref*warped) and three more per gradient.  All of them share one
// density (the combined mask), so one set of separable passes carries the numerators and the
// density together: planar [K + 1][N] arrays, one launch per axis, ping-ponging between two
// buffers.  Each output sample is the reference's line sum in the reference's order (float
// products added in double, reg-lib/cpu/_reg_tools.cpp:1047-1053), so the results equal the
// reference's scalar build bit for bit.

#include "f3d_handle.hpp"
#include "device.hpp"
#include <cmath>

namespace nrh {

namespace {

constexpr int kB = 256;
constexpr int kRedBlocks = 1024;

unsigned blocks_of(size_t n) { return (unsigned)((n + kB - 1) / kB); }

// Combined mask and the numerators of the first convolution set, with the density last.
__global__ void __launch_bounds__(kB)
k_lncc_prepare(const float *__restrict__ ref, const float *__restrict__ war, const int *__restrict__ mask, size_t N,
               uint8_t *__restrict__ comb, float *__restrict__ A) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const float r = ref[i], w = war[i];
    const bool ok = (!mask || mask[i] > -1) && r == r && w == w;
    comb[i] = ok ? 1 : 0;
    A[i] = ok ? r : 0.f;
    A[N + i] = ok ? r * r : 0.f;
    A[2 * N + i] = ok ? w : 0.f;
    A[3 * N + i] = ok ? w * w : 0.f;
    A[4 * N + i] = ok ? r * w : 0.f;
    A[5 * N + i] = ok ? 1.f : 0.f;
}

// One separable pass of reg_tools_kernelConvolution along `axis` for K <= 6 planar arrays.  The K
// sums of a voxel advance together tap by tap (K independent FP64 chains in flight, each in the
// reference's tap order).
constexpr int kConvArrays = 6;

__global__ void __launch_bounds__(kB)
k_conv_axis(const float *__restrict__ in, float *__restrict__ out, int K, size_t N, int nx, int ny, int nz, int axis,
            const float *__restrict__ kern, int radius) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int x, y, z;
    index_xyz(i, nx, ny, x, y, z);
    int l, L;
    size_t stride;
    if (axis == 0) { l = x; L = nx; stride = 1; }
    else if (axis == 1) { l = y; L = ny; stride = (size_t)nx; }
    else { l = z; L = nz; stride = (size_t)nx * ny; }
    int pre = l - radius, k0 = 0;
    const int pst = min(l + radius + 1, L);
    if (pre < 0) { k0 = -pre; pre = 0; }
    const size_t base = i - (size_t)l * stride;
    double acc[kConvArrays] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
    for (int k = pre, kk = k0; k < pst; ++k, ++kk) {
        const float kv = kern[kk];
        const float *src = in + base + (size_t)k * stride;
#pragma unroll
        for (int a = 0; a < kConvArrays; ++a)
            if (a < K) acc[a] += (double)(kv * src[(size_t)a * N]);
    }
#pragma unroll
    for (int a = 0; a < kConvArrays; ++a)
        if (a < K) out[(size_t)a * N + i] = (float)acc[a];
}

// Line-tiled passes.  A workgroup stages a stretch of lines (outputs plus `radius` samples on each
// side, zero outside the image) of one array at a time in LDS, so each input sample is fetched
// from memory about once per pass instead of once per tap.  Every output is still summed tap by
// tap in ascending order in double; the zero samples outside the line add +0 to a sum that starts
// at +0 and cannot become -0, so each sum equals the reference's over the in-range taps only.
constexpr int kLineOut = 64;      // outputs per workgroup along y or z (four waves of 16)
constexpr int kRowsX = 32;        // rows per workgroup for the x pass (four waves of 8)
constexpr int kMaxTileRadius = 96;

size_t lds_yz(int radius) { return (size_t)(kLineOut + 2 * radius) * 64 * sizeof(float); }
size_t lds_x(int radius) { return (size_t)kRowsX * (64 + 2 * radius) * sizeof(float); }

// y or z pass: 64 neighbouring x columns of one plane (blockIdx.y) and kLineOut positions of the
// line (blockIdx.z); LDS tile [kLineOut + 2 radius][64], lanes on x, conflict-free reads.
__global__ void __launch_bounds__(kB)
k_conv_lines_yz(const float *__restrict__ in, float *__restrict__ out, int K, size_t N, int nx, int L,
                size_t outerStride, size_t lineStride, const float *__restrict__ kern, int radius) {
#pragma clang fp contract(off)
    extern __shared__ float tile[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int x = blockIdx.x * 64 + lane;
    const bool col = x < nx;
    const size_t base = (size_t)blockIdx.y * outerStride + (size_t)blockIdx.x * 64 + lane;
    const int t0 = blockIdx.z * kLineOut;
    const int rows = kLineOut + 2 * radius, taps = 2 * radius + 1;
    for (int a = 0; a < K; ++a) {
        const float *src = in + (size_t)a * N + base;
        for (int p = w; p < rows; p += kB / 64) {
            const int pos = t0 - radius + p;
            tile[p * 64 + lane] = col && pos >= 0 && pos < L ? src[(size_t)pos * lineStride] : 0.f;
        }
        __syncthreads();
        float *dst = out + (size_t)a * N + base;
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            const int o0 = w * 16 + g * 8;
            const float *c = tile + o0 * 64 + lane;
            double acc[8] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
            for (int kk = 0; kk < taps; ++kk) {
                const float kv = kern[kk];
#pragma unroll
                for (int o = 0; o < 8; ++o) acc[o] += (double)(kv * c[(o + kk) * 64]);
            }
#pragma unroll
            for (int o = 0; o < 8; ++o) {
                const int pos = t0 + o0 + o;
                if (col && pos < L) dst[(size_t)pos * lineStride] = (float)acc[o];
            }
        }
        __syncthreads();
    }
}

// x pass: kRowsX rows (row = y + ny z, blockIdx.x) and 64 outputs of each (blockIdx.y); LDS tile
// [kRowsX][64 + 2 radius], lanes on x.
__global__ void __launch_bounds__(kB)
k_conv_lines_x(const float *__restrict__ in, float *__restrict__ out, int K, size_t N, int nx, int nrows,
               const float *__restrict__ kern, int radius) {
#pragma clang fp contract(off)
    extern __shared__ float tile[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int W = 64 + 2 * radius, taps = 2 * radius + 1;
    const int x0 = blockIdx.y * 64, row0 = blockIdx.x * kRowsX;
    for (int a = 0; a < K; ++a) {
        const float *src = in + (size_t)a * N;
        for (int idx = threadIdx.x; idx < kRowsX * W; idx += kB) {
            const int r = idx / W, p = idx - r * W;
            const int xx = x0 - radius + p, row = row0 + r;
            tile[idx] = xx >= 0 && xx < nx && row < nrows ? src[(size_t)row * nx + xx] : 0.f;
        }
        __syncthreads();
        const int r0 = w * (kRowsX / 4);
        const float *c = tile + r0 * W + lane;
        double acc[kRowsX / 4] = { 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0 };
        for (int kk = 0; kk < taps; ++kk) {
            const float kv = kern[kk];
#pragma unroll
            for (int r = 0; r < kRowsX / 4; ++r) acc[r] += (double)(kv * c[r * W + kk]);
        }
        float *dst = out + (size_t)a * N;
        const int x = x0 + lane;
#pragma unroll
        for (int r = 0; r < kRowsX / 4; ++r)
            if (x < nx && row0 + r0 + r < nrows) dst[(size_t)(row0 + r0 + r) * nx + x] = (float)acc[r];
        __syncthreads();
    }
}

// Local statistics (UpdateLocalStatImages, reg-lib/cpu/_reg_lncc.cpp:236-244) and the per-voxel
// |local correlation| of reg_getLnccValue (:282-297), reduced per block in a fixed order.
__global__ void __launch_bounds__(kB)
k_lncc_value(const float *__restrict__ C, const uint8_t *__restrict__ comb, size_t N, double *__restrict__ partSum,
             unsigned long long *__restrict__ partCnt) {
#pragma clang fp contract(off)
    __shared__ double sS[kB / 64];
    __shared__ unsigned long long sC[kB / 64];
    double sum = 0.0;
    unsigned long long cnt = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
        if (!comb[i]) continue;
        const float d = C[5 * N + i];
        const float mean = C[i] / d, sq = C[N + i] / d, wmean = C[2 * N + i] / d, wsq = C[3 * N + i] / d;
        const float corr = C[4 * N + i] / d;
        float sdev = sqrtf(sq - mean * mean), wsdev = sqrtf(wsq - wmean * wmean);
        if (sdev < 1.e-06) sdev = 0.f;
        if (wsdev < 1.e-06) wsdev = 0.f;
        const double l = (double)((corr - (mean * wmean)) / (sdev * wsdev));
        if (l == l && !isinf(l)) {
            sum += fabs(l);
            ++cnt;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        sum += __shfl_down(sum, off, 64);
        cnt += __shfl_down(cnt, off, 64);
    }
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        sS[wv] = sum;
        sC[wv] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        unsigned long long c = 0;
        for (int w = 0; w < kB / 64; ++w) {
            s += sS[w];
            c += sC[w];
        }
        partSum[blockIdx.x] = s;
        partCnt[blockIdx.x] = c;
    }
}

// Sum of the block partials in block order; out[0] = sum / count, out[1] = count.
__global__ void k_lncc_finish(const double *__restrict__ partSum, const unsigned long long *__restrict__ partCnt,
                              int nb, double *__restrict__ out) {
    // one wave: lane l sums its run of partials in order, then a fixed tree over the lanes
    const int l = threadIdx.x, per = (nb + 63) / 64;
    double s = 0.0;
    unsigned long long c = 0;
    for (int k = 0; k < per; ++k) {
        const int b = l * per + k;
        if (b < nb) {
            s += partSum[b];
            c += partCnt[b];
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        s += __shfl_down(s, off, 64);
        c += __shfl_down(c, off, 64);
    }
    if (l == 0) {
        out[0] = s / (double)c;
        out[1] = (double)c;
    }
}

// The three per-voxel derivative terms of reg_getVoxelBasedLnccGradient (:441-470), written as
// the numerators of the second convolution set (density last), and the active-voxel count.
__global__ void __launch_bounds__(kB)
k_lncc_terms(const float *__restrict__ C, const uint8_t *__restrict__ comb, size_t N, float *__restrict__ T,
             unsigned long long *__restrict__ partCnt) {
#pragma clang fp contract(off)
    __shared__ unsigned long long sC[kB / 64];
    unsigned long long cnt = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (size_t)gridDim.x * blockDim.x) {
        float t1f = 0.f, t2f = 0.f, t3f = 0.f;
        if (comb[i]) {
            const float d = C[5 * N + i];
            const float mean = C[i] / d, sq = C[N + i] / d, wmean = C[2 * N + i] / d, wsq = C[3 * N + i] / d;
            const float corr = C[4 * N + i] / d;
            float sdev = sqrtf(sq - mean * mean), wsdev = sqrtf(wsq - wmean * wmean);
            if (sdev < 1.e-06) sdev = 0.f;
            if (wsdev < 1.e-06) wsdev = 0.f;
            const double rm = mean, wm = wmean, rs = sdev, ws = wsdev;
            const double c = (double)corr - (rm * wm);
            double t1 = 1.0 / (rs * ws);
            double t2 = c / (rs * ws * ws * ws);
            double t3 = (c * wm) / (rs * ws * ws * ws) - rm / (rs * ws);
            if (t1 == t1 && !isinf(t1) && t2 == t2 && !isinf(t2) && t3 == t3 && !isinf(t3)) {
                if (c < 0) {
                    t1 = -t1;
                    t2 = -t2;
                    t3 = -t3;
                }
                t1f = (float)t1;
                t2f = (float)t2;
                t3f = (float)t3;
                ++cnt;
            }
        }
        T[i] = t1f;
        T[N + i] = t2f;
        T[2 * N + i] = t3f;
        T[3 * N + i] = comb[i] ? 1.f : 0.f;
    }
    for (int off = 32; off > 0; off >>= 1) cnt += __shfl_down(cnt, off, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) sC[wv] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long c = 0;
        for (int w = 0; w < kB / 64; ++w) c += sC[w];
        partCnt[blockIdx.x] = c;
    }
}

__global__ void k_lncc_weight(const unsigned long long *__restrict__ partCnt, int nb, double weight,
                              double *__restrict__ out) {
    unsigned long long c = 0;
    for (int b = threadIdx.x; b < nb; b += 64) c += partCnt[b];
    for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
    if (threadIdx.x == 0) out[0] = weight / (double)c;
}

// grad -= warped gradient * (T1 ref - T2 war + T3) * adjusted weight (:486-500), then non-finite
// entries of every component set to 0 (:502-511).
__global__ void __launch_bounds__(kB)
k_lncc_apply(const float *__restrict__ T, const uint8_t *__restrict__ comb, const float *__restrict__ ref,
             const float *__restrict__ war, const float *__restrict__ wgrad, const double *__restrict__ adjusted,
             size_t N, float *__restrict__ grad) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const double aw = adjusted[0];
    if (comb[i]) {
        const float d = T[3 * N + i];
        const float t1 = T[i] / d, t2 = T[N + i] / d, t3 = T[2 * N + i] / d;
        const double common = (double)(t1 * ref[i] - t2 * war[i] + t3) * aw;
        for (int c = 0; c < 3; ++c) grad[c * N + i] -= (float)((double)wgrad[c * N + i] * common);
    }
    for (int c = 0; c < 3; ++c) {
        const float v = grad[c * N + i];
        if (v != v || isinf(v)) grad[c * N + i] = 0.f;
    }
}

// Gaussian kernel of reg_tools_kernelConvolution (:866-901); returns the radius.
int gaussian_kernel(float sigma, float pixdim, std::vector<float>& k) {
    const double temp = sigma > 0 ? (double)(sigma / pixdim) : std::fabs((double)sigma);
    const int radius = (int)(temp * 3.0f);
    k.assign(2 * std::max(radius, 0) + 1, 0.f);
    for (int i = -radius; i <= radius; ++i)
        k[i + radius] = (float)(std::exp(-(double)(i * i) / (2.0 * (temp * temp))) / (temp * 2.506628274631));
    return radius;
}

void ensure_lncc(nrh_f3d *h) {
    const size_t N = h->N;
    h->lnccA.alloc(6 * N);
    h->lnccB.alloc(6 * N);
    h->lnccComb.alloc(N);
    h->lnccPart.alloc(2 * kRedBlocks);
    h->lnccScalars.alloc(4);
    if (h->lnccKern.empty() || h->lnccSigma != h->params.lncc_sigma) {
        std::vector<float> all;
        const float pix[3] = { h->ref.dx, h->ref.dy, h->ref.dz };
        for (int a = 0; a < 3; ++a) {
            std::vector<float> k;
            h->lnccRadius[a] = gaussian_kernel(h->params.lncc_sigma, pix[a], k);
            h->lnccKernOff[a] = (int)all.size();
            all.insert(all.end(), k.begin(), k.end());
        }
        h->lnccKern.alloc(all.size());
        NRH_HIP(hipMemcpyAsync(h->lnccKern.p, all.data(), all.size() * sizeof(float), hipMemcpyHostToDevice,
                               h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
        h->lnccSigma = h->params.lncc_sigma;
    }
}

// Three axis passes over K + 1 planar arrays starting in `a`; returns the buffer holding the result.
float *convolve_set(nrh_f3d *h, float *a, float *b, int K1) {
    NRH_REQUIRE(K1 >= 1 && K1 <= kConvArrays, NRH_ERR_ARG, "at most six arrays per convolution pass");
    const int dims[3] = { h->ref.nx, h->ref.ny, h->ref.nz };
    float *src = a, *dst = b;
    for (int axis = 0; axis < 3; ++axis) {
        const int r = h->lnccRadius[axis];
        if (dims[axis] <= 1 || r <= 0) continue;
        const float *kern = h->lnccKern.p + h->lnccKernOff[axis];
        if (r > kMaxTileRadius) {
            hipLaunchKernelGGL(k_conv_axis, dim3(blocks_of(h->N)), dim3(kB), 0, h->stream, src, dst, K1, h->N,
                               dims[0], dims[1], dims[2], axis, kern, r);
        } else if (axis == 0) {
            const int nrows = dims[1] * dims[2];
            hipLaunchKernelGGL(k_conv_lines_x, dim3((nrows + kRowsX - 1) / kRowsX, (dims[0] + 63) / 64),
                               dim3(kB), lds_x(r), h->stream, src, dst, K1, h->N, dims[0], nrows, kern, r);
        } else {
            const size_t plane = (size_t)dims[0] * dims[1];
            const int L = dims[axis], outer = axis == 1 ? dims[2] : dims[1];
            const size_t outerStride = axis == 1 ? plane : (size_t)dims[0];
            const size_t lineStride = axis == 1 ? (size_t)dims[0] : plane;
            hipLaunchKernelGGL(k_conv_lines_yz, dim3((dims[0] + 63) / 64, outer, (L + kLineOut - 1) / kLineOut),
                               dim3(kB), lds_yz(r), h->stream, src, dst, K1, h->N, dims[0], L, outerStride,
                               lineStride, kern, r);
        }
        NRH_HIP(hipGetLastError());
        std::swap(src, dst);
    }
    return src;
}

float *lncc_statistics(nrh_f3d *h) {
    NRH_REQUIRE(!h->warped.empty(), NRH_ERR_STATE, "LNCC needs the warped image");
    ensure_lncc(h);
    hipLaunchKernelGGL(k_lncc_prepare, dim3(blocks_of(h->N)), dim3(kB), 0, h->stream, ref_ptr(h), h->warped.p,
                       h->mask.p, h->N, h->lnccComb.p, h->lnccA.p);
    NRH_HIP(hipGetLastError());
    return convolve_set(h, h->lnccA.p, h->lnccB.p, 6);
}

}  // namespace

void op_lncc_value(nrh_f3d *h, double *out_device) {
    const float *C = lncc_statistics(h);
    const unsigned nb = std::min<unsigned>(blocks_of(h->N), kRedBlocks);
    double *ps = h->lnccPart.p;
    auto *pc = reinterpret_cast<unsigned long long *>(h->lnccPart.p + kRedBlocks);
    hipLaunchKernelGGL(k_lncc_value, dim3(nb), dim3(kB), 0, h->stream, C, h->lnccComb.p, h->N, ps, pc);
    NRH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_lncc_finish, dim3(1), dim3(64), 0, h->stream, ps, pc, (int)nb, out_device);
    NRH_HIP(hipGetLastError());
}

void op_lncc_voxel_gradient(nrh_f3d *h, double weight) {
    ensure_voxel_grad(h);
    NRH_REQUIRE(!h->warpedGrad.empty(), NRH_ERR_STATE, "LNCC gradient needs the warped image gradient");
    const float *C = lncc_statistics(h);
    float *T = (C == h->lnccA.p) ? h->lnccB.p : h->lnccA.p;
    const unsigned nb = std::min<unsigned>(blocks_of(h->N), kRedBlocks);
    auto *pc = reinterpret_cast<unsigned long long *>(h->lnccPart.p + kRedBlocks);
    hipLaunchKernelGGL(k_lncc_terms, dim3(nb), dim3(kB), 0, h->stream, C, h->lnccComb.p, h->N, T, pc);
    NRH_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_lncc_weight, dim3(1), dim3(64), 0, h->stream, pc, (int)nb, weight, h->lnccScalars.p);
    NRH_HIP(hipGetLastError());
    // the statistics buffer is free now: it becomes the ping-pong partner of the terms
    const float *S = convolve_set(h, T, const_cast<float *>(C), 4);
    hipLaunchKernelGGL(k_lncc_apply, dim3(blocks_of(h->N)), dim3(kB), 0, h->stream, S, h->lnccComb.p, ref_ptr(h),
                       h->warped.p, h->warpedGrad.p, h->lnccScalars.p, h->N, h->voxelGrad.p);
    NRH_HIP(hipGetLastError());
}

}  // namespace nrh

using namespace nrh;

extern "C" {

int nrh_f3d_lncc_value(nrh_f3d *h, double *value) {
    return guard([&] {
        NRH_REQUIRE(h && value, NRH_ERR_ARG, "null argument");
        ensure_lncc(h);
        op_lncc_value(h, h->lnccScalars.p + 2);
        *value = read_scalar(h, h->lnccScalars.p + 2);
    });
}

int nrh_f3d_lncc_voxel_gradient(nrh_f3d *h, double weight) {
    return guard([&] {
        NRH_REQUIRE(h, NRH_ERR_ARG, "null handle");
        op_lncc_voxel_gradient(h, weight);
    });
}

}  // extern "C"
