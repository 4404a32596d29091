// Bending-energy building blocks shared by the per-operation kernels (f3d_ops.hip) and the
// fused exact-pass launch (f3d_fused.hip): the 27-point basis table, the node bricks and the
// value block body.
#pragma once

#include "device.hpp"

namespace nrh {

struct BeBasis {
    float b[6][27];  // xx, yy, zz, xy, yz, xz
};

// Bending energy on bricks of 8 x 8 x 4 nodes (one workgroup, one node per thread): the control
// points of the brick and its halo are staged in LDS once, so every 27-neighbour stencil reads
// LDS instead of global memory.
constexpr int BBX = 8, BBY = 8, BBZ = 4;

struct BrickGrid {
    int bx, by, bz;  // bricks per axis
};

__device__ __forceinline__ void brick_origin(const BrickGrid& g, int b, int& x0, int& y0, int& z0) {
    x0 = (b % g.bx) * BBX;
    y0 = ((b / g.bx) % g.by) * BBY;
    z0 = (b / (g.bx * g.by)) * BBZ;
}

// reg_spline_approxBendingEnergyValue3D (reg-lib/cpu/_reg_localTrans_regul.cpp:73-153) over the
// interior nodes of the brick; FP64 block partials, summed in block order by the last block.
__device__ __forceinline__ void
be_value_block(const float *__restrict__ cpp, int gnx, int gny, int gnz, BrickGrid bg, BeBasis B,
           double *__restrict__ partials, uint32_t *__restrict__ ticket, double scale, double *__restrict__ out, int bid, int nbk) {
#pragma clang fp contract(off)
    constexpr int HX = BBX + 2, HY = BBY + 2, HZ = BBZ + 2, HN = HX * HY * HZ;
    __shared__ float sC[3][HN];  // deformation values of the brick and a one-node halo
    __shared__ float sB[27][8];  // basis values per neighbour (xx, yy, zz, xy, yz, xz), read as broadcasts
    __shared__ double red[16];
    __shared__ bool last;
    const size_t G = (size_t)gnx * gny * gnz;
    int x0, y0, z0;
    brick_origin(bg, bid, x0, y0, z0);
    if (threadIdx.x < 27 * 6) sB[threadIdx.x / 6][threadIdx.x % 6] = B.b[threadIdx.x % 6][threadIdx.x / 6];
    {
        constexpr int kIter = (3 * HN + 255) / 256;
        float v[kIter];
#pragma unroll
        for (int u = 0; u < kIter; ++u) {
            // clamped, unconditional loads keep v in registers
            const int i = min((int)threadIdx.x + u * 256, 3 * HN - 1);
            const int c = i / HN, r = i % HN, hx = r % HX, hy = (r / HX) % HY, hz = r / (HX * HY);
            const int X = min(max(x0 - 1 + hx, 0), gnx - 1), Y = min(max(y0 - 1 + hy, 0), gny - 1),
                      Z = min(max(z0 - 1 + hz, 0), gnz - 1);
            v[u] = cpp[c * G + ((size_t)Z * gny + Y) * gnx + X];
        }
#pragma unroll
        for (int u = 0; u < kIter; ++u) {
            const int i = threadIdx.x + u * 256;
            if (i < 3 * HN) sC[i / HN][i % HN] = v[u];
        }
    }
    __syncthreads();
    const int tx = threadIdx.x % BBX, ty = (threadIdx.x / BBX) % BBY, tz = threadIdx.x / (BBX * BBY);
    const int x = x0 + tx, y = y0 + ty, z = z0 + tz;
    double s = 0;
    if (x > 0 && y > 0 && z > 0 && x < gnx - 1 && y < gny - 1 && z < gnz - 1) {
        float d[3][6] = {};
        int i = 0;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int b = 0; b < 3; ++b)
#pragma unroll
                for (int a = 0; a < 3; ++a, ++i) {
                    const int h = ((tz + c) * HY + ty + b) * HX + tx + a;
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
        s = (double)(d[0][0] * d[0][0] + d[0][1] * d[0][1] + d[0][2] * d[0][2]) +
            2.0 * (double)(d[0][3] * d[0][3] + d[0][4] * d[0][4] + d[0][5] * d[0][5]);
        for (int k = 1; k < 3; ++k) {
            s += (double)(d[k][0] * d[k][0]);
            s += (double)(d[k][1] * d[k][1]);
            s += (double)(d[k][2] * d[k][2]);
            s += 2.0 * (double)(d[k][3] * d[k][3] + d[k][4] * d[k][4] + d[k][5] * d[k][5]);
        }
    }
    s = block_sum_double(s, red);
    // the last block to finish sums the per-block partials in block order
    if (threadIdx.x == 0) {
        partials[bid] = s;
        __threadfence();
        last = atomicAdd(ticket, 1u) == nbk - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t = 0;
    for (int i = threadIdx.x; i < (int)nbk; i += blockDim.x) t += partials[i];
    t = block_sum_double(t, red);
    if (threadIdx.x == 0) {
        *out = t * scale;
        *ticket = 0u;
    }
}

// The basis table (the reference's 6-decimal literals) and the brick grid of a control-point grid.
BeBasis be_basis();
inline BrickGrid be_brick_grid(int gnx, int gny, int gnz) {
    BrickGrid g;
    g.bx = (gnx + BBX - 1) / BBX;
    g.by = (gny + BBY - 1) / BBY;
    g.bz = (gnz + BBZ - 1) / BBZ;
    return g;
}

}  // namespace nrh
