// Velocity-grid operations of symmetric diffeomorphic F3D (reg_f3d2) for gfx950.
//
// A velocity grid holds cubic B-spline control points (positions in mm) whose displacement is a
// stationary velocity field; its deformation is the exponential map computed by scaling and
// squaring: the flow (the grid evaluated at every voxel's position), its displacement divided by
// 2^s, then s self-compositions with trilinear interpolation (reg-lib/cpu/_reg_localTrans.cpp:
// 3465-3683).  ExponentiateGradient carries the voxel-based gradient along the intermediate
// fields of the opposite direction's grid (reg-lib/Compute.cpp:283-326).
//
// Every kernel is one thread per voxel (or node) and follows the reference's scalar arithmetic
// (float where it computes in float, double where reg_tools promotes, no fused multiply-adds),
// so the results equal the reference CPU build bit for bit.  The flow sums its 64 B-spline terms
// in the reference's order; the compositions and the gradient resampling gather 8 taps from the
// previous field, which stays in L2 for the small displacements the squaring produces.
// Grids with affine extensions (a reg_f3d2 run given an input affine) are not handled here.

#include "device.hpp"
#include "f3d_handle.hpp"
#include "host.hpp"
#include "m33.hpp"
#include <cmath>

namespace nrh {

namespace {

constexpr int kBlock = 256;

struct FieldGeom {
    int nx, ny, nz;
    size_t n;
    Affine xyz, ijk;
};

FieldGeom field_geom(const nrh_geom& g) {
    FieldGeom f;
    f.nx = g.nx; f.ny = g.ny; f.nz = g.nz;
    f.n = voxels(g);
    f.xyz = to_affine(g.xyz);
    f.ijk = to_affine(g.ijk);
    return f;
}

struct VelGrid {
    const float *cp;
    int gnx, gny, gnz;
    size_t G;
    Affine xyz, ijk;
    int order;  // NRH_FIELD_ORDER_*: the summation order of the reference build reproduced
};

// Affine parts of a velocity grid on one field geometry.  pre: the grid's extension 0, applied to
// the identity before the spline (reg-lib/cpu/_reg_localTrans.cpp:1478-1487), removed from the
// flow as the affine-only field A S x (:3508-3515) and added back after the squaring (:3569-3573);
// post: extension 1, composed last (:3576-3577); sub: the opposite content's transformation,
// whose displacement the gradient exponentiation removes from each intermediate field
// (reg-lib/Compute.cpp:303-316).  Rows 0-2 of each matrix.
struct VelAff {
    int pre, post, sub;
    float A[12];   // extension 0
    float AS[12];  // extension 0 times the field's sform
    float B[12];   // extension 1
    float TS[12];  // the transformation times the field's sform
};

// Mat44Mul<double> (reg-lib/cpu/Maths.hpp:272-286) of a position, rounded to float as
// reg_affine_deformationField3D stores it (reg-lib/cpu/_reg_globalTrans.cpp:74-133).
__device__ __forceinline__ float3 mul_d(const float *m, double x, double y, double z) {
#pragma clang fp contract(off)
    float3 r;
    r.x = (float)((double)m[0] * x + (double)m[1] * y + (double)m[2] * z + (double)m[3]);
    r.y = (float)((double)m[4] * x + (double)m[5] * y + (double)m[6] * z + (double)m[7]);
    r.z = (float)((double)m[8] * x + (double)m[9] * y + (double)m[10] * z + (double)m[11]);
    return r;
}

// voxel -> mm in float, as reg_getDeformationFromDisplacement_3D (reg-lib/cpu/_reg_tools.cpp:
// 2107-2150) adds it
__device__ __forceinline__ float3 id_pos(const Affine& m, int x, int y, int z) {
#pragma clang fp contract(off)
    float3 r;
    r.x = m.m[0] * (float)x + m.m[1] * (float)y + m.m[2] * (float)z + m.m[3];
    r.y = m.m[4] * (float)x + m.m[5] * (float)y + m.m[6] * (float)z + m.m[7];
    r.z = m.m[8] * (float)x + m.m[9] * (float)y + m.m[10] * (float)z + m.m[11];
    return r;
}

// mm -> voxel in float (reg-lib/cpu/_reg_localTrans.cpp:940-952, 2384-2396)
__device__ __forceinline__ float3 to_vox(const Affine& m, float3 p) {
#pragma clang fp contract(off)
    float3 r;
    r.x = m.m[0] * p.x + m.m[1] * p.y + m.m[2] * p.z + m.m[3];
    r.y = m.m[4] * p.x + m.m[5] * p.y + m.m[6] * p.z + m.m[7];
    r.z = m.m[8] * p.x + m.m[9] * p.y + m.m[10] * p.z + m.m[11];
    return r;
}

__device__ __forceinline__ int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

// Workgroups are dealt round-robin to the 8 XCDs (workgroup b runs on XCD b % 8).  This maps
// them so that each XCD takes one contiguous run of voxel blocks (a slab of rows and slices):
// the rows a block gathers from its neighbours then sit in the same XCD's L2.  A bijection of
// [0, nb) when nb is a multiple of 8 (launches round their grids up).
__device__ __forceinline__ size_t xcd_block(unsigned b, unsigned nb) {
    return (size_t)(b & 7u) * (nb >> 3) + (b >> 3);
}

// get_SlidedValues 3-D for positions (reg-lib/cpu/_reg_splineBasis.cpp:496-551): the value at the
// clamped index plus the mm offset of the clamped steps.
__device__ __forceinline__ float3 slided(const float *__restrict__ d, size_t n, int nx, int ny, int nz,
                                         const Affine& xyz, int x, int y, int z) {
#pragma clang fp contract(off)
    const int cx = clampi(x, nx), cy = clampi(y, ny), cz = clampi(z, nz);
    const float sx = (float)(x - cx), sy = (float)(y - cy), sz = (float)(z - cz);
    const size_t idx = ((size_t)cz * ny + cy) * nx + cx;
    float3 r;
    r.x = d[idx] + (sx * xyz.m[0] + sy * xyz.m[1] + sz * xyz.m[2]);
    r.y = d[idx + n] + (sx * xyz.m[4] + sy * xyz.m[5] + sz * xyz.m[6]);
    r.z = d[idx + 2 * n] + (sx * xyz.m[8] + sy * xyz.m[9] + sz * xyz.m[10]);
    return r;
}

// Order-preserving integer image of a float, for atomic min / max.
__device__ __forceinline__ uint32_t float_key(float v) {
    const uint32_t u = __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// The velocity grid evaluated at voxel i's position: the composition branch of
// reg_cubic_spline_getDeformationField3D (reg-lib/cpu/_reg_localTrans.cpp:903-1079, scalar
// order) applied to the identity field, as reg_spline_getFlowFieldFromVelocityGrid (:3465-3491)
// does.  Returns the displacement flow - x and the position x.  When every lane of the wave has
// its 4x4x4 support inside the grid, each slab of 16 nodes is loaded at once (48 loads in flight)
// before its terms are summed in the reference's order.  With an extension 0 the spline is
// evaluated at A x and the displacement is taken from the affine-only field A S i in double.
__device__ __forceinline__ void flow_at(const VelGrid& g, const FieldGeom& f, const VelAff& a, size_t i, bool active,
                                        float3& disp, float3& pos) {
#pragma clang fp contract(off)
    const size_t ii = active ? i : 0;
    int x, y, z;
    index_xyz(ii, f.nx, f.ny, x, y, z);
    const float3 p0 = id_pos(f.xyz, x, y, z);
    const float3 p = a.pre ? mul_d(a.A, p0.x, p0.y, p0.z) : p0;
    const float3 v = to_vox(g.ijk, p);
    int pre[3];
    float b[3][4];
    const float vc[3] = { v.x, v.y, v.z };
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        pre[r] = floor_int(vc[r]);
        float t = vc[r] - (float)pre[r];
        pre[r]--;
        if (t < 0) t = 0;
        bspline_basis(t, b[r]);
    }
    const bool inside = pre[0] >= 0 && pre[0] + 3 < g.gnx && pre[1] >= 0 && pre[1] + 3 < g.gny && pre[2] >= 0 &&
                        pre[2] + 3 < g.gnz;
    // scalar build: one sequential sum with weights (x y) z; SSE build (reg-lib/cpu/
    // _reg_localTrans.cpp:1025-1052): four lanes (one per x tap) with weights (y z) x, reduced
    // left to right
    const bool sse = g.order == NRH_FIELD_ORDER_SSE;
    float r0 = 0.f, r1 = 0.f, r2 = 0.f;
    float l0[4] = { 0.f, 0.f, 0.f, 0.f }, l1[4] = { 0.f, 0.f, 0.f, 0.f }, l2[4] = { 0.f, 0.f, 0.f, 0.f };
    if (__all(inside || !active)) {
        const size_t sy = g.gnx, sz = (size_t)g.gnx * g.gny;
        const size_t base = active ? ((size_t)pre[2] * g.gny + pre[1]) * g.gnx + pre[0] : 0;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float cx[16], cy[16], cz[16];
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const size_t idx = base + c * sz + (k >> 2) * sy + (k & 3);
                cx[k] = g.cp[idx];
                cy[k] = g.cp[idx + g.G];
                cz[k] = g.cp[idx + 2 * g.G];
            }
            if (sse) {
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const float w = (b[1][k >> 2] * b[2][c]) * b[0][k & 3];
                    l0[k & 3] = w * cx[k] + l0[k & 3];
                    l1[k & 3] = w * cy[k] + l1[k & 3];
                    l2[k & 3] = w * cz[k] + l2[k & 3];
                }
            } else {
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const float w = b[0][k & 3] * b[1][k >> 2] * b[2][c];
                    r0 += cx[k] * w;
                    r1 += cy[k] * w;
                    r2 += cz[k] * w;
                }
            }
        }
    } else if (active) {
        for (int c = 0; c < 4; ++c) {
            const int Z = pre[2] + c;
            // get_GridValues (reg-lib/cpu/_reg_splineBasis.cpp:602-665): once a row of a slice
            // falls outside the grid, the rest of that slice goes through the sliding path
            bool out = !(Z > -1 && Z < g.gnz);
            for (int bb = 0; bb < 4; ++bb) {
                const int Y = pre[1] + bb;
                if (!(Y > -1 && Y < g.gny && !out)) out = true;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    const int X = pre[0] + a;
                    float3 cv;
                    if (X > -1 && X < g.gnx && !out) {
                        const size_t idx = ((size_t)Z * g.gny + Y) * g.gnx + X;
                        cv.x = g.cp[idx];
                        cv.y = g.cp[idx + g.G];
                        cv.z = g.cp[idx + 2 * g.G];
                    } else {
                        cv = slided(g.cp, g.G, g.gnx, g.gny, g.gnz, g.xyz, X, Y, Z);
                    }
                    if (sse) {
                        const float w = (b[1][bb] * b[2][c]) * b[0][a];
                        l0[a] = w * cv.x + l0[a];
                        l1[a] = w * cv.y + l1[a];
                        l2[a] = w * cv.z + l2[a];
                    } else {
                        const float w = b[0][a] * b[1][bb] * b[2][c];
                        r0 += cv.x * w;
                        r1 += cv.y * w;
                        r2 += cv.z * w;
                    }
                }
            }
        }
    }
    if (sse) {
        r0 = ((l0[0] + l0[1]) + l0[2]) + l0[3];
        r1 = ((l1[0] + l1[1]) + l1[2]) + l1[3];
        r2 = ((l2[0] + l2[1]) + l2[2]) + l2[3];
    }
    if (a.pre) {
        const float3 ao = mul_d(a.AS, (double)x, (double)y, (double)z);
        disp = make_float3((float)((double)r0 - (double)ao.x), (float)((double)r1 - (double)ao.y),
                           (float)((double)r2 - (double)ao.z));
    } else {
        disp = make_float3(r0 - p.x, r1 - p.y, r2 - p.z);
    }
    pos = p0;
}

// The scaled flow, the first field of the squaring (reg-lib/cpu/_reg_localTrans.cpp:3546-3561):
// (flow - x) / scaling + x; a persistent grid-stride loop as k_vel_flow_disp.
__global__ void __launch_bounds__(kBlock) k_vel_flow_scaled(VelGrid g, FieldGeom f, VelAff a, double scaling,
                                                            float *__restrict__ out) {
#pragma clang fp contract(off)
    // scaling is +-2^s: the double quotient equals the product with its exact reciprocal
    const double inv = 1.0 / scaling;
    const size_t stride = (size_t)gridDim.x * kBlock;
    const size_t rounds = (f.n + stride - 1) / stride;
    for (size_t r = 0; r < rounds; ++r) {
        const size_t i = r * stride + (size_t)blockIdx.x * kBlock + threadIdx.x;
        const bool active = i < f.n;
        float3 d, p;
        flow_at(g, f, a, i, active, d, p);
        if (active) {
            out[i] = (float)((double)d.x * inv) + p.x;
            out[i + f.n] = (float)((double)d.y * inv) + p.y;
            out[i + 2 * f.n] = (float)((double)d.z * inv) + p.z;
        }
    }
}

// The flow's displacement flow - x, with its extremes folded into keys[0] (min) and keys[1]
// (max): grid-stride loop, one block-wide reduction and one atomic pair per block.
__global__ void __launch_bounds__(kBlock) k_vel_flow_disp(VelGrid g, FieldGeom f, VelAff a,
                                                          float *__restrict__ out, uint32_t *__restrict__ keys) {
#pragma clang fp contract(off)
    __shared__ float sLo[kBlock / 64], sHi[kBlock / 64];
    float lo = INFINITY, hi = -INFINITY;
    const size_t stride = (size_t)gridDim.x * kBlock;
    const size_t rounds = (f.n + stride - 1) / stride;
    for (size_t r = 0; r < rounds; ++r) {
        const size_t i = r * stride + (size_t)blockIdx.x * kBlock + threadIdx.x;
        const bool active = i < f.n;
        float3 d, p;
        flow_at(g, f, a, i, active, d, p);
        if (active) {
            out[i] = d.x;
            out[i + f.n] = d.y;
            out[i + 2 * f.n] = d.z;
            lo = fminf(lo, fminf(d.x, fminf(d.y, d.z)));
            hi = fmaxf(hi, fmaxf(d.x, fmaxf(d.y, d.z)));
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fminf(lo, __shfl_xor(lo, off, 64));
        hi = fmaxf(hi, __shfl_xor(hi, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        sLo[threadIdx.x >> 6] = lo;
        sHi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) {
            lo = fminf(lo, sLo[w]);
            hi = fmaxf(hi, sHi[w]);
        }
        atomicMin(&keys[0], float_key(lo));
        atomicMax(&keys[1], float_key(hi));
    }
}

// Displacement d -> d / scaling + x (reg_tools_divideValueToImage in double, then
// reg_getDeformationFromDisplacement in float).
__global__ void __launch_bounds__(kBlock) k_vel_scale(const float *disp, FieldGeom f, double scaling, float *out) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= f.n) return;
    int x, y, z;
    index_xyz(i, f.nx, f.ny, x, y, z);
    const float3 p = id_pos(f.xyz, x, y, z);
    const double inv = 1.0 / scaling;  // exact: scaling is +-2^s
    out[i] = (float)((double)disp[i] * inv) + p.x;
    out[i + f.n] = (float)((double)disp[i + f.n] * inv) + p.y;
    out[i + 2 * f.n] = (float)((double)disp[i + 2 * f.n] * inv) + p.z;
}

// The affine parts put back on `count` fields after the squaring: with an extension 0 the
// displacement (float, reg_getDisplacementFromDeformation) plus the affine-only field (double
// sum, reg_tools_addImageToImage); then extension 1 composed (Mat44Mul<double>); then, for the
// gradient exponentiation, the displacement of the opposite content's transformation
// (T S i in double, minus x in float) subtracted in double.
__global__ void __launch_bounds__(kBlock) k_vel_affine_parts(float *fields, int count, FieldGeom f, VelAff a) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= f.n) return;
    int x, y, z;
    index_xyz(i, f.nx, f.ny, x, y, z);
    const float3 p0 = id_pos(f.xyz, x, y, z);
    const float3 ao = a.pre ? mul_d(a.AS, (double)x, (double)y, (double)z) : make_float3(0.f, 0.f, 0.f);
    float3 td = make_float3(0.f, 0.f, 0.f);
    if (a.sub) {
        const float3 t = mul_d(a.TS, (double)x, (double)y, (double)z);
        td = make_float3(t.x - p0.x, t.y - p0.y, t.z - p0.z);
    }
    for (int k = 0; k < count; ++k) {
        float *F = fields + (size_t)k * 3 * f.n;
        float3 v = make_float3(F[i], F[i + f.n], F[i + 2 * f.n]);
        if (a.pre)
            v = make_float3((float)((double)(v.x - p0.x) + (double)ao.x), (float)((double)(v.y - p0.y) + (double)ao.y),
                            (float)((double)(v.z - p0.z) + (double)ao.z));
        if (a.post) v = mul_d(a.B, v.x, v.y, v.z);
        if (a.sub)
            v = make_float3((float)((double)v.x - (double)td.x), (float)((double)v.y - (double)td.y),
                            (float)((double)v.z - (double)td.z));
        F[i] = v.x;
        F[i + f.n] = v.y;
        F[i + 2 * f.n] = v.z;
    }
}

// reg_defField_compose3D without a mask (reg-lib/cpu/_reg_localTrans.cpp:2333-2457): out[i] is
// the trilinear interpolation of field D (geometry d) at position pos[i], with sliding outside
// D.  out may alias pos (each thread reads only its own position) but not D.
__global__ void __launch_bounds__(kBlock) k_compose(const float *__restrict__ D, FieldGeom d, const float *pos,
                                                    size_t m, float *out) {
#pragma clang fp contract(off)
    const size_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    const bool active = i < m;
    const size_t ii = active ? i : 0;
    const float3 real = make_float3(pos[ii], pos[ii + m], pos[ii + 2 * m]);
    const float3 v = to_vox(d.ijk, real);
    const int px = floor_int(v.x), py = floor_int(v.y), pz = floor_int(v.z);
    const float rx1 = v.x - (float)px, ry1 = v.y - (float)py, rz1 = v.z - (float)pz;
    const float rx[2] = { 1.f - rx1, rx1 }, ry[2] = { 1.f - ry1, ry1 }, rz[2] = { 1.f - rz1, rz1 };
    const bool inside = px > -1 && px + 1 < d.nx && py > -1 && py + 1 < d.ny && pz > -1 && pz + 1 < d.nz;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f;
    if (__all(inside || !active) && d.nx > 1 && d.ny > 1 && d.nz > 1) {
        // every tap inside: all 24 loads issued before the sums (reference order kept)
        const size_t base = active ? ((size_t)pz * d.ny + py) * d.nx + px : 0;
        const size_t sy = d.nx, sz = (size_t)d.nx * d.ny;
        float t0[8], t1[8], t2[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const size_t idx = base + (k >> 2) * sz + ((k >> 1) & 1) * sy + (k & 1);
            t0[k] = D[idx];
            t1[k] = D[idx + d.n];
            t2[k] = D[idx + 2 * d.n];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float w = rx[k & 1] * (ry[(k >> 1) & 1] * rz[k >> 2]);
            a0 += t0[k] * w;
            a1 += t1[k] * w;
            a2 += t2[k] * w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int Z = pz + c;
            const bool inZ = Z > -1 && Z < d.nz;
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int Y = py + b;
                const float tb = ry[b] * rz[c];
                const bool inY = Y > -1 && Y < d.ny;
#pragma unroll
                for (int a = 0; a < 2; ++a) {
                    const int X = px + a;
                    float3 dv;
                    if (X > -1 && X < d.nx && inY && inZ) {
                        const size_t idx = ((size_t)Z * d.ny + Y) * d.nx + X;
                        dv = make_float3(D[idx], D[idx + d.n], D[idx + 2 * d.n]);
                    } else {
                        dv = slided(D, d.n, d.nx, d.ny, d.nz, d.xyz, X, Y, Z);
                    }
                    const float w = rx[a] * tb;
                    a0 += dv.x * w;
                    a1 += dv.y * w;
                    a2 += dv.z * w;
                }
            }
        }
    }
    if (!active) return;
    out[i] = a0;
    out[i + m] = a1;
    out[i + 2 * m] = a2;
}

struct ResampleGradArgs {
    const float *grad;      // vector image to resample (planar [3][gn])
    int gnx, gny, gnz;
    size_t gn;
    Affine gradIjk;         // its mm -> voxel
    const float *def;       // positions (planar [3][n]) on the output geometry
    int nx, ny, nz;
    size_t n;
    float reo[9];           // inverse of the polar rotation of the field's sform
    float sp[3];            // real spacing of the output
    float pad;
    const float *accum;     // optional: result + accum[i] (double sum)
    double divisor;         // optional (> 0): then / divisor
    float *out;             // must not alias grad or def
    float rsp[3];           // 1 / sp rounded to float (k_resample_gradient32)
    double invDivisor;      // 1 / divisor, exact (a power of two; k_resample_gradient32)
};

// reg_trilinearResampleGradient (reg-lib/cpu/_reg_resampling.cpp:1547-1766): the vector image
// sampled at each position (Floor / Ceil taps, padding outside) times the field's forward
// difference Jacobian, rotation removed and divided by the real spacing; at the last index of
// an axis the difference steps back one voxel, and so do the other two axes' differences there.
__global__ void __launch_bounds__(kBlock) k_resample_gradient(ResampleGradArgs r) {
#pragma clang fp contract(off)
    const size_t i = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (i >= r.n) return;
    int x, y, z;
    index_xyz(i, r.nx, r.ny, x, y, z);
    const float3 fc = to_vox(r.gradIjk, make_float3(r.def[i], r.def[i + r.n], r.def[i + 2 * r.n]));
    const float fcv[3] = { fc.x, fc.y, fc.z };
    int lo[3], hi[3];
    float bw[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = floor_int(fcv[k]);
        const int t = (int)fcv[k];
        hi[k] = t + (fcv[k] > (float)t);
        bw[k][1] = fabsf(fcv[k] - (float)lo[k]);
        bw[k][0] = 1.f - bw[k][1];
    }
    const int gdim[3] = { r.gnx, r.gny, r.gnz };
    float vx = 0.f, vy = 0.f, vz = 0.f;
    const bool inside = lo[0] > -1 && hi[0] < gdim[0] && lo[1] > -1 && hi[1] < gdim[1] && lo[2] > -1 && hi[2] < gdim[2];
    if (__all(inside)) {
        // every tap inside: the 24 loads issued together, summed in the reference's order
        float t0[8], t1[8], t2[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int X = (k & 1) ? hi[0] : lo[0], Y = ((k >> 1) & 1) ? hi[1] : lo[1], Z = (k >> 2) ? hi[2] : lo[2];
            const size_t idx = ((size_t)Z * r.gny + Y) * r.gnx + X;
            t0[k] = r.grad[idx];
            t1[k] = r.grad[idx + r.gn];
            t2[k] = r.grad[idx + 2 * r.gn];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float w = bw[0][k & 1] * bw[1][(k >> 1) & 1] * bw[2][k >> 2];
            vx += t0[k] * w;
            vy += t1[k] * w;
            vz += t2[k] * w;
        }
    } else {
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        const int Z = c ? hi[2] : lo[2];
        if (Z > -1 && Z < gdim[2]) {
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const int Y = b ? hi[1] : lo[1];
                if (Y > -1 && Y < gdim[1]) {
#pragma unroll
                    for (int a = 0; a < 2; ++a) {
                        const int X = a ? hi[0] : lo[0];
                        const float w = bw[0][a] * bw[1][b] * bw[2][c];
                        if (X > -1 && X < gdim[0]) {
                            const size_t idx = ((size_t)Z * r.gny + Y) * r.gnx + X;
                            vx += r.grad[idx] * w;
                            vy += r.grad[idx + r.gn] * w;
                            vz += r.grad[idx + 2 * r.gn] * w;
                        } else {
                            vx += r.pad * w;
                            vy += r.pad * w;
                            vz += r.pad * w;
                        }
                    }
                } else {
                    const float w = r.pad * bw[1][b] * bw[2][c];
                    vx += w;
                    vy += w;
                    vz += w;
                }
            }
        } else {
            const float w = r.pad * bw[2][c];
            vx += w;
            vy += w;
            vz += w;
        }
    }
    }
    // Forward-difference Jacobian over the eight (c, b, a) pairs of the reference's loop.  Only
    // four pairs carry a non-zero weight: the base pair (a*, b*, c*), where a* = 1 on the last
    // index of the x axis (the difference steps back) and 0 elsewhere, and the base with one of a,
    // b, c flipped.  The other pairs add zeros, which leave the sums unchanged, so the four pairs
    // are visited in the reference's order (ascending c*4 + b*2 + a) with their 12 loads issued
    // together.
    const int ax = x == r.nx - 1, ay = y == r.ny - 1, az = z == r.nz - 1;
    const int base = az * 4 + ay * 2 + ax;
    int pr[4] = { base, base ^ 1, base ^ 2, base ^ 4 };
    // sorting network on four keys
    auto cswap = [](int& u, int& v) {
        const int lo_ = u < v ? u : v, hi_ = u < v ? v : u;
        u = lo_;
        v = hi_;
    };
    cswap(pr[0], pr[1]);
    cswap(pr[2], pr[3]);
    cswap(pr[0], pr[2]);
    cswap(pr[1], pr[3]);
    cswap(pr[1], pr[2]);
    float jw[4][3], jd[4][3];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int a = pr[q] & 1, b = (pr[q] >> 1) & 1, c = pr[q] >> 2;
        // reference weights: basis {1, 0}, derivative {-1, 1}; on the last index the step goes back
        // (index - 1 for the second tap), the basis becomes |basis - 1| and the derivative flips
        const float bx0 = ax ? (float)a : (float)(1 - a), by0 = ay ? (float)b : (float)(1 - b),
                    bz0 = az ? (float)c : (float)(1 - c);
        const float bx1 = ax ? (a ? -1.f : 1.f) : (a ? 1.f : -1.f), by1 = ay ? (b ? -1.f : 1.f) : (b ? 1.f : -1.f),
                    bz1 = az ? (c ? -1.f : 1.f) : (c ? 1.f : -1.f);
        jw[q][0] = bx1 * by0 * bz0;
        jw[q][1] = bx0 * by1 * bz0;
        jw[q][2] = bx0 * by0 * bz1;
        const int X = x + a - (ax && a ? 2 : 0), Y = y + b - (ay && b ? 2 : 0), Z = z + c - (az && c ? 2 : 0);
        // fields with one voxel along an axis would step to index -1: clamp the read (its weight
        // is then zero on that axis in the reference as well)
        const size_t idx = ((size_t)clampi(Z, r.nz) * r.ny + clampi(Y, r.ny)) * r.nx + clampi(X, r.nx);
        jd[q][0] = r.def[idx];
        jd[q][1] = r.def[idx + r.n];
        jd[q][2] = r.def[idx + 2 * r.n];
    }
    float J[3][3] = {};
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            J[m][0] += jw[q][0] * jd[q][m];
            J[m][1] += jw[q][1] * jd[q][m];
            J[m][2] += jw[q][2] * jd[q][m];
        }
    float R[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
            R[a][b] = (r.reo[a * 3] * J[0][b] + r.reo[a * 3 + 1] * J[1][b] + r.reo[a * 3 + 2] * J[2][b]) / r.sp[b];
    float o[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) o[a] = R[a][0] * vx + R[a][1] * vy + R[a][2] * vz;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float v = o[a];
        if (r.accum) v = (float)((double)v + (double)r.accum[i + a * r.n]);
        if (r.divisor > 0.0) v = (float)((double)v / r.divisor);
        r.out[i + a * r.n] = v;
    }
}

// The same operation for arrays below 4 GiB and fields with at least two voxels along each axis,
// with less arithmetic for identical results:
// * loads address 32-bit byte offsets from the three component planes' base pointers;
// * the Jacobian's sums over the four weighted pairs reduce to forward differences at the corner
//   (x - ax, y - ay, z - az), ax = 1 on the last index of x (likewise y, z): every other term is
//   a product with zero, which leaves the sums' values unchanged (the sign of an exact zero
//   difference aside);
// * a / sp is q + (a - q sp) / sp with q = a rsp and the residual taken by fused multiply-adds,
//   rsp = 1 / sp rounded to float: Markstein's correction, correctly rounded;
// * the division by the power of two 2^s is the product with its exact reciprocal.
__device__ __forceinline__ float ld32(const float *base, uint32_t byteOff) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + byteOff);
}

__global__ void __launch_bounds__(kBlock) k_resample_gradient32(ResampleGradArgs r) {
#pragma clang fp contract(off)
    const size_t ib = xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    if (ib >= r.n) return;
    const uint32_t i = (uint32_t)ib;
    int x, y, z;
    index_xyz(i, r.nx, r.ny, x, y, z);
    const float *D0 = r.def, *D1 = r.def + r.n, *D2 = r.def + 2 * r.n;
    const float *G0 = r.grad, *G1 = r.grad + r.gn, *G2 = r.grad + 2 * r.gn;
    const uint32_t io = i * 4u;
    const float3 pos = make_float3(ld32(D0, io), ld32(D1, io), ld32(D2, io));
    const float3 fc = to_vox(r.gradIjk, pos);
    const float fcv[3] = { fc.x, fc.y, fc.z };
    int lo[3], hi[3];
    float bw[3][2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        lo[k] = floor_int(fcv[k]);
        const int t = (int)fcv[k];
        hi[k] = t + (fcv[k] > (float)t);
        bw[k][1] = fabsf(fcv[k] - (float)lo[k]);
        bw[k][0] = 1.f - bw[k][1];
    }
    const int gdim[3] = { r.gnx, r.gny, r.gnz };
    float vx = 0.f, vy = 0.f, vz = 0.f;
    const bool inside = lo[0] > -1 && hi[0] < gdim[0] && lo[1] > -1 && hi[1] < gdim[1] && lo[2] > -1 && hi[2] < gdim[2];
    if (__all(inside)) {
        const uint32_t sy = (uint32_t)r.gnx, sz = (uint32_t)r.gnx * (uint32_t)r.gny;
        const uint32_t xo[2] = { (uint32_t)lo[0], (uint32_t)hi[0] };
        const uint32_t yo[2] = { (uint32_t)lo[1] * sy, (uint32_t)hi[1] * sy };
        const uint32_t zo[2] = { (uint32_t)lo[2] * sz, (uint32_t)hi[2] * sz };
        float t0[8], t1[8], t2[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const uint32_t o = (zo[k >> 2] + yo[(k >> 1) & 1] + xo[k & 1]) * 4u;
            t0[k] = ld32(G0, o);
            t1[k] = ld32(G1, o);
            t2[k] = ld32(G2, o);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const float w = bw[0][k & 1] * bw[1][(k >> 1) & 1] * bw[2][k >> 2];
            vx += t0[k] * w;
            vy += t1[k] * w;
            vz += t2[k] * w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const int Z = c ? hi[2] : lo[2];
            if (Z > -1 && Z < gdim[2]) {
#pragma unroll
                for (int b = 0; b < 2; ++b) {
                    const int Y = b ? hi[1] : lo[1];
                    if (Y > -1 && Y < gdim[1]) {
#pragma unroll
                        for (int a = 0; a < 2; ++a) {
                            const int X = a ? hi[0] : lo[0];
                            const float w = bw[0][a] * bw[1][b] * bw[2][c];
                            if (X > -1 && X < gdim[0]) {
                                const uint32_t o = (((uint32_t)Z * (uint32_t)r.gny + (uint32_t)Y) * (uint32_t)r.gnx +
                                                    (uint32_t)X) * 4u;
                                vx += ld32(G0, o) * w;
                                vy += ld32(G1, o) * w;
                                vz += ld32(G2, o) * w;
                            } else {
                                vx += r.pad * w;
                                vy += r.pad * w;
                                vz += r.pad * w;
                            }
                        }
                    } else {
                        const float w = r.pad * bw[1][b] * bw[2][c];
                        vx += w;
                        vy += w;
                        vz += w;
                    }
                }
            } else {
                const float w = r.pad * bw[2][c];
                vx += w;
                vy += w;
                vz += w;
            }
        }
    }
    const int ax = x == r.nx - 1, ay = y == r.ny - 1, az = z == r.nz - 1;
    const uint32_t sx = 4u, sy = (uint32_t)r.nx * 4u, sz = (uint32_t)r.nx * (uint32_t)r.ny * 4u;
    float3 d0 = pos;
    uint32_t o0 = io;
    if (ax | ay | az) {
        o0 = io - (ax ? sx : 0u) - (ay ? sy : 0u) - (az ? sz : 0u);
        d0 = make_float3(ld32(D0, o0), ld32(D1, o0), ld32(D2, o0));
    }
    const float dxv[3] = { ld32(D0, o0 + sx), ld32(D1, o0 + sx), ld32(D2, o0 + sx) };
    const float dyv[3] = { ld32(D0, o0 + sy), ld32(D1, o0 + sy), ld32(D2, o0 + sy) };
    const float dzv[3] = { ld32(D0, o0 + sz), ld32(D1, o0 + sz), ld32(D2, o0 + sz) };
    const float d0v[3] = { d0.x, d0.y, d0.z };
    float J[3][3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
        J[m][0] = dxv[m] - d0v[m];
        J[m][1] = dyv[m] - d0v[m];
        J[m][2] = dzv[m] - d0v[m];
    }
    float R[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float num = r.reo[a * 3] * J[0][b] + r.reo[a * 3 + 1] * J[1][b] + r.reo[a * 3 + 2] * J[2][b];
            const float q = num * r.rsp[b];
            R[a][b] = fmaf(fmaf(-q, r.sp[b], num), r.rsp[b], q);
        }
    float o[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) o[a] = R[a][0] * vx + R[a][1] * vy + R[a][2] * vz;
    float *O[3] = { r.out, r.out + r.n, r.out + 2 * r.n };
    const float *A[3] = { r.accum, r.accum + r.n, r.accum + 2 * r.n };
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float v = o[a];
        if (r.accum) v = (float)((double)v + (double)ld32(A[a], io));
        if (r.divisor > 0.0) v = (float)((double)v * r.invDivisor);
        *reinterpret_cast<float *>(reinterpret_cast<char *>(O[a]) + io) = v;
    }
}

// Compute::UpdateVelocityField (reg-lib/Compute.cpp:336-348): grid + gradient * scale, the
// gradient zeroed on the axes not optimised; reg_tools arithmetic in double.
__global__ void __launch_bounds__(kBlock) k_update_velocity(float *__restrict__ cp, const float *__restrict__ grad,
                                                            float scale, size_t G, int ox, int oy, int oz) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= 3 * G) return;
    const int c = (int)(i / G);
    const bool on = c == 0 ? ox : (c == 1 ? oy : oz);
    const float s = on ? (float)((double)grad[i] * (double)scale) : 0.f;
    cp[i] = (float)((double)cp[i] + (double)s);
}

// Compute::SymmetriseVelocityFields (reg-lib/Compute.cpp:357-391): displacements f, b of the two
// grids (same geometry) become (f - b) / 2 and (b - f) / 2.
__global__ void __launch_bounds__(kBlock) k_symmetrise(float *__restrict__ fw, float *__restrict__ bw, FieldGeom g) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.n) return;
    int x, y, z;
    index_xyz(i, g.nx, g.ny, x, y, z);
    const float3 p = id_pos(g.xyz, x, y, z);
    const float pv[3] = { p.x, p.y, p.z };
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const size_t k = i + c * g.n;
        const float f = fw[k] - pv[c], b = bw[k] - pv[c];
        const float nf = (float)((double)f - (double)b), nb = (float)((double)b - (double)f);
        fw[k] = (float)((double)nf * 0.5) + pv[c];
        bw[k] = (float)((double)nb * 0.5) + pv[c];
    }
}

inline dim3 blocks_for(size_t n) { return dim3((unsigned)((n + kBlock - 1) / kBlock)); }

// grid of a kernel indexed through xcd_block: a multiple of 8 workgroups
inline dim3 blocks8_for(size_t n) { return dim3((unsigned)(((n + kBlock - 1) / kBlock + 7) / 8 * 8)); }

// k_resample_gradient32 when its conditions hold (planes below 4 GiB, at least two voxels per axis,
// a power-of-two divisor), else k_resample_gradient.
void launch_resample_gradient(ResampleGradArgs r, hipStream_t s) {
    const size_t lim = 0xffffffffull / 4;
    int e = 0;
    const bool pow2 = r.divisor <= 0.0 || std::frexp(r.divisor, &e) == 0.5;
    if (3 * r.n <= lim && 3 * r.gn <= lim && r.nx >= 2 && r.ny >= 2 && r.nz >= 2 && pow2) {
        for (int k = 0; k < 3; ++k) r.rsp[k] = (float)(1.0 / (double)r.sp[k]);
        r.invDivisor = r.divisor > 0.0 ? 1.0 / r.divisor : 0.0;
        hipLaunchKernelGGL(k_resample_gradient32, blocks8_for(r.n), dim3(kBlock), 0, s, r);
    } else {
        hipLaunchKernelGGL(k_resample_gradient, blocks8_for(r.n), dim3(kBlock), 0, s, r);
    }
}


// The flow kernels' persistent grid: enough workgroups to fill the device (8 per CU); the
// extremes variant ends with one atomic pair per workgroup.
inline unsigned flow_blocks(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)std::min<size_t>(b, 2048);
}

VelGrid vel_grid(const nrh_f3d *h) {
    // the flow's fast path reads a 4x4x4 window from node 0 on lanes without a voxel
    NRH_REQUIRE(h->cpp.nx >= 4 && h->cpp.ny >= 4 && h->cpp.nz >= 4, NRH_ERR_UNSUPPORTED,
                "velocity grids need at least 4 nodes along each axis");
    VelGrid g;
    g.cp = h->cppPos.p;
    g.gnx = h->cpp.nx; g.gny = h->cpp.ny; g.gnz = h->cpp.nz;
    g.G = h->G;
    g.xyz = to_affine(h->cpp.xyz);
    g.ijk = to_affine(h->cpp.ijk);
    g.order = h->params.field_order;
    return g;
}

// reg-lib/cpu/_reg_localTrans.cpp:3520-3544: the number of squarings that brings the largest
// displacement component below 0.28 mm (3-D) or 0.35 mm (2-D) per step, at least 6.
int squaring_number(float mn, float mx, bool is3d) {
    float extrema = std::fabs(mn) < mx ? mx : std::fabs(mn);
    const float maxLength = is3d ? 0.28f : 0.35f;
    int sq = 1;
    while (extrema / (float)std::pow(2.0, sq) >= maxLength && sq < 1000) sq++;
    return sq < 6 ? 6 : sq;
}

float key_to_float(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

// The affine parts of grid handle `grid` on field geometry `field`; `sub` (may be null) supplies
// the transformation removed by the gradient exponentiation.
VelAff vel_aff(const nrh_f3d *grid, const nrh_geom& field, const nrh_f3d *sub) {
    VelAff a{};
    a.pre = grid->velExt0;
    a.post = grid->velExt1;
    a.sub = sub && sub->velTrans;
    float m[16];
    if (a.pre) {
        std::memcpy(a.A, grid->velExt0M, sizeof a.A);
        mat44_product(grid->velExt0M, field.xyz, m);
        std::memcpy(a.AS, m, sizeof a.AS);
    }
    if (a.post) std::memcpy(a.B, grid->velExt1M, sizeof a.B);
    if (a.sub) {
        mat44_product(sub->velTransM, field.xyz, m);
        std::memcpy(a.TS, m, sizeof a.TS);
    }
    return a;
}

double scaling_of(int steps) {
    const int sq = steps < 0 ? -steps : steps;
    const double s = (double)(float)std::pow(2.0, sq);
    return steps < 0 ? -s : s;
}

void launch_compose(const float *D, const FieldGeom& d, const float *pos, size_t m, float *out, hipStream_t s) {
    hipLaunchKernelGGL(k_compose, blocks8_for(m), dim3(kBlock), 0, s, D, d, pos, m, out);
}

// The rotation-free, spacing-normalised Jacobian factors of reg_trilinearResampleGradient for a
// field on geometry g: reorientation from the sform (reg-lib/cpu/_reg_resampling.cpp:1580) and
// the real spacing (reg_getRealImageSpacing, reg-lib/cpu/_reg_tools.cpp:253-274) when the sform
// is set, else the voxel size.  Host arithmetic as the reference: niftilib's float mat33 helpers
// and Mat44Mul<float>.
void resample_gradient_factors(const nrh_geom& g, float reo[9], float sp[3]) {
    M33 s;
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) s.m[a][b] = g.xyz[a * 4 + b];
    const M33 R = m33_inverse(m33_polar(s));
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) reo[a * 3 + b] = R.m[a][b];
    if (g.sform) {
        auto mul = [&](const float in[3], float out[3]) {
            for (int r = 0; r < 3; ++r)
                out[r] = (float)((double)g.xyz[r * 4] * (double)in[0] + (double)g.xyz[r * 4 + 1] * (double)in[1] +
                                 (double)g.xyz[r * 4 + 2] * (double)in[2] + (double)g.xyz[r * 4 + 3]);
        };
        const float o[3] = { 0.f, 0.f, 0.f };
        float r1[3];
        mul(o, r1);
        for (int k = 0; k < 3; ++k) {
            float e[3] = { 0.f, 0.f, 0.f }, r2[3];
            e[k] = 1.f;
            mul(e, r2);
            volatile float a = r1[0] - r2[0], b = r1[1] - r2[1], c = r1[2] - r2[2];
            volatile float a2 = a * a, b2 = b * b, c2 = c * c;
            volatile float sum = a2 + b2;
            sum = sum + c2;
            sp[k] = std::sqrt((float)sum);
        }
        if (g.nz <= 1) sp[2] = g.dz;
    } else {
        sp[0] = g.dx; sp[1] = g.dy; sp[2] = g.dz;
    }
}

}  // namespace

// Compute::GetDefFieldFromVelocityGrid (reg-lib/Compute.cpp:223-228):
// reg_spline_getDefFieldFromVelocityGrid into the handle's deformation field.
void op_vel_def_field(nrh_f3d *h, bool update) {
    NRH_REQUIRE(h->velocity, NRH_ERR_STATE, "the control points are not a velocity grid");
    // the field already holds this exponentiation: same control points and the same step rule
    if (h->velFieldVersion == h->cppVersion && !h->field.empty()) {
        if (!update && h->velSteps == h->velFieldSteps) return;
        if (update && h->velFieldUpdated && (h->velSteps >= 0) == (h->velFieldSteps > 0)) {
            h->velSteps = h->velFieldSteps;
            return;
        }
    }
    h->velFieldVersion = 0;
    h->velChainVersion = 0;
    h->pairStateVersion = 0;
    ensure_field(h);
    const FieldGeom f = field_geom(h->ref);
    const VelGrid g = vel_grid(h);
    const VelAff a = vel_aff(h, h->ref, nullptr);
    int steps = h->velSteps;
    const bool chosen = update || steps == 0;
    int sq;
    if (chosen) {
        h->velTmp.alloc(3 * h->N);
        const uint32_t init[2] = { 0xffffffffu, 0u };
        h->velKeys.alloc(2);
        NRH_HIP(hipMemcpyAsync(h->velKeys.p, init, sizeof init, hipMemcpyHostToDevice, h->stream));
        hipLaunchKernelGGL(k_vel_flow_disp, dim3(flow_blocks(f.n)), dim3(kBlock), 0, h->stream, g, f, a,
                           h->velTmp.p, h->velKeys.p);
        NRH_HIP(hipGetLastError());
        uint32_t keys[2];
        NRH_HIP(hipMemcpyAsync(keys, h->velKeys.p, sizeof keys, hipMemcpyDeviceToHost, h->stream));
        NRH_HIP(hipStreamSynchronize(h->stream));
        sq = squaring_number(key_to_float(keys[0]), key_to_float(keys[1]), h->ref.nz > 1);
        steps = steps >= 0 ? sq : -sq;
    } else {
        sq = steps < 0 ? -steps : steps;
    }
    // the squaring chain: fields[k] for k < sq in velChain (the intermediate fields of
    // reg_spline_getIntermediateDefFieldFromVelGrid, which the opposite direction's gradient
    // exponentiation reuses on a shared geometry), the last one in the deformation field
    h->velChain.alloc((size_t)std::max(sq, 1) * 3 * f.n);
    auto field_k = [&](int k) { return k == sq ? h->field.p : h->velChain.p + (size_t)k * 3 * f.n; };
    if (chosen) {
        hipLaunchKernelGGL(k_vel_scale, blocks_for(f.n), dim3(kBlock), 0, h->stream, h->velTmp.p, f,
                           scaling_of(steps), field_k(0));
    } else {
        hipLaunchKernelGGL(k_vel_flow_scaled, dim3(flow_blocks(f.n)), dim3(kBlock), 0, h->stream, g, f, a,
                           scaling_of(steps), field_k(0));
    }
    NRH_HIP(hipGetLastError());
    for (int k = 0; k < sq; ++k) launch_compose(field_k(k), f, field_k(k), f.n, field_k(k + 1), h->stream);
    if (a.pre || a.post)
        hipLaunchKernelGGL(k_vel_affine_parts, blocks_for(f.n), dim3(kBlock), 0, h->stream, h->field.p, 1, f, a);
    NRH_HIP(hipGetLastError());
    h->velSteps = steps;
    h->velFieldVersion = h->cppVersion;
    h->velFieldSteps = steps;
    h->velFieldUpdated = chosen;
    h->velChainVersion = sq > 0 ? h->cppVersion : 0;
}

// reg_spline_getIntermediateDefFieldFromVelGrid (reg-lib/cpu/_reg_localTrans.cpp:3616-3683) of
// grid `g` (steps squarings) on geometry f: fields[0..count-1] (count <= |steps| + 1).
// The affine parts (extensions of the grid, the transformation removed) follow the squaring.
static void intermediate_fields(const VelGrid& g, const FieldGeom& f, const VelAff& a, int steps, float *fields,
                                int count, hipStream_t s) {
    hipLaunchKernelGGL(k_vel_flow_scaled, dim3(flow_blocks(f.n)), dim3(kBlock), 0, s, g, f, a, scaling_of(steps),
                       fields);
    for (int k = 0; k + 1 < count; ++k)
        launch_compose(fields + (size_t)k * 3 * f.n, f, fields + (size_t)k * 3 * f.n, f.n,
                       fields + (size_t)(k + 1) * 3 * f.n, s);
    if (a.pre || a.post || a.sub)
        hipLaunchKernelGGL(k_vel_affine_parts, blocks_for(f.n), dim3(kBlock), 0, s, fields, count, f, a);
    NRH_HIP(hipGetLastError());
}

// Compute::ExponentiateGradient (reg-lib/Compute.cpp:283-326): the voxel-based gradient of h,
// resampled along the intermediate fields of bw's velocity grid (evaluated on h's reference
// geometry, less the displacement of bw's transformation matrix) and accumulated, then divided by
// 2^|steps of bw|.
void op_exponentiate_gradient(nrh_f3d *h, nrh_f3d *bw, bool syncOther) {
    NRH_REQUIRE(bw->velocity, NRH_ERR_STATE, "the backward control points are not a velocity grid");
    NRH_REQUIRE(!h->voxelGrad.empty(), NRH_ERR_STATE, "voxel-based gradient not computed");
    const int steps = bw->velSteps;
    const int sq = steps < 0 ? -steps : steps;
    if (sq == 0) return;
    // the backward grid's last update ran on its own stream
    if (syncOther) NRH_HIP(hipStreamSynchronize(bw->stream));
    const FieldGeom f = field_geom(h->ref);
    // bw's own squaring chain holds these fields when both handles share the reference geometry
    // and bw's last exponentiation used its current control points and this step count
    const bool sameGeom = bw->ref.nx == h->ref.nx && bw->ref.ny == h->ref.ny && bw->ref.nz == h->ref.nz &&
                          std::memcmp(bw->ref.xyz, h->ref.xyz, sizeof h->ref.xyz) == 0 &&
                          std::memcmp(bw->ref.ijk, h->ref.ijk, sizeof h->ref.ijk) == 0;
    // (the chain holds the fields before their affine parts)
    const VelAff a = vel_aff(bw, h->ref, bw);
    const bool affine = a.pre || a.post || a.sub;
    const float *fields;
    if (sameGeom && !affine && bw->velChainVersion == bw->cppVersion && bw->velFieldSteps == steps) {
        fields = bw->velChain.p;
    } else {
        h->velFields.alloc((size_t)sq * 3 * f.n);
        intermediate_fields(vel_grid(bw), f, a, steps, h->velFields.p, sq, h->stream);
        fields = h->velFields.p;
    }
    h->velTmp.alloc(3 * f.n);
    ResampleGradArgs r{};
    r.gnx = h->ref.nx; r.gny = h->ref.ny; r.gnz = h->ref.nz;
    r.gn = f.n;
    r.gradIjk = f.ijk;
    r.nx = f.nx; r.ny = f.ny; r.nz = f.nz;
    r.n = f.n;
    resample_gradient_factors(h->ref, r.reo, r.sp);
    r.pad = 0.f;
    float *src = h->voxelGrad.p, *dst = h->velTmp.p;
    const double div = std::pow(2.0, sq);
    for (int k = 0; k < sq; ++k) {
        r.grad = src;
        r.accum = src;
        r.def = fields + (size_t)k * 3 * f.n;
        r.divisor = (k == sq - 1) ? div : 0.0;
        r.out = dst;
        launch_resample_gradient(r, h->stream);
        std::swap(src, dst);
    }
    NRH_HIP(hipGetLastError());
    if (src != h->voxelGrad.p)
        NRH_HIP(hipMemcpyAsync(h->voxelGrad.p, src, 3 * f.n * sizeof(float), hipMemcpyDeviceToDevice, h->stream));
}

// Compute::ResampleGradient(1, padding) (reg-lib/Compute.cpp:399-405): the voxel-based gradient
// resampled through the handle's deformation field into out (device, planar [3][N]).
void op_resample_gradient(nrh_f3d *h, float pad, float *out) {
    NRH_REQUIRE(!h->voxelGrad.empty(), NRH_ERR_STATE, "voxel-based gradient not computed");
    NRH_REQUIRE(!h->field.empty(), NRH_ERR_STATE, "deformation field not computed");
    const FieldGeom f = field_geom(h->ref);
    ResampleGradArgs r{};
    r.grad = h->voxelGrad.p;
    r.gnx = h->ref.nx; r.gny = h->ref.ny; r.gnz = h->ref.nz;
    r.gn = f.n;
    r.gradIjk = f.ijk;
    r.def = h->field.p;
    r.nx = f.nx; r.ny = f.ny; r.nz = f.nz;
    r.n = f.n;
    resample_gradient_factors(h->ref, r.reo, r.sp);
    r.pad = pad;
    r.accum = nullptr;
    r.divisor = 0.0;
    r.out = out;
    launch_resample_gradient(r, h->stream);
    NRH_HIP(hipGetLastError());
}

// Compute::DefFieldCompose (reg-lib/Compute.cpp:393-396): the handle's deformation field becomes
// `def` (device, geometry g) evaluated at its positions.
void op_compose_field(nrh_f3d *h, const nrh_geom& g, const float *def) {
    NRH_REQUIRE(!h->field.empty(), NRH_ERR_STATE, "deformation field not computed");
    h->velFieldVersion = 0;
    h->pairStateVersion = 0;
    launch_compose(def, field_geom(g), h->field.p, h->N, h->field.p, h->stream);
    NRH_HIP(hipGetLastError());
}

void op_update_velocity(nrh_f3d *h, float scale, bool ox, bool oy, bool oz) {
    NRH_REQUIRE(h->velocity, NRH_ERR_STATE, "the control points are not a velocity grid");
    hipLaunchKernelGGL(k_update_velocity, blocks_for(3 * h->G), dim3(kBlock), 0, h->stream, h->cppPos.p,
                       h->nodeGrad.p, scale, h->G, (int)ox, (int)oy, (int)oz);
    NRH_HIP(hipGetLastError());
    h->cppVersion++;
}

void op_symmetrise_velocity(nrh_f3d *fw, nrh_f3d *bw) {
    NRH_REQUIRE(fw->velocity && bw->velocity, NRH_ERR_STATE, "the control points are not velocity grids");
    NRH_REQUIRE(fw->cpp.nx == bw->cpp.nx && fw->cpp.ny == bw->cpp.ny && fw->cpp.nz == bw->cpp.nz &&
                    std::memcmp(fw->cpp.xyz, bw->cpp.xyz, sizeof fw->cpp.xyz) == 0,
                NRH_ERR_ARG, "forward and backward velocity grids differ in geometry");
    // the backward grid's last update ran on its own stream
    NRH_HIP(hipStreamSynchronize(bw->stream));
    hipLaunchKernelGGL(k_symmetrise, blocks_for(fw->G), dim3(kBlock), 0, fw->stream, fw->cppPos.p, bw->cppPos.p,
                       field_geom(fw->cpp));
    NRH_HIP(hipGetLastError());
    NRH_HIP(hipStreamSynchronize(fw->stream));
    fw->cppVersion++;
    bw->cppVersion++;
}

}  // namespace nrh
