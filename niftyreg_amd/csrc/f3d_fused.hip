// Fused F3D evaluation for gfx950.
//
// One objective evaluation is one sweep over the reference grid (k_sweep<false>): B-spline
// field, trilinear warp and joint-histogram fill, with no field or warped image written to HBM.
// One gradient evaluation is a second sweep (k_sweep<true>): the field and warp are recomputed
// together with the image gradient and the NMI voxel gradient, which is written once (planar
// [3][N]); three separable passes then project it onto the control points.
//
// Tiles.  A workgroup owns a 64 x 8 x 8 voxel tile (x fastest).  The control points that
// influence the tile (a window of at most 20 x 6 x 6 nodes when the grid spacing is >= 4 voxels)
// are staged in LDS and contracted along z and y for every row of the tile at once; each voxel
// then needs only its 4-tap x contraction, and the voxels of a tile run without barriers.
//
// Node gradient.  The reference convolves the voxel gradient with the cubic B-spline kernel of
// radius 2*spacing along x, y, z (each pass normalised by the in-image kernel sum) and samples the
// result at the nodes (reg-lib/Compute.cpp:230-281).  When the grid spacing is an integer number
// of voxels and the nodes fall on voxels, the weight of voxel v on node n is exactly the field's
// basis value B_{n - pre(v)}(t(v)) on each axis, divided by the per-axis normalisation at the
// node.  The projection is therefore the transpose of the separable field evaluation: a pass
// along x, one along y, one along z (with normalisation, reorientation and weight), each summing
// in a fixed order (deterministic).
//
// Voxels on voxel planes.  The reference evaluates the field in float; where a floating-image
// position lands within rounding of a voxel plane, its last bit decides the trilinear cell (and
// so the image gradient, which jumps there) and, on the image border, whether the voxel is inside
// at all.  On the identity grid every registration starts from, every voxel is such a voxel.  The
// sweeps therefore list each voxel whose position lies within a tolerance of a plane (above the
// reference's own float error) instead of using it, and a second pass (k_sweep_fixup) evaluates
// the listed voxels as the reference does: its float field in the summation order of the selected
// build, its resampling, image gradient and FP64 NMI voxel term.

#include "device.hpp"
#include "be_blocks.hpp"
#include "f3d_handle.hpp"
#include "f3d_batch.hpp"
#include <cmath>
#include <map>
#include <mutex>

namespace nrh {

namespace {

constexpr int TX = 64, TY = 8, TZ = 8;
constexpr int WXM = 20, WYM = 6, WZM = 6;  // node-window bounds for spacing >= 4 voxels
constexpr int kThreads = 256;
constexpr int kTileVoxels = TX * TY * TZ;

struct FusedArgs {
    const void *ref;
    const void *flo;
    const int *mask;
    int nx, ny, nz;
    int fnx, fny, fnz;
    Affine floIjk;
    float pad;
    int rb, fb;
    const float *cpp;
    int gnx, gny, gnz;
    size_t G;
    const int *preX, *preY, *preZ;
    const float4 *bX, *bY, *bZ;
    int tilesX, tilesY, tilesZ, numTiles;
    Affine cppXyz;  // grid voxel -> mm
    uint32_t *hist;
    const double *ent;
    const float *kTab;  // gradient sweep: K[r + w rb] = log p_r + log p_w - NMI log p_rw (padded to 4)
    float *gvox;      // gradient sweep: the NMI voxel gradient (planar [3][N])
    float *fieldOut;  // optional: the deformation field of the sweep (planar [3][N]), for checking
    // voxels left to the exact pass: positions within ambTol (floating voxels) of a voxel plane;
    // tile t lists its voxels in ambList[t * kTileVoxels ..] and their number in ambCounts[t]
    uint32_t *ambList;
    uint32_t *ambCounts;
    float ambTol;
    int fieldOrder;  // NRH_FIELD_ORDER_*
    int lut;         // 5-voxel spacing on every axis (the reference's table path)
    const double *logh;  // the reference's jointHistogramLog (exact pass of the gradient sweep)
};

template<class T> __device__ __forceinline__ float ld(const T *p, size_t i);
template<> __device__ __forceinline__ float ld<float>(const float *p, size_t i) { return p[i]; }
template<> __device__ __forceinline__ float ld<__half>(const __half *p, size_t i) { return __half2float(p[i]); }

// Append the voxels of a wave that ask for it to the tile's list (one LDS atomic per wave on the
// tile's count).  The tile's list has a slot for each of its voxels, so it cannot overflow.
__device__ __forceinline__ void append_voxels(bool want, uint32_t idx, uint32_t *list, uint32_t *count) {
    const uint64_t m = __ballot(want);
    if (m == 0) return;
    const int lane = (int)__lane_id();
    const int leader = __ffsll((unsigned long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (want) list[base + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = idx;
}

// Tile index with the tiles of one XCD contiguous: consecutive workgroups are dealt round-robin
// to the 8 XCDs, so workgroup b runs on XCD b % 8; giving XCD k a contiguous block of tiles keeps
// neighbouring tiles (which read overlapping floating-image lines) in the same L2.  A bijection
// of [0, numTiles) only when numTiles is a multiple of 8: launches round their grids up.
__device__ __forceinline__ int xcd_tile(int b, int numTiles) {
    const int per = (numTiles + 7) / 8;
    const int t = (b % 8) * per + b / 8;
    return t;
}

// One voxel of a sweep, in two steps so that a wave can have the floating-image gathers of
// several voxels in flight: sweep_fetch evaluates the B-spline field from the tile's
// y/z-contracted window, maps it into the floating image and loads the 8 trilinear taps
// (branch-free: clamped addresses, out-of-image taps replaced by the padding value afterwards);
// sweep_finish interpolates and either counts the joint histogram (value sweep) or forms the NMI
// voxel gradient (gradient sweep).
struct VoxelTaps {
    float v[2][2][2];
    float rx, ry, rz;
    float rv;
    bool active;
};

// Position of a voxel in the floating image.  The identity part of the field is the affine map
// of (pre + t + 1) per axis (cubic B-splines reproduce linear functions), split as idX[c] (the
// lane's x term, double) + idYZ[c] (the row's y, z and offset terms, double); the window holds
// displacements, contracted here along x.
struct VoxelPos {
    int fx, fy, fz;
    float rx, ry, rz;
    bool amb;  // within ambTol of a voxel plane on some axis
};

__device__ __forceinline__ VoxelPos sweep_position(const FusedArgs& a, const float *ayz, const float4 b,
                                                   const int pxRel, const double idX[3], const double *idYZ,
                                                   float *fieldOut, const uint32_t idx, const bool active) {
    float w[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float *q = ayz + pxRel * 3 + c;
        const float d = b.x * q[0] + b.y * q[3] + b.z * q[6] + b.w * q[9];
        w[c] = (float)((idX[c] + idYZ[c]) + (double)d);
    }
    if (fieldOut && active) {
        const uint32_t N = (uint32_t)a.nx * a.ny * a.nz;
        fieldOut[idx] = w[0];
        fieldOut[idx + N] = w[1];
        fieldOut[idx + 2 * N] = w[2];
    }
    const float3 p = affine_apply_f(a.floIjk, w[0], w[1], w[2]);
    VoxelPos r;
    r.amb = fminf(fminf(fabsf(p.x - rintf(p.x)), fabsf(p.y - rintf(p.y))), fabsf(p.z - rintf(p.z))) < a.ambTol;
    r.fx = floor_int(p.x);
    r.fy = floor_int(p.y);
    r.fz = floor_int(p.z);
    r.rx = p.x - (float)r.fx;
    r.ry = p.y - (float)r.fy;
    r.rz = p.z - (float)r.fz;
    return r;
}

__device__ __forceinline__ bool taps_inside(const FusedArgs& a, const VoxelPos& p) {
    return p.fx >= 0 && p.fx + 1 < a.fnx && p.fy >= 0 && p.fy + 1 < a.fny && p.fz >= 0 && p.fz + 1 < a.fnz;
}

// The 8 trilinear taps; `inside`: every tap lies in the image (one base address, fixed strides),
// otherwise clamped addresses with out-of-image taps replaced by the padding value.  Offsets are
// 32-bit: the fused path requires images below 2^31 voxels.
template<class TImg>
__device__ __forceinline__ void load_taps(const FusedArgs& a, const TImg *__restrict__ flo, const VoxelPos& p,
                                          bool inside, float v[2][2][2]) {
    const uint32_t sy = (uint32_t)a.fnx, sz = (uint32_t)a.fnx * a.fny;
    if (inside) {
        const TImg *q = flo + ((uint32_t)p.fz * sz + (uint32_t)p.fy * sy + (uint32_t)p.fx);
        v[0][0][0] = ld(q, 0);       v[0][0][1] = ld(q, 1);
        v[0][1][0] = ld(q, sy);      v[0][1][1] = ld(q, sy + 1);
        v[1][0][0] = ld(q, sz);      v[1][0][1] = ld(q, sz + 1);
        v[1][1][0] = ld(q, sz + sy); v[1][1][1] = ld(q, sz + sy + 1);
    } else {
        const uint32_t X[2] = { (uint32_t)min(max(p.fx, 0), a.fnx - 1), (uint32_t)min(max(p.fx + 1, 0), a.fnx - 1) };
        const uint32_t Y[2] = { (uint32_t)min(max(p.fy, 0), a.fny - 1), (uint32_t)min(max(p.fy + 1, 0), a.fny - 1) };
        const uint32_t Z[2] = { (uint32_t)min(max(p.fz, 0), a.fnz - 1), (uint32_t)min(max(p.fz + 1, 0), a.fnz - 1) };
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int bb = 0; bb < 2; ++bb)
#pragma unroll
                for (int aa = 0; aa < 2; ++aa) {
                    const bool in = (unsigned)(p.fx + aa) < (unsigned)a.fnx &&
                                    (unsigned)(p.fy + bb) < (unsigned)a.fny && (unsigned)(p.fz + c) < (unsigned)a.fnz;
                    const float t = ld(flo, Z[c] * sz + Y[bb] * sy + X[aa]);
                    v[c][bb][aa] = in ? t : a.pad;
                }
    }
}

// GRAD = false: counts the voxel's (reference, warped) bin, or, when the warped value lies within
// its error bound of an integer (where the reference's bin index (int)war, and the in-range test,
// flip with the last bits), lists the voxel for the exact pass instead.  The bound: a position
// error of up to ambTol floating voxels per axis moves the trilinear value by at most the taps'
// spread per axis, plus eight ulps for the interpolation's own rounding.  Called by every lane.
template<bool GRAD>
__device__ __forceinline__ void sweep_finish(const FusedArgs& a, VoxelTaps t, float *dyn, double coef, float& g0,
                                             float& g1, float& g2, uint32_t idx, uint32_t *ambList,
                                             uint32_t *ambCount) {
    g0 = g1 = g2 = 0.f;
    if (t.rx < 0) t.rx = 0;
    if (t.ry < 0) t.ry = 0;
    if (t.rz < 0) t.rz = 0;
    const float rx = t.rx, ry = t.ry, rz = t.rz;
    const float x0b = 1.f - rx, y0b = 1.f - ry, z0b = 1.f - rz;
    const float e00 = t.v[0][0][0] * x0b + t.v[0][0][1] * rx, e01 = t.v[0][1][0] * x0b + t.v[0][1][1] * rx;
    const float e10 = t.v[1][0][0] * x0b + t.v[1][0][1] * rx, e11 = t.v[1][1][0] * x0b + t.v[1][1][1] * rx;
    const float f0 = e00 * y0b + e01 * ry, f1 = e10 * y0b + e11 * ry;
    const float war = f0 * z0b + f1 * rz;
    const float rv = t.rv;
    if (!GRAD) {
        bool edge = false;
        if (t.active && rv == rv && war == war) {
            float lo = t.v[0][0][0], hi = lo;
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int e = 0; e < 2; ++e) {
                        lo = fminf(lo, t.v[c][b][e]);
                        hi = fmaxf(hi, t.v[c][b][e]);
                    }
            const float tol = 3.f * (hi - lo) * a.ambTol + 8.f * 1.1920929e-7f * fabsf(war);
            edge = fabsf(war - rintf(war)) <= tol;
            if (!edge && rv >= 0.f && rv < (float)a.rb && war >= 0.f && war < (float)a.fb)
                atomicAdd(&((uint32_t *)dyn)[(int)rv + (int)war * a.rb], 1u);
        }
        append_voxels(edge, idx, ambList, ambCount);
    } else if (t.active && rv == rv && war == war) {
        // image gradient in floating voxel units (TrilinearImageGradient)
        const float d00 = t.v[0][0][1] - t.v[0][0][0], d01 = t.v[0][1][1] - t.v[0][1][0];
        const float d10 = t.v[1][0][1] - t.v[1][0][0], d11 = t.v[1][1][1] - t.v[1][1][0];
        float gx = (d00 * y0b + d01 * ry) * z0b + (d10 * y0b + d11 * ry) * rz;
        float gy = (e01 - e00) * z0b + (e11 - e10) * rz;
        float gz = f1 - f0;
        if (gx != gx) gx = 0.f;
        if (gy != gy) gy = 0.f;
        if (gz != gz) gz = 0.f;
        // NMI derivative (reg_getVoxelBasedNmiGradient3d, reg-lib/cpu/_reg_nmi.cpp:501-593):
        //   (sum_r B_r log p_r)(sum_w B'_w) + (sum_r B_r)(sum_w B'_w log p_w)
        //   - NMI sum_r sum_w B_r B'_w log p_rw  =  sum_r sum_w B_r B'_w K[r, w]
        // over the 4 bins around each intensity (bins outside the histogram get weight 0, as the
        // reference skips them).  B_r: cubic B-spline weights of the reference intensity's
        // fraction; B'_w: their derivative for the warped one.
        const float rfl = floorf(rv), wfl = floorf(war);
        const float f = rv - rfl, gg = war - wfl;
        const float f2 = f * f, f3 = f2 * f, omf = 1.f - f;
        float br[4] = { omf * omf * omf * (1.f / 6.f), 0.66666669f - f2 + 0.5f * f3,
                        (-3.f * f3 + 3.f * f2 + 3.f * f + 1.f) * (1.f / 6.f), f3 * (1.f / 6.f) };
        const float omg = 1.f - gg;
        float dw[4] = { -0.5f * omg * omg, (1.5f * gg - 2.f) * gg, (0.5f + 1.5f * gg) * omg, 0.5f * gg * gg };
        const int rb0 = (int)rfl - 1, wb0 = (int)wfl - 1;
        int ri[4], wc[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const bool okr = (unsigned)(rb0 + k) < (unsigned)a.rb;
            const bool okw = (unsigned)(wb0 + k) < (unsigned)a.fb;
            ri[k] = okr ? rb0 + k : 0;
            br[k] = okr ? br[k] : 0.f;
            wc[k] = (okw ? wb0 + k : 0) * a.rb;
            dw[k] = okw ? dw[k] : 0.f;
        }
        // the 16 terms summed in FP64, as the reference sums them
        double S = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const double T = (double)br[0] * (double)dyn[ri[0] + wc[j]] + (double)br[1] * (double)dyn[ri[1] + wc[j]] +
                             (double)br[2] * (double)dyn[ri[2] + wc[j]] + (double)br[3] * (double)dyn[ri[3] + wc[j]];
            S += (double)dw[j] * T;
        }
        S *= coef;
        g0 = (float)(S * (double)gx);
        g1 = (float)(S * (double)gy);
        g2 = (float)(S * (double)gz);
    }
}

// The exact pass: a voxel a sweep listed, evaluated as the reference evaluates it (float field in
// the selected build's order, Mat44Mul, ResampleImage3D, and for the gradient
// TrilinearImageGradient and the FP64 voxel term of reg_getVoxelBasedNmiGradient3d).
// The 8 trilinear taps around floating position p (loaded together), padding outside the image.
template<class TImg>
__device__ __forceinline__ void exact_taps(const TImg *__restrict__ img, int nx, int ny, int nz, int px, int py,
                                           int pz, float pad, float v[2][2][2]) {
#pragma unroll
    for (int c = 0; c < 2; ++c)
#pragma unroll
        for (int bb = 0; bb < 2; ++bb)
#pragma unroll
            for (int aa = 0; aa < 2; ++aa) {
                const int X = px + aa, Y = py + bb, Z = pz + c;
                v[c][bb][aa] = (X > -1 && X < nx && Y > -1 && Y < ny && Z > -1 && Z < nz)
                                   ? ld(img, ((size_t)Z * ny + Y) * nx + X) : pad;
            }
}

// ResampleImage3D's trilinear value from the taps (reg-lib/cpu/_reg_resampling.cpp).
__device__ __forceinline__ float sample_taps(const float v[2][2][2], float3 p, int px, int py, int pz) {
#pragma clang fp contract(off)
    float rx = p.x - (float)px, ry = p.y - (float)py, rz = p.z - (float)pz;
    if (rx < 0) rx = 0;
    if (ry < 0) ry = 0;
    if (rz < 0) rz = 0;
    const float xb[2] = { 1.f - rx, rx }, yb[2] = { 1.f - ry, ry }, zb[2] = { 1.f - rz, rz };
    float intensity = 0.f;
    for (int c = 0; c < 2; ++c) {
        float yT = 0.f;
        for (int bb = 0; bb < 2; ++bb) {
            float xT = 0.f;
            for (int aa = 0; aa < 2; ++aa) xT += v[c][bb][aa] * xb[aa];
            yT += xT * yb[bb];
        }
        intensity += yT * zb[c];
    }
    return intensity;
}

// TrilinearImageGradient from the taps (reg-lib/cpu/_reg_resampling.cpp:1797-1915).
__device__ __forceinline__ float3 gradient_taps(const float v[2][2][2], float3 p, int px, int py, int pz) {
#pragma clang fp contract(off)
    float g0 = 0.f, g1 = 0.f, g2 = 0.f;
    float r = p.x - (float)px;
    const float xb[2] = { (float)(1.0 - (double)r), r };
    r = p.y - (float)py;
    const float yb[2] = { (float)(1.0 - (double)r), r };
    r = p.z - (float)pz;
    const float zb[2] = { (float)(1.0 - (double)r), r };
    const float deriv[2] = { -1.f, 1.f };
    for (int c = 0; c < 2; ++c) {
        float xx = 0.f, yy = 0.f, zz = 0.f;
        for (int bb = 0; bb < 2; ++bb) {
            float xT = 0.f, yT = 0.f;
            for (int aa = 0; aa < 2; ++aa) {
                const float coeff = v[c][bb][aa];
                xT += coeff * deriv[aa];
                yT += coeff * xb[aa];
            }
            xx += xT * yb[bb];
            yy += yT * deriv[bb];
            zz += yT * yb[bb];
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

// One listed voxel idx of pair a: the histogram pass counts it into the tile's LDS histogram h,
// the gradient pass writes its voxel gradient.
template<class TImg, bool GRAD>
__device__ __forceinline__ void fixup_voxel(const FusedArgs& a, const uint32_t idx, uint32_t *h) {
    const TImg *__restrict__ ref = (const TImg *)a.ref;
    const TImg *__restrict__ flo = (const TImg *)a.flo;
    const int x = (int)(idx % (uint32_t)a.nx), y = (int)((idx / (uint32_t)a.nx) % (uint32_t)a.ny);
    const int z = (int)(idx / ((uint32_t)a.nx * a.ny));
    const float4 bx = a.bX[x], by = a.bY[y], bz = a.bZ[z];
    const float xb[4] = { bx.x, bx.y, bx.z, bx.w }, yb[4] = { by.x, by.y, by.z, by.w };
    const float zb[4] = { bz.x, bz.y, bz.z, bz.w };
    const float3 w = ref_spline_point(a.cpp, a.G, a.gnx, a.gny, a.preX[x], a.preY[y], a.preZ[z], xb, yb, zb,
                                      a.fieldOrder, a.lut != 0);
    const float3 p = affine_apply_d(a.floIjk, w.x, w.y, w.z);
    const int px = floor_int(p.x), py = floor_int(p.y), pz = floor_int(p.z);
    float v[2][2][2];
    exact_taps(flo, a.fnx, a.fny, a.fnz, px, py, pz, a.pad, v);
    const float rv = ld(ref, idx);
    const float war = sample_taps(v, p, px, py, pz);
    if (!GRAD) {
        if (rv == rv && war == war && rv >= 0.f && rv < (float)a.rb && war >= 0.f && war < (float)a.fb)
            atomicAdd(&h[(int)rv + (int)war * a.rb], 1u);
    } else {
        float d[3] = { 0.f, 0.f, 0.f };
        if (rv == rv && war == war) {
            const float3 g = gradient_taps(v, p, px, py, pz);
            const float gg[3] = { g.x, g.y, g.z };
            nmi_voxel_term(rv, war, gg, a.logh, a.ent, a.rb, a.fb, 1.0, d);
        }
        const uint32_t N = (uint32_t)a.nx * a.ny * a.nz;
        a.gvox[idx] = d[0];
        a.gvox[idx + N] = d[1];
        a.gvox[idx + 2 * N] = d[2];
    }
}

// The exact pass over the voxels that tiles [t0, t0 + nt) of pair a listed, one workgroup for
// the chunk: the lists are concatenated (offsets in LDS) so that every thread takes listed voxels;
// the histogram pass counts them in one LDS histogram h per workgroup and merges it once.
constexpr int kFixChunk = 8;  // tiles per exact-pass workgroup

template<class TImg, bool GRAD>
__device__ __forceinline__ void fixup_tiles(const FusedArgs& a, const int t0, const int nt, uint32_t *h) {
    __shared__ uint32_t sOff[kFixChunk + 1];
    const int tid = threadIdx.x;
    if (tid == 0) {
        uint32_t o = 0;
        for (int k = 0; k < nt; ++k) {
            sOff[k] = o;
            o += a.ambCounts[t0 + k];
        }
        for (int k = nt; k <= kFixChunk; ++k) sOff[k] = o;
    }
    __syncthreads();
    const uint32_t total = sOff[kFixChunk];
    if (total == 0) return;  // uniform over the workgroup
    if (!GRAD) {
        for (int i = tid; i < a.rb * a.fb; i += blockDim.x) h[i] = 0u;
        __syncthreads();
    }
    for (uint32_t i = tid; i < total; i += blockDim.x) {
        int k = 0;
        while (i >= sOff[k + 1]) ++k;
        const uint32_t idx = a.ambList[(size_t)(t0 + k) * kTileVoxels + (i - sOff[k])];
        fixup_voxel<TImg, GRAD>(a, idx, h);
    }
    if (!GRAD) {
        __syncthreads();
        for (int i = tid; i < a.rb * a.fb; i += blockDim.x)
            if (h[i]) atomicAdd(&a.hist[i], h[i]);
    }
}

template<class TImg, bool GRAD>
__global__ void __launch_bounds__(256) k_sweep_fixup(FusedArgs a) {
    extern __shared__ __align__(16) uint32_t fixLds[];  // histogram counts (objective pass)
    const int t0 = blockIdx.x * kFixChunk;
    fixup_tiles<TImg, GRAD>(a, t0, min(kFixChunk, a.numTiles - t0), fixLds);
}

// The histogram pass's exact pass with the bending energy in the same launch: workgroups
// [0, be.nb) evaluate the bending-energy bricks (be_value_block, its result into be.out), the
// others the exact-pass chunks.  The two are independent (the bricks read only the control
// points), and both are latency-bound, so one launch overlaps them without a second stream.
struct BeValueArgs {
    const float *cpp;
    int gnx, gny, gnz;
    BrickGrid bg;
    double *partials;
    uint32_t *ticket;
    double scale;
    double *out;
    int nb;  // bricks
};

template<class TImg>
__global__ void __launch_bounds__(256) k_sweep_fixup_be(FusedArgs a, BeValueArgs be, BeBasis B) {
    extern __shared__ __align__(16) uint32_t fixLds[];
    if ((int)blockIdx.x < be.nb) {  // uniform over the workgroup
        be_value_block(be.cpp, be.gnx, be.gny, be.gnz, be.bg, B, be.partials, be.ticket, be.scale, be.out,
                       (int)blockIdx.x, be.nb);
        return;
    }
    const int t0 = ((int)blockIdx.x - be.nb) * kFixChunk;
    fixup_tiles<TImg, false>(a, t0, min(kFixChunk, a.numTiles - t0), fixLds);
}

// Batched: one workgroup per chunk (pair, first tile, tiles) of the exact-pass chunk list.
template<class TImg, bool GRAD>
__global__ void __launch_bounds__(256)
k_sweep_fixup_batch(const FusedArgs *__restrict__ A, const int4 *__restrict__ chunks, const int *active) {
    extern __shared__ __align__(16) uint32_t fixLds[];
    const int4 c = chunks[blockIdx.x];
    if (active && !active[c.x]) return;
    fixup_tiles<TImg, GRAD>(A[c.x], c.y, c.z, fixLds);
}

// A 64 x 8 x 8 tile per workgroup.  The control-point window is staged in LDS and contracted
// along z and y for every (z, y) row of the tile in one pass; the voxels then run without
// barriers (one wave per x row), which keeps many floating-image gathers in flight.  The window
// shares LDS with the histogram / NMI table, which are filled after the contraction.
struct SweepShared {
    float sAyz[TZ][TY][WXM][3];
    double sIdYZ[TZ][TY][3];  // identity terms of each row: m[c][1] uy + m[c][2] uz + m[c][3]
    float4 sBx[TX], sBy[TY], sBz[TZ];
    int sPx[TX], sPy[TY], sPz[TZ];
    uint32_t ambCount;  // voxels of the tile listed for the exact pass
};

// One tile t of the pair described by a; dyn: the dynamic LDS (window [3][WZM][WYM][WXM], then
// the histogram or the table K).  Called by every thread of the workgroup.
template<class TImg, bool GRAD>
__device__ __forceinline__ void sweep_tile(const FusedArgs& a, const int t, SweepShared& S, float *dyn) {
    auto& sAyz = S.sAyz;
    auto& sIdYZ = S.sIdYZ;
    auto& sBx = S.sBx;
    auto& sBy = S.sBy;
    auto& sBz = S.sBz;
    auto& sPx = S.sPx;
    auto& sPy = S.sPy;
    auto& sPz = S.sPz;
    const int tid = threadIdx.x;
    if (tid == 0) S.ambCount = 0u;  // read after the tile's barriers
    uint32_t *ambList = a.ambList + (size_t)t * kTileVoxels;
    const int tx = t % a.tilesX, ty = (t / a.tilesX) % a.tilesY, tz = t / (a.tilesX * a.tilesY);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    const int nX = min(TX, a.nx - x0), nY = min(TY, a.ny - y0), nZ = min(TZ, a.nz - z0);

    if (tid < TX) {
        const int x = x0 + min(tid, nX - 1);
        sPx[tid] = a.preX[x];
        sBx[tid] = a.bX[x];
    } else if (tid < TX + TY) {
        const int i = tid - TX, y = y0 + min(i, nY - 1);
        sPy[i] = a.preY[y];
        sBy[i] = a.bY[y];
    } else if (tid < TX + TY + TZ) {
        const int i = tid - TX - TY, z = z0 + min(i, nZ - 1);
        sPz[i] = a.preZ[z];
        sBz[i] = a.bZ[z];
    }
    const int px0 = a.preX[x0], py0 = a.preY[y0], pz0 = a.preZ[z0];
    const int WX = a.preX[x0 + nX - 1] - px0 + 4;
    const int WY = a.preY[y0 + nY - 1] - py0 + 4;
    const int WZ = a.preZ[z0 + nZ - 1] - pz0 + 4;
    float *win = dyn;  // [c][iz][iy][ix] with strides WZM*WYM*WXM, WYM*WXM, WXM
    {
        // every load of the window is issued before the first LDS store; thread (row slot, ix) with
        // compile-time row strides, so no index needs a division by a run-time size
        constexpr int kRows = 3 * WZM * WYM, kSlots = kThreads / 32, kIter = (kRows + kSlots - 1) / kSlots;
        const int ix = tid & 31;
        float cv[kIter];
#pragma unroll
        for (int u = 0; u < kIter; ++u) {
            const int r = (tid >> 5) + u * kSlots, c = r / (WZM * WYM), iz = (r / WYM) % WZM, iy = r % WYM;
            if (r < kRows && ix < WX && iz < WZ && iy < WY)
                cv[u] = a.cpp[c * a.G + ((size_t)(pz0 + iz) * a.gny + (py0 + iy)) * a.gnx + (px0 + ix)];
        }
#pragma unroll
        for (int u = 0; u < kIter; ++u) {
            const int r = (tid >> 5) + u * kSlots, c = r / (WZM * WYM), iz = (r / WYM) % WZM, iy = r % WYM;
            if (r < kRows && ix < WX && iz < WZ && iy < WY) {
                const double node = (double)a.cppXyz.m[c * 4] * (px0 + ix) +
                                    (double)a.cppXyz.m[c * 4 + 1] * (py0 + iy) +
                                    (double)a.cppXyz.m[c * 4 + 2] * (pz0 + iz) + (double)a.cppXyz.m[c * 4 + 3];
                win[r * WXM + ix] = (float)((double)cv[u] - node);
            }
        }
    }
    __syncthreads();
    // z contraction of the window for every slice of the tile (into the LDS after the window),
    // then y contraction for every (z, y) row
    float *az = dyn + 3 * WZM * WYM * WXM;  // [c][z][iy][ix] with strides TZ*WYM*WXM, WYM*WXM, WXM
    for (int i = tid; i < nZ * WY * WX; i += kThreads) {
        const int ix = i % WX, iy = (i / WX) % WY, z = i / (WX * WY);
        const float4 bz = sBz[z];
        const int oz = sPz[z] - pz0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float *w0 = win + ((c * WZM + oz) * WYM + iy) * WXM + ix;
            az[((c * TZ + z) * WYM + iy) * WXM + ix] = bz.x * w0[0] + bz.y * w0[WYM * WXM] +
                                                      bz.z * w0[2 * WYM * WXM] + bz.w * w0[3 * WYM * WXM];
        }
    }
    __syncthreads();
    for (int i = tid; i < nZ * nY * WX; i += kThreads) {
        const int ix = i % WX, y = (i / WX) % nY, z = i / (WX * nY);
        const float4 b = sBy[y];
        const int o = sPy[y] - py0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float *a0 = az + ((c * TZ + z) * WYM + o) * WXM + ix;
            sAyz[z][y][ix][c] = b.x * a0[0] + b.y * a0[WXM] + b.z * a0[2 * WXM] + b.w * a0[3 * WXM];
        }
    }
    if (tid < nZ * TY * 3) {
        const int c = tid % 3, y = (tid / 3) % TY, z = tid / (3 * TY);
        const int yy = min(y, nY - 1);
        const float4 by = sBy[yy], bz = sBz[z];
        const double uy = (double)sPy[yy] + (double)(by.y + 2.f * by.z + 3.f * by.w);
        const double uz = (double)sPz[z] + (double)(bz.y + 2.f * bz.z + 3.f * bz.w);
        sIdYZ[z][y][c] = (double)a.cppXyz.m[c * 4 + 1] * uy + (double)a.cppXyz.m[c * 4 + 2] * uz +
                         (double)a.cppXyz.m[c * 4 + 3];
    }
    __syncthreads();
    double coef = 0.0;
    if (GRAD) {
        // the table K of k_nmi_table: the voxel NMI derivative is the bicubic contraction
        // sum_r sum_w B(ref - r) B'(war - w) K[r, w]; four float4 loads in flight per thread
        const int n4 = (a.rb * a.fb + 3) / 4;
        const float4 *src = (const float4 *)a.kTab;
        float4 *dst4 = (float4 *)dyn;
        for (int base = 0; base < n4; base += 4 * kThreads) {
            // unconditional loads (clamped index) keep v in registers with all four in flight
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = src[min(base + u * kThreads + tid, n4 - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = base + u * kThreads + tid;
                if (i < n4) dst4[i] = v[u];
            }
        }
        coef = 1.0 / (a.ent[2] * a.ent[3]);
    } else {
        uint32_t *h = (uint32_t *)dyn;
        for (int i = tid; i < a.rb * a.fb; i += kThreads) h[i] = 0u;
    }
    __syncthreads();

    const int lane = tid & 63, wave = tid >> 6;
    const TImg *__restrict__ ref = (const TImg *)a.ref;
    const TImg *__restrict__ flo = (const TImg *)a.flo;
    if (lane < nX) {
        const float4 b = sBx[lane];
        const int px = sPx[lane], pxRel = px - px0;
        const double ux = (double)px + (double)(b.y + 2.f * b.z + 3.f * b.w);
        const double idX[3] = { (double)a.cppXyz.m[0] * ux, (double)a.cppXyz.m[4] * ux, (double)a.cppXyz.m[8] * ux };
        const uint32_t N = (uint32_t)a.nx * a.ny * a.nz;
        float *fieldOut = GRAD ? nullptr : a.fieldOut;
        constexpr int kRowsPerWave = TY / (kThreads / 64);
        for (int z = 0; z < nZ; ++z) {
            // the rows of this wave in slice z: every tap is loaded before any is used
            VoxelPos pos[kRowsPerWave];
            VoxelTaps t[kRowsPerWave];
            uint32_t idx[kRowsPerWave];
            bool inside = true;
#pragma unroll
            for (int k = 0; k < kRowsPerWave; ++k) {
                const int y = min(wave + k * (kThreads / 64), nY - 1);
                idx[k] = ((uint32_t)(z0 + z) * a.ny + (uint32_t)(y0 + y)) * a.nx + (uint32_t)(x0 + lane);
                t[k].active = !a.mask || a.mask[idx[k]] > -1;
                pos[k] = sweep_position(a, &sAyz[z][y][0][0], b, pxRel, idX, &sIdYZ[z][y][0], fieldOut, idx[k],
                                        t[k].active);
                // a voxel near a plane goes to the exact pass (rows past the tile's last are
                // clamped duplicates and list nothing)
                const bool amb = pos[k].amb && t[k].active && wave + k * (kThreads / 64) < nY;
                append_voxels(amb, idx[k], ambList, &S.ambCount);
                t[k].active = t[k].active && !pos[k].amb;
                inside = inside && taps_inside(a, pos[k]);
            }
            // wave-uniform choice between the fixed-stride and the clamped gathers (lanes past
            // the tile's last column sit out this vote, so they never force the clamped path)
            const bool allInside = __all(inside);
#pragma unroll
            for (int k = 0; k < kRowsPerWave; ++k) {
                load_taps<TImg>(a, flo, pos[k], allInside, t[k].v);
                t[k].rx = pos[k].rx;
                t[k].ry = pos[k].ry;
                t[k].rz = pos[k].rz;
                t[k].rv = ld(ref, idx[k]);
            }
#pragma unroll
            for (int k = 0; k < kRowsPerWave; ++k) {
                const int y = wave + k * (kThreads / 64);
                if (y >= nY) continue;  // uniform over the wave
                float g0, g1, g2;
                sweep_finish<GRAD>(a, t[k], dyn, coef, g0, g1, g2, idx[k], ambList, &S.ambCount);
                if (GRAD) {
                    a.gvox[idx[k]] = g0;
                    a.gvox[idx[k] + N] = g1;
                    a.gvox[idx[k] + 2 * N] = g2;
                }
            }
        }
    }
    __syncthreads();
    if (tid == 0) a.ambCounts[t] = S.ambCount;
    if (!GRAD) {
        const uint32_t *h = (const uint32_t *)dyn;
        for (int i = tid; i < a.rb * a.fb; i += kThreads)
            if (h[i]) atomicAdd(&a.hist[i], h[i]);
    }
}

template<class TImg, bool GRAD>
__global__ void __launch_bounds__(kThreads) k_sweep(FusedArgs a) {
    __shared__ SweepShared S;
    extern __shared__ __align__(16) float dyn[];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= a.numTiles) return;
    sweep_tile<TImg, GRAD>(a, t, S, dyn);
}

// Persistent sweep over the (pair, tile) work list of a batch of pairs: item = pair << 20 | tile,
// sorted by pair.  The list is cut into 8 contiguous ranges, one per XCD (workgroup b runs on XCD
// b % 8), so a pair's tiles, which read overlapping floating-image lines, share one L2; the
// workgroups of an XCD take the items of its range in turn.  Items of pairs whose flag in
// `active` is 0 are skipped (a line search that has ended for that pair).
template<class TImg, bool GRAD>
__global__ void __launch_bounds__(kThreads)
k_sweep_batch(const FusedArgs *__restrict__ args, const uint32_t *__restrict__ items, int nItems,
              const int *__restrict__ active) {
    __shared__ SweepShared S;
    extern __shared__ __align__(16) float dyn[];
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8, perXcd = gridDim.x / 8;
    const int per = (nItems + 7) / 8;
    const int lo = xcd * per, hi = min(nItems, lo + per);
    for (int i = lo + slot; i < hi; i += perXcd) {
        const uint32_t it = __builtin_amdgcn_readfirstlane(items[i]);  // uniform: scalar loads of the arguments
        const int pair = (int)(it >> 20);
        if (active && !active[pair]) continue;  // uniform over the workgroup
        sweep_tile<TImg, GRAD>(args[pair], (int)(it & 0xfffffu), S, dyn);
        __syncthreads();  // the next tile reuses the LDS
    }
}

// Projection of the voxel gradient onto the nodes, one axis at a time.  The reference convolves
// the voxel gradient with the cubic B-spline kernel along x, y, z (each pass normalised by the
// in-image kernel sum) and samples the result at the nodes (reg-lib/Compute.cpp:230-281); with
// nodes on voxels and integer spacing that equals the transposed field evaluation
// sum_v B_x B_y B_z g(v), divided per axis by the kernel sum at the node.  Three separable passes
// (x: [3][nz][ny][gnx], y: [3][nz][gny][gnx], z: the nodes) sum in a fixed order.
// x pass: X[c][row][ix] = sum_x B(x, ix) g[c][row][x] for rows = (z, y).  A workgroup stages R
// whole rows in LDS with coalesced loads.  Every voxel of a cell (the voxels sharing pre(x))
// feeds the same four nodes, so one thread per (row, cell) first forms the cell's four weighted
// sums per component, then one thread per (row, node) adds the four cells that reach the node.
struct ProjXArgs {
    const float *g;       // [3][rows][nx]
    float *out;           // [3][rows][gnx]
    int nx, gnx;
    size_t rows;
    int rowsPerBlock;
    int pre0, cells;      // pre(0) and the number of cells along x
    const int2 *cellX;    // voxels [first, last] of each cell
    const float4 *bX;
};

constexpr int kCellStride = 13;  // 12 sums per cell, padded to an odd stride (no LDS bank conflicts)

__device__ __forceinline__ void project_x_block(const ProjXArgs& p, const int block, float *ldsX) {
    // ldsX: rows [3][R][nx], then cell sums [R][cells][4][3]
    const int R = p.rowsPerBlock;
    const size_t row0 = (size_t)block * R;
    const int nr = (int)min((size_t)R, p.rows - row0);
    const size_t planeIn = p.rows * (size_t)p.nx, planeOut = p.rows * (size_t)p.gnx;
    float *rowsL = ldsX, *cellL = ldsX + (size_t)3 * R * p.nx;
    const int per = nr * p.nx;
    for (int base = 0; base < per; base += 8 * 256) {
        // the three components' loads in flight together
        float v[3][8];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int u = 0; u < 8; ++u)
                v[c][u] = p.g[c * planeIn + row0 * p.nx + min(base + u * 256 + (int)threadIdx.x, per - 1)];
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = base + u * 256 + (int)threadIdx.x;
                if (i < per) rowsL[(size_t)c * R * p.nx + i] = v[c][u];
            }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nr * p.cells; i += blockDim.x) {
        const int cell = i % p.cells, r = i / p.cells;
        const int2 xr = p.cellX[cell];
        float acc[4][3] = {};
        for (int x = xr.x; x <= xr.y; ++x) {
            const float4 b = p.bX[x];
            const float v0 = rowsL[r * p.nx + x], v1 = rowsL[(R + r) * p.nx + x], v2 = rowsL[(2 * R + r) * p.nx + x];
            acc[0][0] += b.x * v0; acc[0][1] += b.x * v1; acc[0][2] += b.x * v2;
            acc[1][0] += b.y * v0; acc[1][1] += b.y * v1; acc[1][2] += b.y * v2;
            acc[2][0] += b.z * v0; acc[2][1] += b.z * v1; acc[2][2] += b.z * v2;
            acc[3][0] += b.w * v0; acc[3][1] += b.w * v1; acc[3][2] += b.w * v2;
        }
        float *o = cellL + ((size_t)r * p.cells + cell) * kCellStride;
        for (int k = 0; k < 4; ++k)
            for (int c = 0; c < 3; ++c) o[k * 3 + c] = acc[k][c];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < nr * p.gnx; i += blockDim.x) {
        const int node = i % p.gnx, r = i / p.gnx;
        float s0 = 0.f, s1 = 0.f, s2 = 0.f;
        for (int k = 3; k >= 0; --k) {  // cells in x order: node - 3 first
            const int cell = node - k - p.pre0;
            if (cell < 0 || cell >= p.cells) continue;
            const float *o = cellL + ((size_t)r * p.cells + cell) * kCellStride + k * 3;
            s0 += o[0];
            s1 += o[1];
            s2 += o[2];
        }
        const size_t q = (row0 + r) * p.gnx + node;
        p.out[q] = s0;
        p.out[q + planeOut] = s1;
        p.out[q + 2 * planeOut] = s2;
    }
}

// y pass: Y[c][z][iy][ix] = sum_y B(y, iy) X[c][z][y][ix] (consecutive threads read consecutive ix).
struct ProjYArgs {
    const float *in;   // [3][nz][ny][gnx]
    float *out;        // [3][nz][gny][gnx]
    int ny, nz, gnx, gny;
    const int *preY;
    const float4 *bY;
    const int2 *rangeY;
};

// Taps of the z pass loaded per batch before any is summed: a node's whole support (four cells)
// at 5-voxel spacing, so each thread has one round of loads in flight.
constexpr int kTapBatch = 20;

__device__ __forceinline__ void project_y_at(const ProjYArgs& p, const size_t t) {
    if (t >= (size_t)p.nz * p.gny * p.gnx) return;
    int ix, iy, z;
    index_xyz(t, p.gnx, p.gny, ix, iy, z);
    const size_t planeIn = (size_t)p.nz * p.ny * p.gnx, planeOut = (size_t)p.nz * p.gny * p.gnx;
    const int2 r = p.rangeY[iy];
    float s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll 4
    for (int y = r.x; y <= r.y; ++y) {
        const int k = iy - p.preY[y];
        const float4 b = p.bY[y];
        const float w = k == 0 ? b.x : k == 1 ? b.y : k == 2 ? b.z : b.w;
        const size_t q = ((size_t)z * p.ny + y) * p.gnx + ix;
        s0 += w * p.in[q];
        s1 += w * p.in[q + planeIn];
        s2 += w * p.in[q + 2 * planeIn];
    }
    p.out[t] = s0;
    p.out[t + planeOut] = s1;
    p.out[t + 2 * planeOut] = s2;
}

// z pass fused with the per-node normalisation, the reorientation by the floating image's
// mm->voxel 3x3 (transposed, as reg_voxelCentricToNodeCentric) and the weight; nodes whose voxel
// position lies outside the reference image receive 0.
struct NodeArgs {
    const float *in;  // [3][nz][gny][gnx]
    int nz, gnx, gny, gnz;
    size_t G;
    const int *preZ;
    const float4 *bZ;
    const int2 *rangeZ;
    const float *invNx, *invNy, *invNz;
    Affine reorient;
    float weight;
    float *nodeGrad;
    // optional bending-energy term (reg_spline_approxBendingEnergyGradient3D's gathered sums,
    // computed beside the sweep): grad += beRatio * beGrad, then the points' in-place
    // (p - node) + node round trip (reg-lib/cpu/_reg_localTrans_regul.cpp:312, 450)
    const float *beGrad;
    float beRatio;
    float *cpp;
    Affine cppXyz;
    uint32_t *maxBits;  // optional: max length of the final gradient (float bits, atomicMax)
};

// a + r * b rounded after the product, as the reference's separate multiply and add
__device__ __forceinline__ float add_product(float a, float r, float b) {
#pragma clang fp contract(off)
    return a + r * b;
}

// sqrt(x^2 + y^2 + z^2) with every product rounded (reg_getMaximalLength)
__device__ __forceinline__ float length3(float x, float y, float z) {
#pragma clang fp contract(off)
    return sqrtf(x * x + y * y + z * z);
}

// Returns the float bits of the final gradient's length at node n (0 past the grid).
__device__ __forceinline__ uint32_t project_nodes_at(const NodeArgs& r, const size_t n) {
    if (n >= r.G) return 0u;
    int i, j, k;
    index_xyz(n, r.gnx, r.gny, i, j, k);
    const float inv = r.invNx[i] * r.invNy[j] * r.invNz[k];
    float u0 = 0.f, u1 = 0.f, u2 = 0.f;
    if (inv != 0.f) {
        const size_t plane = (size_t)r.nz * r.gny * r.gnx, col = (size_t)j * r.gnx + i;
        const int2 rz = r.rangeZ[k];
        for (int z0 = rz.x; z0 <= rz.y; z0 += kTapBatch) {
            float v[kTapBatch][3], w[kTapBatch];
#pragma unroll
            for (int u = 0; u < kTapBatch; ++u) {
                const int z = min(z0 + u, rz.y);  // clamped: taps past the range are loaded, not summed
                const int q = k - r.preZ[z];
                const float4 b = r.bZ[z];
                w[u] = q == 0 ? b.x : q == 1 ? b.y : q == 2 ? b.z : b.w;
                const size_t a = (size_t)z * r.gny * r.gnx + col;
                v[u][0] = r.in[a];
                v[u][1] = r.in[a + plane];
                v[u][2] = r.in[a + 2 * plane];
            }
#pragma unroll
            for (int u = 0; u < kTapBatch; ++u) {
                if (z0 + u <= rz.y) {
                    u0 += w[u] * v[u][0];
                    u1 += w[u] * v[u][1];
                    u2 += w[u] * v[u][2];
                }
            }
        }
        u0 *= inv;
        u1 *= inv;
        u2 *= inv;
    }
    const float *R = r.reorient.m;
    float g0 = (R[0] * u0 + R[4] * u1 + R[8] * u2) * r.weight;
    float g1 = (R[1] * u0 + R[5] * u1 + R[9] * u2) * r.weight;
    float g2 = (R[2] * u0 + R[6] * u1 + R[10] * u2) * r.weight;
    if (r.beGrad) {
        g0 = add_product(g0, r.beRatio, r.beGrad[n]);
        g1 = add_product(g1, r.beRatio, r.beGrad[n + r.G]);
        g2 = add_product(g2, r.beRatio, r.beGrad[n + 2 * r.G]);
        for (int c = 0; c < 3; ++c) {
            const float node = node_pos(r.cppXyz, c, i, j, k);
            r.cpp[n + c * r.G] = (r.cpp[n + c * r.G] - node) + node;
        }
    }
    r.nodeGrad[n] = g0;
    r.nodeGrad[n + r.G] = g1;
    r.nodeGrad[n + 2 * r.G] = g2;
    return __float_as_uint(length3(g0, g1, g2));
}

__global__ void __launch_bounds__(256) k_project_x(ProjXArgs p) {
    extern __shared__ float ldsX[];
    project_x_block(p, blockIdx.x, ldsX);
}

// blocks of one XCD take a contiguous range of z slabs, so each slab's rows go through one L2
__global__ void __launch_bounds__(256) k_project_y(ProjYArgs p) {
    project_y_at(p, (size_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256) k_project_nodes(NodeArgs r) {
    uint32_t bm = project_nodes_at(r, (size_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
    if (r.maxBits) {
        // reg_getMaximalLength of the final gradient (lengths >= 0 order like their bit patterns)
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t o = __shfl_down(bm, off, 64);
            bm = o > bm ? o : bm;
        }
        if ((threadIdx.x & 63) == 0) atomicMax(r.maxBits, bm);
    }
}

// Batched projections: item (pair, block of that pair); each output element is computed by the
// same body, in the same order, as in the pair's own launch.
__global__ void __launch_bounds__(256)
k_project_x_batch(const ProjXArgs *__restrict__ A, const int2 *__restrict__ items, const int *active) {
    extern __shared__ float ldsX[];
    const int2 it = items[blockIdx.x];
    if (active && !active[it.x]) return;
    project_x_block(A[it.x], it.y, ldsX);
}

__global__ void __launch_bounds__(256)
k_project_y_batch(const ProjYArgs *__restrict__ A, const int2 *__restrict__ items, const int *active) {
    const int2 it = items[blockIdx.x];
    if (active && !active[it.x]) return;
    project_y_at(A[it.x], (size_t)it.y * blockDim.x + threadIdx.x);
}

__global__ void __launch_bounds__(256)
k_project_nodes_batch(const NodeArgs *__restrict__ A, const int2 *__restrict__ items, const int *active) {
    const int2 it = items[blockIdx.x];
    if (active && !active[it.x]) return;
    project_nodes_at(A[it.x], (size_t)it.y * blockDim.x + threadIdx.x);
}

// K[r + w rb] = log p_r + log p_w - NMI log p_rw, evaluated in FP64 from the reference's log
// histogram and rounded once to float: the table the gradient sweep contracts.
__global__ void __launch_bounds__(256) k_nmi_table(const double *__restrict__ logh, const double *__restrict__ ent,
                                                   int rb, int fb, float *__restrict__ K) {
#pragma clang fp contract(off)
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rb * fb) return;
    const double nmi = (ent[0] + ent[1]) / ent[2];
    const double *lr = logh + rb * fb, *lw = lr + rb;
    K[i] = (float)((lr[i % rb] + lw[i / rb]) - nmi * logh[i]);
}


// Per pair (blockIdx.y): the NMI table K of the gradient sweep.
__global__ void __launch_bounds__(256) k_nmi_table_batch(const FusedArgs *__restrict__ A, const int *active) {
#pragma clang fp contract(off)
    if (active && !active[blockIdx.y]) return;
    const FusedArgs& a = A[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.rb * a.fb) return;
    const double nmi = (a.ent[0] + a.ent[1]) / a.ent[2];
    const double *lr = a.logh + a.rb * a.fb, *lw = lr + a.rb;
    ((float *)a.kTab)[i] = (float)((lr[i % a.rb] + lw[i / a.rb]) - nmi * a.logh[i]);
}

__global__ void k_axis_table(int n, float gvs, int lut, int *pre, float4 *b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int p;
    float t, v[4];
    cell_of(i, gvs, lut != 0, p, t);
    bspline_basis(t, v);
    pre[i] = p;
    b[i] = make_float4(v[0], v[1], v[2], v[3]);
}

__global__ void k_to_half(const float *in, __half *out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = __float2half(in[i]);
}

}  // namespace

// Device tables of the fused path, built once per handle.
struct FusedPlan {
    bool ok = false;
    int tilesX = 0, tilesY = 0, tilesZ = 0, numTiles = 0;
    DevBuf<int> preX, preY, preZ;
    DevBuf<float4> bX, bY, bZ;
    DevBuf<int2> rX, rY, rZ;     // voxels [first, last] in the support of each node, per axis
    DevBuf<int2> cellX;          // voxels [first, last] of each x cell (voxels sharing pre(x))
    int pre0X = 0, cellsX = 0;
    DevBuf<float> invNx, invNy, invNz;
    DevBuf<float> gvox, projX, projY;  // voxel gradient [3][N]; after the x pass; after the y pass
    DevBuf<float> kTab;                // NMI derivative table, rb * fb rounded up to 4 floats
    DevBuf<uint32_t> ambList;          // voxels left to the exact pass, kTileVoxels slots per tile
    DevBuf<uint32_t> ambCounts;        // listed voxels per tile
    float ambTol = 0.f;                // floating voxels
    bool lut = false;
};

// The plan is kept beside the handle in a side table keyed by handle address.
static std::map<const nrh_f3d *, FusedPlan *> g_planMap;
static std::mutex g_planMutex;

void fused_release(const nrh_f3d *h) {
    std::lock_guard<std::mutex> lk(g_planMutex);
    auto it = g_planMap.find(h);
    if (it != g_planMap.end()) {
        delete it->second;
        g_planMap.erase(it);
    }
}

static FusedPlan *plan_of(nrh_f3d *h) {
    {
        std::lock_guard<std::mutex> lk(g_planMutex);
        auto it = g_planMap.find(h);
        if (it != g_planMap.end()) return it->second;
    }
    FusedPlan *p = new FusedPlan;
    // Conditions of the fused path: integer grid spacing >= 4 voxels on every axis and nodes on
    // voxels (reference voxel -> node map = diag(s) with offset -s), float images.
    const float gv[3] = { h->cpp.dx / h->ref.dx, h->cpp.dy / h->ref.dy, h->cpp.dz / h->ref.dz };
    bool ok = true;
    int s[3];
    for (int i = 0; i < 3; ++i) {
        s[i] = (int)std::lround(gv[i]);
        ok = ok && std::fabs(gv[i] - (float)s[i]) < 1e-5f && s[i] >= 4 && s[i] <= 64;
    }
    // T = ref.ijk * cpp.xyz must be diag(s) with translation -s
    for (int r = 0; r < 3 && ok; ++r)
        for (int c = 0; c < 4; ++c) {
            double v = 0;
            for (int k = 0; k < 4; ++k) v += (double)h->ref.ijk[r * 4 + k] * (double)h->cpp.xyz[k * 4 + c];
            const double want = c == r ? s[r] : c == 3 ? -s[r] : 0.0;
            if (std::fabs(v - want) > 1e-3) ok = false;
        }
    // grid must cover the image: pre(n-1) + 3 < grid dim
    ok = ok && (h->ref.nx - 1) / s[0] + 3 < h->cpp.nx && (h->ref.ny - 1) / s[1] + 3 < h->cpp.ny &&
         (h->ref.nz - 1) / s[2] + 3 < h->cpp.nz;
    p->ok = ok;
    if (ok) {
        const bool lut = gv[0] == 5.f && gv[1] == 5.f && gv[2] == 5.f;
        const int dims[3] = { h->ref.nx, h->ref.ny, h->ref.nz };
        const int gdims[3] = { h->cpp.nx, h->cpp.ny, h->cpp.nz };
        const int tsz[3] = { TX, TY, TZ };
        const int wmax[3] = { WXM, WYM, WZM };
        DevBuf<int> *pre[3] = { &p->preX, &p->preY, &p->preZ };
        DevBuf<float4> *bb[3] = { &p->bX, &p->bY, &p->bZ };
        DevBuf<int2> *rr[3] = { &p->rX, &p->rY, &p->rZ };
        DevBuf<float> *inv[3] = { &p->invNx, &p->invNy, &p->invNz };
        int tiles[3];
        const float pd[3] = { h->ref.dx, h->ref.dy, h->ref.dz };
        const float sp[3] = { h->cpp.dx, h->cpp.dy, h->cpp.dz };
        for (int ax = 0; ax < 3; ++ax) {
            const int n = dims[ax];
            pre[ax]->alloc(n);
            bb[ax]->alloc(n);
            hipLaunchKernelGGL(k_axis_table, dim3((n + 255) / 256), dim3(256), 0, h->stream, n, gv[ax], lut ? 1 : 0,
                               pre[ax]->p, bb[ax]->p);
            std::vector<int> hpre(n);
            NRH_HIP(hipMemcpyAsync(hpre.data(), pre[ax]->p, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
            NRH_HIP(hipStreamSynchronize(h->stream));
            tiles[ax] = (n + tsz[ax] - 1) / tsz[ax];
            for (int t = 0; t < tiles[ax]; ++t) {
                const int a0 = t * tsz[ax], a1 = std::min(n, a0 + tsz[ax]) - 1;
                if (hpre[a1] - hpre[a0] + 4 > wmax[ax]) ok = false;
            }
            // voxels whose 4-node support contains each node (pre is non-decreasing)
            std::vector<int2> range(gdims[ax], make_int2(1, 0));
            for (int v = 0; v < n; ++v)
                for (int k = 0; k < 4; ++k) {
                    const int q = hpre[v] + k;
                    if (q < 0 || q >= gdims[ax]) continue;
                    if (range[q].x > range[q].y) range[q] = make_int2(v, v);
                    else range[q].y = v;
                }
            if (ax == 0) {
                p->pre0X = hpre[0];
                p->cellsX = hpre[n - 1] - hpre[0] + 1;
                std::vector<int2> cells(p->cellsX, make_int2(1, 0));
                for (int v = 0; v < n; ++v) {
                    int2& c = cells[hpre[v] - hpre[0]];
                    if (c.x > c.y) c = make_int2(v, v);
                    else c.y = v;
                }
                p->cellX.alloc(cells.size());
                NRH_HIP(hipMemcpyAsync(p->cellX.p, cells.data(), cells.size() * sizeof(int2), hipMemcpyHostToDevice,
                                       h->stream));
            }
            rr[ax]->alloc(gdims[ax]);
            NRH_HIP(hipMemcpyAsync(rr[ax]->p, range.data(), range.size() * sizeof(int2), hipMemcpyHostToDevice,
                                   h->stream));
            // per-node inverse normalisation: kernel of reg_tools_kernelConvolution (reg-lib/cpu/_reg_tools.cpp:868-904)
            // summed over the taps inside the image, as its density sum (float products, double sum)
            const double temp = (double)(sp[ax] / pd[ax]);
            const int radius = (int)(temp * 2.0f);
            std::vector<float> kern(2 * radius + 1);
            for (int i = -radius; i <= radius; ++i) {
                const double rel = std::fabs(i / temp);
                float kv = 0.f;
                if (rel < 1.0) kv = (float)(2.0 / 3.0 - rel * rel + 0.5 * rel * rel * rel);
                else if (rel < 2.0) kv = (float)(-((rel - 2.0) * (rel - 2.0) * (rel - 2.0)) / 6.0);
                kern[i + radius] = kv;
            }
            std::vector<float> invn(gdims[ax], 0.f);
            for (int q = 0; q < gdims[ax]; ++q) {
                const int pv = s[ax] * (q - 1);
                if (pv < 0 || pv >= n) continue;
                double ds = 0;
                for (int k = -radius; k <= radius; ++k)
                    if (pv + k >= 0 && pv + k < n) ds += (double)(kern[k + radius] * 1.f);
                invn[q] = 1.f / (float)ds;
            }
            inv[ax]->alloc(gdims[ax]);
            NRH_HIP(hipMemcpyAsync(inv[ax]->p, invn.data(), invn.size() * sizeof(float), hipMemcpyHostToDevice, h->stream));
        }
        p->tilesX = tiles[0];
        p->tilesY = tiles[1];
        p->tilesZ = tiles[2];
        p->numTiles = tiles[0] * tiles[1] * tiles[2];
        p->ok = ok;
        p->lut = lut;
        // Tolerance of the exact pass: the reference's float field is within a few float ulps of
        // its largest coordinates of the exact sum (measured: 1.4e-4 mm at 256 mm, about 9 ulps);
        // 48 ulps of the grid's largest corner coordinate (with a margin for displacements),
        // mapped into floating voxels by the largest row of the floating mm -> voxel matrix.
        double M = 0.0;
        for (int corner = 0; corner < 8; ++corner) {
            const double v[3] = { (corner & 1) ? h->cpp.nx - 1.0 : 0.0, (corner & 2) ? h->cpp.ny - 1.0 : 0.0,
                                  (corner & 4) ? h->cpp.nz - 1.0 : 0.0 };
            for (int r = 0; r < 3; ++r) {
                const double w = h->cpp.xyz[r * 4] * v[0] + h->cpp.xyz[r * 4 + 1] * v[1] +
                                 h->cpp.xyz[r * 4 + 2] * v[2] + h->cpp.xyz[r * 4 + 3];
                M = std::max(M, std::fabs(w));
            }
        }
        double rowMax = 0.0;
        for (int r = 0; r < 3; ++r)
            rowMax = std::max(rowMax, (double)std::fabs(h->flo.ijk[r * 4]) + std::fabs(h->flo.ijk[r * 4 + 1]) +
                                          std::fabs(h->flo.ijk[r * 4 + 2]));
        p->ambTol = (float)(48.0 * std::ldexp(1.0, -24) * (1.25 * M + 8.0) * rowMax);
        p->ambList.alloc((size_t)p->numTiles * kTileVoxels);
        p->ambCounts.alloc(p->numTiles);
        NRH_HIP(hipStreamSynchronize(h->stream));
    }
    std::lock_guard<std::mutex> lk(g_planMutex);
    g_planMap[h] = p;
    return p;
}

// The sweeps keep the joint histogram / NMI table in LDS: at most 12288 entries (bins <= 106).
// The voxel offsets of the sweeps are 32-bit: 3 N and the floating image below 2^32 elements.
bool fused_supported(nrh_f3d *h) {
    return plan_of(h)->ok && (size_t)h->params.ref_bins * h->params.flo_bins <= 12288 &&
           3 * (size_t)h->N <= 0xffffffffull && (size_t)h->Nf <= 0xffffffffull;
}

void collect_sweep_times(nrh_f3d *h) {
    for (int k = 0; k < 2; ++k) {
        if (!h->evPending[k]) continue;
        // a sweep still running (a gradient enqueued behind the objective) is collected later
        const hipError_t q = hipEventQuery(h->ev[2 * k + 1]);
        if (q == hipErrorNotReady) continue;
        NRH_HIP(q);
        float ms = 0.f;
        NRH_HIP(hipEventElapsedTime(&ms, h->ev[2 * k], h->ev[2 * k + 1]));
        h->msSweep[k] += ms;
        h->nSweep[k] += 1;
        h->evPending[k] = false;
    }
}

static FusedArgs make_args(nrh_f3d *h, FusedPlan *p) {
    FusedArgs a{};
    const bool half = h->params.image_storage == 1;
    if (half) {
        if (h->refH.empty() || h->halfDirty) {
            h->refH.alloc(h->N);
            h->floH.alloc(h->Nf);
            hipLaunchKernelGGL(k_to_half, dim3(2048), dim3(256), 0, h->stream, h->refImg.p, h->refH.p, h->N);
            hipLaunchKernelGGL(k_to_half, dim3(2048), dim3(256), 0, h->stream, h->floImg.p, h->floH.p, h->Nf);
            h->halfDirty = false;
        }
        a.ref = h->refH.p;
        a.flo = h->floH.p;
    } else {
        a.ref = h->refImg.p;
        a.flo = h->floImg.p;
    }
    a.mask = h->mask.p;
    a.nx = h->ref.nx; a.ny = h->ref.ny; a.nz = h->ref.nz;
    a.fnx = h->flo.nx; a.fny = h->flo.ny; a.fnz = h->flo.nz;
    a.floIjk = to_affine(h->flo.ijk);
    a.cppXyz = to_affine(h->cpp.xyz);
    a.pad = h->params.padding;
    a.rb = h->params.ref_bins;
    a.fb = h->params.flo_bins;
    a.cpp = h->cppPos.p;
    a.gnx = h->cpp.nx; a.gny = h->cpp.ny; a.gnz = h->cpp.nz;
    a.G = h->G;
    a.preX = p->preX.p; a.preY = p->preY.p; a.preZ = p->preZ.p;
    a.bX = p->bX.p; a.bY = p->bY.p; a.bZ = p->bZ.p;
    a.tilesX = p->tilesX; a.tilesY = p->tilesY; a.tilesZ = p->tilesZ;
    a.numTiles = p->numTiles;
    a.ambList = p->ambList.p;
    a.ambCounts = p->ambCounts.p;
    a.ambTol = p->ambTol;
    a.fieldOrder = h->params.field_order;
    a.lut = p->lut ? 1 : 0;
    a.logh = h->histLog.p;
    a.ent = h->ent.p;
    return a;
}



static unsigned grid_for(int numTiles) { return (unsigned)((numTiles + 7) / 8 * 8); }

// Dynamic LDS of a sweep: the control-point window and its z contraction, then the histogram / NMI
// table in their place.
static size_t sweep_lds(const FusedArgs& a) {
    // window [3][WZM][WYM][WXM] + its z contraction [3][TZ][WYM][WXM]; then histogram / table K
    return std::max<size_t>(((size_t)a.rb * a.fb + 3) / 4 * 4, (size_t)3 * (WZM + TZ) * WYM * WXM) * sizeof(float);
}

void fused_histogram(nrh_f3d *h, float *fieldOut, bool beValue, const double *pubS1, int pubN1) {
    FusedPlan *p = plan_of(h);
    NRH_REQUIRE(fused_supported(h), NRH_ERR_UNSUPPORTED, "fused path unavailable for this geometry or bin count");
    FusedArgs a = make_args(h, p);
    a.hist = h->histCount.p;
    a.fieldOut = fieldOut;
    // histCount is zero here: allocation zeroes it and every finalisation consumes it
    const size_t lds = sweep_lds(a);
    const bool timed = h->profile && h->profileSeen[0]++ % h->profilePeriod == 0;
    if (timed) NRH_HIP(hipEventRecord(h->ev[0], h->stream));
    if (h->params.image_storage == 1)
        hipLaunchKernelGGL((k_sweep<__half, false>), dim3(grid_for(p->numTiles)), dim3(kThreads), lds, h->stream, a);
    else
        hipLaunchKernelGGL((k_sweep<float, false>), dim3(grid_for(p->numTiles)), dim3(kThreads), lds, h->stream, a);
    NRH_HIP(hipGetLastError());
    if (timed) {
        NRH_HIP(hipEventRecord(h->ev[1], h->stream));
        h->evPending[0] = true;
    }
    const unsigned fixBlocks = (unsigned)((p->numTiles + kFixChunk - 1) / kFixChunk);
    const size_t fixLds = (size_t)a.rb * a.fb * sizeof(uint32_t);
    if (beValue) {
        BeValueArgs be;
        be.cpp = h->cppPos.p;
        be.gnx = h->cpp.nx;
        be.gny = h->cpp.ny;
        be.gnz = h->cpp.nz;
        be.bg = be_brick_grid(h->cpp.nx, h->cpp.ny, h->cpp.nz);
        be.nb = be.bg.bx * be.bg.by * be.bg.bz;
        h->partials.alloc(std::max<size_t>(h->partials.n, 2 * (size_t)be.nb + 64));
        be.partials = h->partials.p;
        be.ticket = h->maxBits.p + 6;
        be.scale = 1.0 / (double)(3 * h->G);
        be.out = h->scalars.p + 0;
        const dim3 grid(fixBlocks + (unsigned)be.nb);
        if (h->params.image_storage == 1)
            hipLaunchKernelGGL(k_sweep_fixup_be<__half>, grid, dim3(256), fixLds, h->stream, a, be, be_basis());
        else
            hipLaunchKernelGGL(k_sweep_fixup_be<float>, grid, dim3(256), fixLds, h->stream, a, be, be_basis());
    } else if (h->params.image_storage == 1) {
        hipLaunchKernelGGL((k_sweep_fixup<__half, false>), dim3(fixBlocks), dim3(256), fixLds, h->stream, a);
    } else {
        hipLaunchKernelGGL((k_sweep_fixup<float, false>), dim3(fixBlocks), dim3(256), fixLds, h->stream, a);
    }
    NRH_HIP(hipGetLastError());
    // the finalisation also writes the gradient sweep's NMI table
    p->kTab.alloc((size_t)(a.rb * a.fb + 3) / 4 * 4);
    op_nmi_finalize(h, p->kTab.p, pubS1, pubN1);
    h->histVersion = h->cppVersion;
}

// Arguments of the three projection passes of the voxel gradient g (planar [3][N]) onto the
// nodes of h (into h->nodeGrad); the x pass's LDS size in *ldsX.
static void projection_args(nrh_f3d *h, FusedPlan *p, const float *g, float weight, ProjXArgs& px, ProjYArgs& py,
                            NodeArgs& r, size_t *ldsX) {
    const int nx = h->ref.nx, ny = h->ref.ny, nz = h->ref.nz;
    const int gnx = h->cpp.nx, gny = h->cpp.ny, gnz = h->cpp.nz;
    p->projX.alloc((size_t)3 * nz * ny * gnx);
    p->projY.alloc((size_t)3 * nz * gny * gnx);
    px = ProjXArgs{};
    px.g = g;
    px.out = p->projX.p;
    px.nx = nx;
    px.gnx = gnx;
    px.rows = (size_t)ny * nz;
    px.rowsPerBlock = std::max(1, 1024 / nx);  // rows: <= 12 KB of LDS for nx <= 1024 (six blocks per CU)
    px.pre0 = p->pre0X;
    px.cells = p->cellsX;
    px.cellX = p->cellX.p;
    px.bX = p->bX.p;
    *ldsX = ((size_t)3 * px.rowsPerBlock * nx + (size_t)px.rowsPerBlock * px.cells * kCellStride) * sizeof(float);
    NRH_REQUIRE(*ldsX <= 64 * 1024, NRH_ERR_UNSUPPORTED, "image rows too long for the fused projection");
    py = ProjYArgs{};
    py.in = p->projX.p;
    py.out = p->projY.p;
    py.ny = ny;
    py.nz = nz;
    py.gnx = gnx;
    py.gny = gny;
    py.preY = p->preY.p;
    py.bY = p->bY.p;
    py.rangeY = p->rY.p;
    r = NodeArgs{};
    r.in = p->projY.p;
    r.nz = nz;
    r.gnx = gnx; r.gny = gny; r.gnz = gnz;
    r.G = h->G;
    r.preZ = p->preZ.p;
    r.bZ = p->bZ.p;
    r.rangeZ = p->rZ.p;
    r.invNx = p->invNx.p; r.invNy = p->invNy.p; r.invNz = p->invNz.p;
    r.reorient = to_affine(h->flo.ijk);
    r.weight = node_weight_ratio(h, weight);
    r.nodeGrad = h->nodeGrad.p;
    r.beGrad = nullptr;
    r.maxBits = nullptr;
}

// The voxel gradient g (planar [3][N]) projected onto the nodes: x pass (rows staged in LDS,
// cell sums), y pass, z pass with normalisation, reorientation and weight, into h->nodeGrad.
// beWeight > 0: the bending-energy sums the side stream computes (op_bending_energy_gradient_raw)
// are joined in and added by the z pass, which also round-trips the points and, with maxBits,
// tracks the final gradient's max length.
static void project_to_nodes(nrh_f3d *h, FusedPlan *p, const float *g, float weight, float beWeight = 0.f,
                             uint32_t *maxBits = nullptr) {
    ProjXArgs px;
    ProjYArgs py;
    NodeArgs r;
    size_t ldsX = 0;
    projection_args(h, p, g, weight, px, py, r, &ldsX);
    if (beWeight > 0.f) {
        // the bending-energy gradient sums on the side stream beside the x and y passes
        fork_side(h);
        op_bending_energy_gradient_raw(h, h->side);
        r.beGrad = h->beGrad.p;
        r.beRatio = beWeight / (float)h->G;
        r.cpp = h->cppPos.p;
        r.cppXyz = to_affine(h->cpp.xyz);
        r.maxBits = maxBits;
    }
    hipLaunchKernelGGL(k_project_x, dim3((unsigned)((px.rows + px.rowsPerBlock - 1) / px.rowsPerBlock)), dim3(256),
                       ldsX, h->stream, px);
    // xcd_tile needs a grid that is a multiple of 8 (extra blocks find no work)
    auto blocks8 = [](size_t n) { return dim3((unsigned)(((n + 255) / 256 + 7) / 8 * 8)); };
    hipLaunchKernelGGL(k_project_y, blocks8((size_t)py.nz * py.gny * py.gnx), dim3(256), 0, h->stream, py);
    if (beWeight > 0.f) join_side(h);
    hipLaunchKernelGGL(k_project_nodes, blocks8(h->G), dim3(256), 0, h->stream, r);
    NRH_HIP(hipGetLastError());
}

void fused_gradient(nrh_f3d *h, float weight, float beWeight, uint32_t *maxBits) {
    FusedPlan *p = plan_of(h);
    NRH_REQUIRE(fused_supported(h), NRH_ERR_UNSUPPORTED, "fused path unavailable for this geometry or bin count");
    FusedArgs a = make_args(h, p);
    a.ent = h->ent.p;
    p->gvox.alloc(3 * h->N);
    a.gvox = p->gvox.p;
    const int nbins = a.rb * a.fb;
    p->kTab.alloc((size_t)(nbins + 3) / 4 * 4);
    if (!h->kTabCurrent)  // the histogram came from another path: the table from its logs
        hipLaunchKernelGGL(k_nmi_table, dim3((nbins + 255) / 256), dim3(256), 0, h->stream, h->histLog.p, a.ent, a.rb,
                           a.fb, p->kTab.p);
    a.kTab = p->kTab.p;
    const size_t lds = sweep_lds(a);
    const bool timed = h->profile && h->profileSeen[1]++ % h->profilePeriod == 0;
    if (timed) NRH_HIP(hipEventRecord(h->ev[2], h->stream));
    if (h->params.image_storage == 1)
        hipLaunchKernelGGL((k_sweep<__half, true>), dim3(grid_for(p->numTiles)), dim3(kThreads), lds, h->stream, a);
    else
        hipLaunchKernelGGL((k_sweep<float, true>), dim3(grid_for(p->numTiles)), dim3(kThreads), lds, h->stream, a);
    NRH_HIP(hipGetLastError());
    if (timed) {
        NRH_HIP(hipEventRecord(h->ev[3], h->stream));
        h->evPending[1] = true;
    }
    const dim3 fixGrid((unsigned)((p->numTiles + kFixChunk - 1) / kFixChunk));
    if (h->params.image_storage == 1)
        hipLaunchKernelGGL((k_sweep_fixup<__half, true>), fixGrid, dim3(256), 0, h->stream, a);
    else
        hipLaunchKernelGGL((k_sweep_fixup<float, true>), fixGrid, dim3(256), 0, h->stream, a);
    NRH_HIP(hipGetLastError());
    project_to_nodes(h, p, p->gvox.p, weight, beWeight, maxBits);
    if (beWeight > 0.f) h->cppVersion++;  // the round trip may have changed the last bit of some points
}

// (a grid extension 0 puts the nodes off the voxel lattice: the convolution path handles it)
bool fused_projection_supported(nrh_f3d *h) { return !(h->velocity && h->velExt0) && plan_of(h)->ok; }

void fused_project(nrh_f3d *h, const float *g, float weight) {
    FusedPlan *p = plan_of(h);
    NRH_REQUIRE(p->ok, NRH_ERR_UNSUPPORTED, "separable node projection unavailable for this geometry");
    project_to_nodes(h, p, g, weight);
}

// ------------------------------------------------------------------ batches of pairs
// Every pair must take the fused path (integer spacing of at least 4 voxels, nodes on voxels).
void batch_fused_setup(nrh_batch *b) {
    std::vector<FusedArgs> args(b->P);
    std::vector<ProjXArgs> pxs(b->P);
    std::vector<ProjYArgs> pys(b->P);
    std::vector<NodeArgs> nds(b->P);
    std::vector<uint32_t> items;
    std::vector<int2> itX, itY, itN;
    b->sweepLds = b->projXLds = b->fixLds = 0;
    std::vector<int4> fixChunks;
    const int storage = b->pairs[0]->params.image_storage;
    for (int q = 0; q < b->P; ++q) {
        nrh_f3d *h = b->pairs[q];
        NRH_REQUIRE(fused_supported(h), NRH_ERR_UNSUPPORTED,
                    "batched pairs need the fused path (integer grid spacing >= 4 voxels, nodes on voxels)");
        NRH_REQUIRE(h->params.image_storage == storage, NRH_ERR_ARG, "batched pairs share one image storage type");
        FusedPlan *p = plan_of(h);
        NRH_REQUIRE(p->numTiles < (1 << 20) && b->P < 4096, NRH_ERR_UNSUPPORTED, "batch too large for the work list");
        FusedArgs a = make_args(h, p);
        a.hist = h->histCount.p;
        p->gvox.alloc(3 * h->N);
        a.gvox = p->gvox.p;
        const int nbins = a.rb * a.fb;
        p->kTab.alloc((size_t)(nbins + 3) / 4 * 4);
        a.kTab = p->kTab.p;
        args[q] = a;
        b->sweepLds = std::max(b->sweepLds, sweep_lds(a));
        for (int t = 0; t < p->numTiles; ++t) items.push_back(((uint32_t)q << 20) | (uint32_t)t);
        for (int t = 0; t < p->numTiles; t += kFixChunk)
            fixChunks.push_back(make_int4(q, t, std::min(kFixChunk, p->numTiles - t), 0));
        b->fixLds = std::max(b->fixLds, (size_t)nbins * sizeof(uint32_t));
        size_t ldsX = 0;
        projection_args(h, p, p->gvox.p, h->params.similarity_weight, pxs[q], pys[q], nds[q], &ldsX);
        b->projXLds = std::max(b->projXLds, ldsX);
        const int nbx = (int)((pxs[q].rows + pxs[q].rowsPerBlock - 1) / pxs[q].rowsPerBlock);
        for (int k = 0; k < nbx; ++k) itX.push_back(make_int2(q, k));
        const int nby = (int)(((size_t)pys[q].nz * pys[q].gny * pys[q].gnx + 255) / 256);
        for (int k = 0; k < nby; ++k) itY.push_back(make_int2(q, k));
        const int nbn = (int)((h->G + 255) / 256);
        for (int k = 0; k < nbn; ++k) itN.push_back(make_int2(q, k));
    }
    b->rb = args[0].rb;
    b->fb = args[0].fb;
    auto upload = [&](nrh::DevBuf<uint8_t>& dst, const void *src, size_t bytes) {
        dst.alloc(bytes);
        NRH_HIP(hipMemcpyAsync(dst.p, src, bytes, hipMemcpyHostToDevice, b->stream));
    };
    upload(b->fusedArgs, args.data(), args.size() * sizeof(FusedArgs));
    upload(b->projXArgs, pxs.data(), pxs.size() * sizeof(ProjXArgs));
    upload(b->projYArgs, pys.data(), pys.size() * sizeof(ProjYArgs));
    upload(b->nodeArgs, nds.data(), nds.size() * sizeof(NodeArgs));
    auto uploadItems = [&](nrh::DevBuf<int2>& dst, const std::vector<int2>& v, int& n) {
        dst.alloc(v.size());
        NRH_HIP(hipMemcpyAsync(dst.p, v.data(), v.size() * sizeof(int2), hipMemcpyHostToDevice, b->stream));
        n = (int)v.size();
    };
    uploadItems(b->itProjX, itX, b->nProjX);
    uploadItems(b->itProjY, itY, b->nProjY);
    uploadItems(b->itNodes, itN, b->nNodes);
    b->fixChunks.alloc(fixChunks.size());
    NRH_HIP(hipMemcpyAsync(b->fixChunks.p, fixChunks.data(), fixChunks.size() * sizeof(int4), hipMemcpyHostToDevice,
                           b->stream));
    b->nFixChunks = (int)fixChunks.size();
    b->sweepItems.alloc(items.size());
    NRH_HIP(hipMemcpyAsync(b->sweepItems.p, items.data(), items.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                           b->stream));
    b->nSweepItems = (int)items.size();
    if (b->sweepBlocks == 0) batch_set_sweep_blocks(b, b->tunedBlocksPerXcd > 0 ? b->tunedBlocksPerXcd : 128);
    NRH_HIP(hipStreamSynchronize(b->stream));
}

// The persistent grid: blocksPerXcd workgroups on each of the 8 XCDs, never more than the items.
void batch_set_sweep_blocks(nrh_batch *b, int blocksPerXcd) {
    const int perXcd = std::max(1, std::min(blocksPerXcd, (b->nSweepItems + 7) / 8));
    b->sweepBlocks = 8 * perXcd;
}

void batch_histogram(nrh_batch *b, const int *active) {
    const FusedArgs *A = (const FusedArgs *)b->fusedArgs.p;
    // the pairs' counts are zero here: allocation zeroes them and every finalisation consumes them
    const bool half = b->pairs[0]->params.image_storage == 1;
    if (half) {
        hipLaunchKernelGGL((k_sweep_batch<__half, false>), dim3(b->sweepBlocks), dim3(kThreads), b->sweepLds, b->stream,
                           A, b->sweepItems.p, b->nSweepItems, active);
        hipLaunchKernelGGL((k_sweep_fixup_batch<__half, false>), dim3(b->nFixChunks), dim3(256), b->fixLds, b->stream,
                           A, b->fixChunks.p, active);
    } else {
        hipLaunchKernelGGL((k_sweep_batch<float, false>), dim3(b->sweepBlocks), dim3(kThreads), b->sweepLds, b->stream,
                           A, b->sweepItems.p, b->nSweepItems, active);
        hipLaunchKernelGGL((k_sweep_fixup_batch<float, false>), dim3(b->nFixChunks), dim3(256), b->fixLds, b->stream,
                           A, b->fixChunks.p, active);
    }
    NRH_HIP(hipGetLastError());
    batch_finalize(b, active);
}

void batch_similarity_gradient(nrh_batch *b, const int *active) {
    const FusedArgs *A = (const FusedArgs *)b->fusedArgs.p;
    hipLaunchKernelGGL(k_nmi_table_batch, dim3((b->rb * b->fb + 255) / 256, b->P), dim3(256), 0, b->stream, A, active);
    const bool half = b->pairs[0]->params.image_storage == 1;
    if (half) {
        hipLaunchKernelGGL((k_sweep_batch<__half, true>), dim3(b->sweepBlocks), dim3(kThreads), b->sweepLds, b->stream,
                           A, b->sweepItems.p, b->nSweepItems, active);
        hipLaunchKernelGGL((k_sweep_fixup_batch<__half, true>), dim3(b->nFixChunks), dim3(256), 0, b->stream, A,
                           b->fixChunks.p, active);
    } else {
        hipLaunchKernelGGL((k_sweep_batch<float, true>), dim3(b->sweepBlocks), dim3(kThreads), b->sweepLds, b->stream,
                           A, b->sweepItems.p, b->nSweepItems, active);
        hipLaunchKernelGGL((k_sweep_fixup_batch<float, true>), dim3(b->nFixChunks), dim3(256), 0, b->stream, A,
                           b->fixChunks.p, active);
    }
    hipLaunchKernelGGL(k_project_x_batch, dim3(b->nProjX), dim3(256), b->projXLds, b->stream,
                       (const ProjXArgs *)b->projXArgs.p, b->itProjX.p, active);
    hipLaunchKernelGGL(k_project_y_batch, dim3(b->nProjY), dim3(256), 0, b->stream, (const ProjYArgs *)b->projYArgs.p,
                       b->itProjY.p, active);
    hipLaunchKernelGGL(k_project_nodes_batch, dim3(b->nNodes), dim3(256), 0, b->stream,
                       (const NodeArgs *)b->nodeArgs.p, b->itNodes.p, active);
    NRH_HIP(hipGetLastError());
}

}  // namespace nrh
