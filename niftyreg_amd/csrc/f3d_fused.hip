// Fused F3D evaluation for gfx950.
//
// One objective evaluation is one sweep over the reference grid (k_fused_pass<false>):
// B-spline field, trilinear warp and joint-histogram fill, with no field or warped image
// written to HBM.  One gradient evaluation is a second sweep (k_fused_pass<true>): the field
// and warp are recomputed together with the image gradient and the NMI voxel gradient, and
// the voxel gradient is projected straight onto the control points.
//
// Tiles.  A workgroup owns a 64 x 8 x 8 voxel tile (x fastest, one wave per 64-voxel row).
// The control points that influence the tile (a window of at most 20 x 6 x 6 nodes when the
// grid spacing is >= 4 voxels) are staged in LDS, and the cubic B-spline sum is evaluated
// separably: z contraction for the whole tile, y contraction per z slice, x contraction per
// voxel -- about 17 FMA per voxel and component instead of 64.
//
// Node gradient.  The reference convolves the voxel gradient with the cubic B-spline kernel
// of radius 2*spacing along x, y, z (each pass normalised by the in-image kernel sum) and
// samples the result at the nodes (reg-lib/Compute.cpp:230-281).  When the grid spacing is an
// integer number of voxels and the nodes fall on voxels, the weight of voxel v on node n is
// exactly the field's basis value B_{n - pre(v)}(t(v)) on each axis, divided by the per-axis
// normalisation at the node.  The gradient sweep therefore applies the transpose of the
// separable field evaluation to the voxel gradient (x, then y, then z) and writes one
// partial node window per tile; k_node_reduce sums the windows in a fixed order
// (deterministic), normalises, reorients and weights them.

#include "device.hpp"
#include "f3d_handle.hpp"
#include <cmath>
#include <map>
#include <mutex>

namespace nrh {

namespace {

constexpr int TX = 64, TY = 8, TZ = 8;
constexpr int WXM = 20, WYM = 6, WZM = 6;  // node-window bounds for spacing >= 4 voxels
constexpr int WIN = WXM * WYM * WZM;      // floats per component of one tile window
constexpr int kThreads = 256;

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
    const float *logTab;
    const double *ent;
    float *partials;
    float *fieldOut;  // optional: the deformation field of the sweep (planar [3][N]), for checking
};

template<class T> __device__ __forceinline__ float ld(const T *p, size_t i);
template<> __device__ __forceinline__ float ld<float>(const float *p, size_t i) { return p[i]; }
template<> __device__ __forceinline__ float ld<__half>(const __half *p, size_t i) { return __half2float(p[i]); }

// Tile index with the tiles of one XCD contiguous: consecutive workgroups are dealt round-robin
// to the 8 XCDs, so workgroup b runs on XCD b % 8; giving XCD k a contiguous block of tiles keeps
// neighbouring tiles (which read overlapping floating-image lines) in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int numTiles) {
    const int per = (numTiles + 7) / 8;
    const int t = (b % 8) * per + b / 8;
    return t;
}

template<class TImg, bool GRAD>
__global__ void __launch_bounds__(kThreads, 2) k_fused_pass(FusedArgs a) {
    // ---- LDS
    __shared__ float sCp[3][WZM][WYM][WXM];
    __shared__ float sAyz[TY][WXM][3];
    __shared__ float4 sBx[TX], sBy[TY], sBz[TZ];
    __shared__ int sPx[TX], sPy[TY], sPz[TZ];
    __shared__ int sXFirst[WXM], sXLast[WXM];
    extern __shared__ float dyn[];  // histogram (value) or log table (gradient)
    __shared__ float sG[GRAD ? TY : 1][GRAD ? TX : 1][3];
    __shared__ float sU1[GRAD ? TY : 1][GRAD ? WXM : 1][3];
    __shared__ float sOut[GRAD ? WZM : 1][GRAD ? WYM : 1][GRAD ? WXM : 1][3];

    const int tid = threadIdx.x;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    if (t >= a.numTiles) return;
    const int tx = t % a.tilesX, ty = (t / a.tilesX) % a.tilesY, tz = t / (a.tilesX * a.tilesY);
    const int x0 = tx * TX, y0 = ty * TY, z0 = tz * TZ;
    const int nX = min(TX, a.nx - x0), nY = min(TY, a.ny - y0), nZ = min(TZ, a.nz - z0);
    const TImg *ref = (const TImg *)a.ref;
    const TImg *flo = (const TImg *)a.flo;

    // ---- per-axis basis tables of the tile
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

    if (GRAD) {
        // K[r + w rb] = log p_r + log p_w - NMI log p_rw: the voxel NMI derivative is then the
        // bicubic contraction sum_r sum_w B(ref - r) B'(war - w) K[r, w] (see below)
        const float nmiF = (float)a.ent[4];
        const float *lj = a.logTab, *lr = a.logTab + a.rb * a.fb, *lw = lr + a.rb;
        for (int i = tid; i < a.rb * a.fb; i += kThreads) dyn[i] = (lr[i % a.rb] + lw[i / a.rb]) - nmiF * lj[i];
    } else {
        uint32_t *h = (uint32_t *)dyn;
        for (int i = tid; i < a.rb * a.fb; i += kThreads) h[i] = 0u;
    }
    // ---- control-point window
    for (int i = tid; i < 3 * WZ * WY * WX; i += kThreads) {
        const int ix = i % WX, iy = (i / WX) % WY, iz = (i / (WX * WY)) % WZ, c = i / (WX * WY * WZ);
        const double node = (double)a.cppXyz.m[c * 4] * (px0 + ix) + (double)a.cppXyz.m[c * 4 + 1] * (py0 + iy) +
                            (double)a.cppXyz.m[c * 4 + 2] * (pz0 + iz) + (double)a.cppXyz.m[c * 4 + 3];
        sCp[c][iz][iy][ix] =
            (float)((double)a.cpp[c * a.G + ((size_t)(pz0 + iz) * a.gny + (py0 + iy)) * a.gnx + (px0 + ix)] - node);
    }
    __syncthreads();
    if (GRAD && tid < WX) {
        // voxels of the tile whose 4-node support contains window node tid
        int f = nX, l = -1;
        for (int x = 0; x < nX; ++x) {
            const int o = tid - (sPx[x] - px0);
            if (o >= 0 && o < 4) { f = min(f, x); l = max(l, x); }
        }
        sXFirst[tid] = f;
        sXLast[tid] = l;
    }
    if (GRAD) {
        for (int i = tid; i < WZM * WYM * WXM * 3; i += kThreads) (&sOut[0][0][0][0])[i] = 0.f;
    }
    __syncthreads();

    const int lane = tid & 63, row = tid >> 6;
    double coefD = 0, nmiD = 0;
    if (GRAD) {
        nmiD = a.ent[4];
        coefD = 1.0 / (a.ent[2] * a.ent[3]);
    }
    const float coef = (float)coefD, nmi = (float)nmiD;
    const int nb2 = a.rb * a.fb;

    for (int z = 0; z < nZ; ++z) {
        // ---- stage B: z then y contraction of the window for this slice
        {
            const float4 bz = sBz[z];
            const int oz = sPz[z] - pz0;
            for (int i = tid; i < nY * WX; i += kThreads) {
                const int ix = i % WX, y = i / WX;
                const float4 b = sBy[y];
                const int o = sPy[y] - py0;
                for (int c = 0; c < 3; ++c) {
                    float az[4];
                    for (int j = 0; j < 4; ++j)
                        az[j] = bz.x * sCp[c][oz][o + j][ix] + bz.y * sCp[c][oz + 1][o + j][ix] +
                                bz.z * sCp[c][oz + 2][o + j][ix] + bz.w * sCp[c][oz + 3][o + j][ix];
                    sAyz[y][ix][c] = b.x * az[0] + b.y * az[1] + b.z * az[2] + b.w * az[3];
                }
            }
        }
        __syncthreads();
        // ---- per voxel: field, warp (+ gradient)
        for (int y = row; y < TY; y += kThreads / 64) {
            const int x = lane;
            float g0 = 0.f, g1 = 0.f, g2 = 0.f;
            if (x < nX && y < nY) {
                const size_t idx = ((size_t)(z0 + z) * a.ny + (y0 + y)) * a.nx + (x0 + x);
                const bool active = !a.mask || a.mask[idx] > -1;
                if (active) {
                    const float4 b = sBx[x];
                    const int o = sPx[x] - px0;
                    // field = identity + displacement: the B-spline sum of the node positions is
                    // the affine map of (pre + t + 1) on each axis (cubic B-splines reproduce
                    // linear functions), evaluated in double; the window holds displacements.
                    const float4 by = sBy[y], bz = sBz[z];
                    const double ux = (double)sPx[x] + (double)(b.y + 2.f * b.z + 3.f * b.w);
                    const double uy = (double)sPy[y] + (double)(by.y + 2.f * by.z + 3.f * by.w);
                    const double uz = (double)sPz[z] + (double)(bz.y + 2.f * bz.z + 3.f * bz.w);
                    float w[3];
                    for (int c = 0; c < 3; ++c) {
                        const float d = b.x * sAyz[y][o][c] + b.y * sAyz[y][o + 1][c] + b.z * sAyz[y][o + 2][c] +
                                        b.w * sAyz[y][o + 3][c];
                        const double id = (double)a.cppXyz.m[c * 4] * ux + (double)a.cppXyz.m[c * 4 + 1] * uy +
                                          (double)a.cppXyz.m[c * 4 + 2] * uz + (double)a.cppXyz.m[c * 4 + 3];
                        w[c] = (float)(id + (double)d);
                    }
                    if (a.fieldOut) {
                        a.fieldOut[idx] = w[0];
                        a.fieldOut[idx + (size_t)a.nx * a.ny * a.nz] = w[1];
                        a.fieldOut[idx + 2 * (size_t)a.nx * a.ny * a.nz] = w[2];
                    }
                    const float3 p = affine_apply_f(a.floIjk, w[0], w[1], w[2]);
                    const int fx = floor_int(p.x), fy = floor_int(p.y), fz = floor_int(p.z);
                    float rx = p.x - (float)fx, ry = p.y - (float)fy, rz = p.z - (float)fz;
                    float v[2][2][2];
                    const bool inside = fx >= 0 && fx + 1 < a.fnx && fy >= 0 && fy + 1 < a.fny && fz >= 0 && fz + 1 < a.fnz;
                    if (inside) {
                        const size_t base = ((size_t)fz * a.fny + fy) * a.fnx + fx;
                        const size_t sy = a.fnx, sz = (size_t)a.fnx * a.fny;
                        v[0][0][0] = ld(flo, base);          v[0][0][1] = ld(flo, base + 1);
                        v[0][1][0] = ld(flo, base + sy);     v[0][1][1] = ld(flo, base + sy + 1);
                        v[1][0][0] = ld(flo, base + sz);     v[1][0][1] = ld(flo, base + sz + 1);
                        v[1][1][0] = ld(flo, base + sz + sy); v[1][1][1] = ld(flo, base + sz + sy + 1);
                    } else {
                        for (int c = 0; c < 2; ++c)
                            for (int bb = 0; bb < 2; ++bb)
                                for (int aa = 0; aa < 2; ++aa) {
                                    const int X = fx + aa, Y = fy + bb, Z = fz + c;
                                    v[c][bb][aa] = (X >= 0 && X < a.fnx && Y >= 0 && Y < a.fny && Z >= 0 && Z < a.fnz)
                                        ? ld(flo, ((size_t)Z * a.fny + Y) * a.fnx + X) : a.pad;
                                }
                    }
                    if (rx < 0) rx = 0;
                    if (ry < 0) ry = 0;
                    if (rz < 0) rz = 0;
                    const float x0b = 1.f - rx, y0b = 1.f - ry, z0b = 1.f - rz;
                    const float e00 = v[0][0][0] * x0b + v[0][0][1] * rx, e01 = v[0][1][0] * x0b + v[0][1][1] * rx;
                    const float e10 = v[1][0][0] * x0b + v[1][0][1] * rx, e11 = v[1][1][0] * x0b + v[1][1][1] * rx;
                    const float f0 = e00 * y0b + e01 * ry, f1 = e10 * y0b + e11 * ry;
                    const float war = f0 * z0b + f1 * rz;
                    const float rv = ld(ref, idx);
                    if (!GRAD) {
                        if (rv == rv && war == war && rv >= 0.f && rv < (float)a.rb && war >= 0.f && war < (float)a.fb)
                            atomicAdd(&((uint32_t *)dyn)[(int)rv + (int)war * a.rb], 1u);
                    } else if (rv == rv && war == war) {
                        // image gradient in floating voxel units (TrilinearImageGradient)
                        const float d00 = v[0][0][1] - v[0][0][0], d01 = v[0][1][1] - v[0][1][0];
                        const float d10 = v[1][0][1] - v[1][0][0], d11 = v[1][1][1] - v[1][1][0];
                        float gx = (d00 * y0b + d01 * ry) * z0b + (d10 * y0b + d11 * ry) * rz;
                        float gy = (e01 - e00) * z0b + (e11 - e10) * rz;
                        float gz = f1 - f0;
                        if (gx != gx) gx = 0.f;
                        if (gy != gy) gy = 0.f;
                        if (gz != gz) gz = 0.f;
                        // NMI derivative (reg_getVoxelBasedNmiGradient3d, reg-lib/cpu/_reg_nmi.cpp:501-593):
                        //   (sum_r B_r log p_r)(sum_w B'_w) + (sum_r B_r)(sum_w B'_w log p_w)
                        //   - NMI sum_r sum_w B_r B'_w log p_rw  =  sum_r sum_w B_r B'_w K[r, w]
                        // over the 4 bins around each intensity (bins outside the histogram get
                        // weight 0, as the reference skips them).  B_r: cubic B-spline weights of the
                        // reference intensity's fraction; B'_w: their derivative for the warped one.
                        const float rfl = floorf(rv), wfl = floorf(war);
                        const float f = rv - rfl, gg = war - wfl;
                        const float f2 = f * f, f3 = f2 * f, omf = 1.f - f;
                        float br[4] = { omf * omf * omf * (1.f / 6.f), 0.66666669f - f2 + 0.5f * f3,
                                        (-3.f * f3 + 3.f * f2 + 3.f * f + 1.f) * (1.f / 6.f), f3 * (1.f / 6.f) };
                        const float omg = 1.f - gg;
                        float dw[4] = { -0.5f * omg * omg, (1.5f * gg - 2.f) * gg, (0.5f + 1.5f * gg) * omg,
                                        0.5f * gg * gg };
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
                        float S = 0.f;
#pragma unroll
                        for (int j = 0; j < 4; ++j) {
                            const float T = br[0] * dyn[ri[0] + wc[j]] + br[1] * dyn[ri[1] + wc[j]] +
                                            br[2] * dyn[ri[2] + wc[j]] + br[3] * dyn[ri[3] + wc[j]];
                            S += dw[j] * T;
                        }
                        S *= coef;
                        g0 = S * gx;
                        g1 = S * gy;
                        g2 = S * gz;
                    }
                }
            }
            if (GRAD) {
                sG[y][x][0] = g0;
                sG[y][x][1] = g1;
                sG[y][x][2] = g2;
            }
        }
        if (GRAD) {
            __syncthreads();
            // ---- stage C^T: x adjoint, per (y, window node)
            for (int i = tid; i < nY * WX; i += kThreads) {
                const int ix = i % WX, y = i / WX;
                float s0 = 0.f, s1 = 0.f, s2 = 0.f;
                for (int x = sXFirst[ix]; x <= sXLast[ix]; ++x) {
                    const int o = ix - (sPx[x] - px0);
                    const float4 b = sBx[x];
                    const float w = o == 0 ? b.x : o == 1 ? b.y : o == 2 ? b.z : b.w;
                    s0 += w * sG[y][x][0];
                    s1 += w * sG[y][x][1];
                    s2 += w * sG[y][x][2];
                }
                sU1[y][ix][0] = s0;
                sU1[y][ix][1] = s1;
                sU1[y][ix][2] = s2;
            }
            __syncthreads();
            // ---- stage B^T: y adjoint of the slice, then its z-adjoint share added to the
            // tile's node window (slices in z order, one owner thread per (node y, node x))
            const float4 bzz = sBz[z];
            const int ozz = sPz[z] - pz0;
            for (int i = tid; i < WY * WX; i += kThreads) {
                const int ix = i % WX, iy = i / WX;
                float s0 = 0.f, s1 = 0.f, s2 = 0.f;
                for (int y = 0; y < nY; ++y) {
                    const int o = iy - (sPy[y] - py0);
                    if (o < 0 || o > 3) continue;
                    const float4 b = sBy[y];
                    const float w = o == 0 ? b.x : o == 1 ? b.y : o == 2 ? b.z : b.w;
                    s0 += w * sU1[y][ix][0];
                    s1 += w * sU1[y][ix][1];
                    s2 += w * sU1[y][ix][2];
                }
                const float bzk[4] = { bzz.x, bzz.y, bzz.z, bzz.w };
                for (int k = 0; k < 4; ++k) {
                    sOut[ozz + k][iy][ix][0] += bzk[k] * s0;
                    sOut[ozz + k][iy][ix][1] += bzk[k] * s1;
                    sOut[ozz + k][iy][ix][2] += bzk[k] * s2;
                }
            }
        }
        __syncthreads();
    }

    if (!GRAD) {
        const uint32_t *h = (const uint32_t *)dyn;
        for (int i = tid; i < nb2; i += kThreads)
            if (h[i]) atomicAdd(&a.hist[i], h[i]);
    } else {
        // ---- the tile's partial node window
        float *out = a.partials + (size_t)t * 3 * WIN;
        for (int i = tid; i < WZ * WY * WX; i += kThreads) {
            const int ix = i % WX, iy = (i / WX) % WY, iz = i / (WX * WY);
            const int wi = (iz * WYM + iy) * WXM + ix;
            out[wi] = sOut[iz][iy][ix][0];
            out[WIN + wi] = sOut[iz][iy][ix][1];
            out[2 * WIN + wi] = sOut[iz][iy][ix][2];
        }
    }
}

// Sum of the tile windows covering each node (fixed tile order), per-axis normalisation,
// reorientation by the floating image's mm->voxel 3x3 (transposed), weighting; nodes whose
// voxel position lies outside the reference image receive 0 (reg_voxelCentricToNodeCentric
// samples nothing there).
struct ReduceArgs {
    const float *partials;
    int gnx, gny, gnz;
    size_t G;
    int tilesX, tilesY;
    const int2 *tx, *ty, *tz;          // first / last tile covering each node index, per axis
    const int *wx0, *wy0, *wz0;        // window origin (node index) of each tile, per axis
    const float *invNx, *invNy, *invNz;  // 1 / kernel sum at the node, 0 outside the image
    Affine reorient;
    float weight;
    float *nodeGrad;
};

__global__ void __launch_bounds__(256) k_node_reduce(ReduceArgs r) {
    const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= r.G) return;
    const int i = (int)(n % r.gnx), j = (int)((n / r.gnx) % r.gny), k = (int)(n / ((size_t)r.gnx * r.gny));
    const float inv = r.invNx[i] * r.invNy[j] * r.invNz[k];
    float u0 = 0.f, u1 = 0.f, u2 = 0.f;
    if (inv != 0.f) {
        const int2 ax = r.tx[i], ay = r.ty[j], az = r.tz[k];
        const int na = ax.y - ax.x + 1, nab = na * (ay.y - ay.x + 1);
        const int cnt = nab * (az.y - az.x + 1);
        // windows in (z, y, x) tile order, loaded four at a time so the loads overlap; the sums
        // keep that order
        for (int q0 = 0; q0 < cnt; q0 += 4) {
            float v[4][3];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int q = q0 + u;
                if (q < cnt) {
                    const int a = ax.x + q % na, b = ay.x + (q % nab) / na, c = az.x + q / nab;
                    const size_t tile = ((size_t)c * r.tilesY + b) * r.tilesX + a;
                    const int wi = ((k - r.wz0[c]) * WYM + (j - r.wy0[b])) * WXM + (i - r.wx0[a]);
                    const float *p = r.partials + tile * 3 * WIN + wi;
                    v[u][0] = p[0];
                    v[u][1] = p[WIN];
                    v[u][2] = p[2 * WIN];
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (q0 + u < cnt) {
                    u0 += v[u][0];
                    u1 += v[u][1];
                    u2 += v[u][2];
                }
            }
        }
        u0 *= inv;
        u1 *= inv;
        u2 *= inv;
    }
    const float *R = r.reorient.m;
    r.nodeGrad[n] = (R[0] * u0 + R[4] * u1 + R[8] * u2) * r.weight;
    r.nodeGrad[n + r.G] = (R[1] * u0 + R[5] * u1 + R[9] * u2) * r.weight;
    r.nodeGrad[n + 2 * r.G] = (R[2] * u0 + R[6] * u1 + R[10] * u2) * r.weight;
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
    DevBuf<int2> tX, tY, tZ;
    DevBuf<int> wx0, wy0, wz0;
    DevBuf<float> invNx, invNy, invNz;
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
        DevBuf<int2> *tt[3] = { &p->tX, &p->tY, &p->tZ };
        DevBuf<int> *w0[3] = { &p->wx0, &p->wy0, &p->wz0 };
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
            std::vector<int> org(tiles[ax]);
            std::vector<int2> cover(gdims[ax], make_int2(1, 0));
            for (int t = 0; t < tiles[ax]; ++t) {
                const int a0 = t * tsz[ax], a1 = std::min(n, a0 + tsz[ax]) - 1;
                org[t] = hpre[a0];
                const int wn = hpre[a1] - hpre[a0] + 4;
                if (wn > wmax[ax]) ok = false;
                for (int q = hpre[a0]; q < hpre[a0] + wn && q < gdims[ax]; ++q) {
                    if (cover[q].x > cover[q].y) cover[q] = make_int2(t, t);
                    else cover[q].y = t;
                }
            }
            w0[ax]->alloc(tiles[ax]);
            tt[ax]->alloc(gdims[ax]);
            NRH_HIP(hipMemcpyAsync(w0[ax]->p, org.data(), org.size() * sizeof(int), hipMemcpyHostToDevice, h->stream));
            NRH_HIP(hipMemcpyAsync(tt[ax]->p, cover.data(), cover.size() * sizeof(int2), hipMemcpyHostToDevice, h->stream));
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
        NRH_HIP(hipStreamSynchronize(h->stream));
    }
    std::lock_guard<std::mutex> lk(g_planMutex);
    g_planMap[h] = p;
    return p;
}

bool fused_supported(nrh_f3d *h) { return plan_of(h)->ok; }

void collect_sweep_times(nrh_f3d *h) {
    for (int k = 0; k < 2; ++k) {
        if (!h->evPending[k]) continue;
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
    return a;
}

static unsigned grid_for(int numTiles) { return (unsigned)((numTiles + 7) / 8 * 8); }

void fused_histogram(nrh_f3d *h, float *fieldOut) {
    FusedPlan *p = plan_of(h);
    NRH_REQUIRE(p->ok, NRH_ERR_UNSUPPORTED, "fused path unavailable for this geometry");
    FusedArgs a = make_args(h, p);
    a.hist = h->histCount.p;
    a.fieldOut = fieldOut;
    op_fill_u32(h->histCount.p, (size_t)a.rb * a.fb, 0u, h->stream);
    const size_t lds = (size_t)a.rb * a.fb * sizeof(uint32_t);
    if (h->profile) NRH_HIP(hipEventRecord(h->ev[0], h->stream));
    if (h->params.image_storage == 1)
        hipLaunchKernelGGL((k_fused_pass<__half, false>), dim3(grid_for(p->numTiles)), dim3(kThreads), lds, h->stream, a);
    else
        hipLaunchKernelGGL((k_fused_pass<float, false>), dim3(grid_for(p->numTiles)), dim3(kThreads), lds, h->stream, a);
    NRH_HIP(hipGetLastError());
    if (h->profile) {
        NRH_HIP(hipEventRecord(h->ev[1], h->stream));
        h->evPending[0] = true;
    }
    op_nmi_finalize(h);
    h->histVersion = h->cppVersion;
}

void fused_gradient(nrh_f3d *h, float weight) {
    FusedPlan *p = plan_of(h);
    NRH_REQUIRE(p->ok, NRH_ERR_UNSUPPORTED, "fused path unavailable for this geometry");
    FusedArgs a = make_args(h, p);
    a.logTab = h->histLogF.p;
    a.ent = h->ent.p;
    h->nodePartials.alloc((size_t)p->numTiles * 3 * WIN);
    a.partials = h->nodePartials.p;
    const size_t lds = (size_t)a.rb * a.fb * sizeof(float);
    if (h->profile) NRH_HIP(hipEventRecord(h->ev[2], h->stream));
    if (h->params.image_storage == 1)
        hipLaunchKernelGGL((k_fused_pass<__half, true>), dim3(grid_for(p->numTiles)), dim3(kThreads), lds, h->stream, a);
    else
        hipLaunchKernelGGL((k_fused_pass<float, true>), dim3(grid_for(p->numTiles)), dim3(kThreads), lds, h->stream, a);
    NRH_HIP(hipGetLastError());
    if (h->profile) {
        NRH_HIP(hipEventRecord(h->ev[3], h->stream));
        h->evPending[1] = true;
    }
    ReduceArgs r{};
    r.partials = h->nodePartials.p;
    r.gnx = h->cpp.nx; r.gny = h->cpp.ny; r.gnz = h->cpp.nz;
    r.G = h->G;
    r.tilesX = p->tilesX; r.tilesY = p->tilesY;
    r.tx = p->tX.p; r.ty = p->tY.p; r.tz = p->tZ.p;
    r.wx0 = p->wx0.p; r.wy0 = p->wy0.p; r.wz0 = p->wz0.p;
    r.invNx = p->invNx.p; r.invNy = p->invNy.p; r.invNz = p->invNz.p;
    r.reorient = to_affine(h->flo.ijk);
    r.weight = node_weight_ratio(h, weight);
    r.nodeGrad = h->nodeGrad.p;
    hipLaunchKernelGGL(k_node_reduce, dim3((unsigned)((h->G + 255) / 256)), dim3(256), 0, h->stream, r);
    NRH_HIP(hipGetLastError());
}

}  // namespace nrh
