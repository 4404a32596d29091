// Host side of a full F3D registration: image pyramid, control-point grid creation and
// refinement, and the conjugate-gradient / line-search loop of reg_base<T>::Run
// (reg-lib/_reg_base.cpp:727-819, reg-lib/Optimiser.cpp:129-317, reg-lib/_reg_f3d.cpp:92-159).
// Every objective and gradient evaluation runs on the device through the F3D handle.

#include "f3d_handle.hpp"
#include "f3d_batch.hpp"
#include "host.hpp"
#include <algorithm>
#include <cmath>
#include <limits>
#include <memory>
#include <chrono>
#include <thread>

namespace nrh {

namespace {

// Gaussian smoothing of reg_tools_kernelConvolution (reg-lib/cpu/_reg_tools.cpp:832-1112) with all
// three axes in one call: intensity and density are convolved axis after axis and divided once.
// Each line sum follows the reference's default x86 build (USE_SSE with SSE3, :1017-1047): float
// products accumulated in four float lanes over groups of four taps (lane j takes tap 3 - j of
// each group), the lanes reduced as (l0 + l1) + (l2 + l3), then the remaining taps added in double.
void gaussian_smooth(const nrh_geom& g, std::vector<float>& img, float sigmaVox) {
#pragma clang fp contract(off)
    const int dims[3] = { g.nx, g.ny, g.nz };
    const size_t N = voxels(g);
    std::vector<float> dens(N);
    std::vector<uint8_t> isnan_(N);
    for (size_t i = 0; i < N; ++i) {
        const bool ok = img[i] == img[i];
        dens[i] = ok ? 1.f : 0.f;
        isnan_[i] = !ok;
        if (!ok) img[i] = 0.f;
    }
    const double temp = std::fabs((double)sigmaVox);
    const int radius = (int)(temp * 3.0f);
    std::vector<float> kern(2 * radius + 1);
    for (int i = -radius; i <= radius; ++i)
        kern[i + radius] = (float)(std::exp(-(double)(i * i) / (2.0 * (temp * temp))) / (temp * 2.506628274631));
    std::vector<float> bi, bd;
    for (int n = 0; n < 3; ++n) {
        if (dims[n] <= 1 || radius <= 0) continue;
        const int L = dims[n];
        const size_t stride = n == 0 ? 1 : n == 1 ? (size_t)dims[0] : (size_t)dims[0] * dims[1];
        const size_t planes = N / L;
        bi.resize(L);
        bd.resize(L);
        for (size_t p = 0; p < planes; ++p) {
            size_t start;
            if (n == 0) start = p * dims[0];
            else if (n == 1) start = (p / dims[0]) * (size_t)dims[0] * dims[1] + p % dims[0];
            else start = p;
            for (int l = 0; l < L; ++l) {
                bi[l] = img[start + l * stride];
                bd[l] = dens[start + l * stride];
            }
            for (int l = 0; l < L; ++l) {
                int pre = l - radius, pst = l + radius + 1, k0 = 0;
                if (pre < 0) { k0 = -pre; pre = 0; }
                if (pst > L) pst = L;
                float li[4] = { 0.f, 0.f, 0.f, 0.f }, ld[4] = { 0.f, 0.f, 0.f, 0.f };
                int k = pre, kk = k0;
                for (; k < pst - 3; k += 4, kk += 4)
                    for (int j = 0; j < 4; ++j) {
                        li[3 - j] = kern[kk + j] * bi[k + j] + li[3 - j];
                        ld[3 - j] = kern[kk + j] * bd[k + j] + ld[3 - j];
                    }
                double is = (double)((li[0] + li[1]) + (li[2] + li[3]));
                double ds = (double)((ld[0] + ld[1]) + (ld[2] + ld[3]));
                for (; k < pst; ++k, ++kk) {
                    is += (double)(kern[kk] * bi[k]);
                    ds += (double)(kern[kk] * bd[k]);
                }
                img[start + l * stride] = (float)is;
                dens[start + l * stride] = (float)ds;
            }
        }
    }
    for (size_t i = 0; i < N; ++i) img[i] = isnan_[i] ? std::numeric_limits<float>::quiet_NaN() : img[i] / dens[i];
}

}  // namespace

// reg_downsampleImage<float>(image, 1, axes) (reg-lib/cpu/_reg_tools.cpp:1343-1467): Gaussian
// smoothing (sigma 0.7355 voxel) then nearest-neighbour decimation by 2 along the flagged axes.
// The new voxel -> mm map doubles the flagged columns and keeps the origin.
void downsample(const nrh_geom& g, const std::vector<float>& in, const bool axes[3], nrh_geom& out,
                std::vector<float>& res) {
    std::vector<float> img(in);
    gaussian_smooth(g, img, -0.7355f);
    const int od[3] = { g.nx, g.ny, g.nz };
    int nd[3];
    float pd[3] = { g.dx, g.dy, g.dz };
    for (int i = 0; i < 3; ++i) {
        nd[i] = od[i];
        if (od[i] > 1 && axes[i]) nd[i] = (int)std::ceil(od[i] / 2.0);
        if (axes[i]) pd[i] *= 2.f;
    }
    out = g;
    out.nx = nd[0]; out.ny = nd[1]; out.nz = nd[2];
    out.dx = pd[0]; out.dy = pd[1]; out.dz = pd[2];
    for (int c = 0; c < 3; ++c)
        if (axes[c])
            for (int r = 0; r < 3; ++r) out.xyz[r * 4 + c] = g.xyz[r * 4 + c] * 2.f;
    mat44_inverse(out.xyz, out.ijk);
    const float *m = out.xyz, *r2v = g.ijk;
    auto round_int = [](float x) { return (int)(x + (x >= 0 ? 0.5f : -0.5f)); };
    res.assign(voxels(out), 0.f);
    size_t idx = 0;
    for (int z = 0; z < nd[2]; ++z)
        for (int y = 0; y < nd[1]; ++y)
            for (int x = 0; x < nd[0]; ++x, ++idx) {
#pragma clang fp contract(off)
                float real[3];
                for (int r = 0; r < 3; ++r)
                    real[r] = (float)x * m[r * 4] + (float)y * m[r * 4 + 1] + (float)z * m[r * 4 + 2] + m[r * 4 + 3];
                int pos[3];
                for (int r = 0; r < 3; ++r)
                    pos[r] = round_int(real[0] * r2v[r * 4] + real[1] * r2v[r * 4 + 1] + real[2] * r2v[r * 4 + 2] + r2v[r * 4 + 3]);
                if (od[2] == 1) pos[2] = 0;
                float v = std::numeric_limits<float>::quiet_NaN();
                if (pos[0] > -1 && pos[0] < od[0] && pos[1] > -1 && pos[1] < od[1] && pos[2] > -1 && pos[2] < od[2])
                    v = img[((size_t)pos[2] * od[1] + pos[1]) * od[0] + pos[0]];
                res[idx] = v;
            }
}

// reg_spline_refineControlPointGrid3D (reg-lib/cpu/_reg_localTrans.cpp:1809-2120): cubic B-spline
// subdivision.  Node 2x-1 of the new grid takes (1 6 1)/8 of old nodes x-1, x, x+1 and node 2x
// takes (1 1)/2 of x, x+1 on each axis; old nodes outside the grid count as 0.  The new header is
// the control-point grid of the level's reference image at half the spacing.
static void subdivide(const nrh_geom& old, const std::vector<float>& oldData, const nrh_geom& grid,
                      std::vector<float>& data);

void refine_grid(const nrh_geom& ref, const nrh_geom& old, const std::vector<float>& oldData, nrh_geom& grid,
                 std::vector<float>& data) {
    const float sp[3] = { old.dx / 2.0f, old.dy / 2.0f, old.dz / 2.0f };
    create_grid(ref, sp, grid, nullptr);
    subdivide(old, oldData, grid, data);
}

// reg_spline_refineControlPointGrid3D without a reference image (reg-lib/cpu/_reg_localTrans.cpp:
// 1825-1829, 2218-2236), as reg_f3d2 refines its velocity grids: (n - 3) * 2 + 3 nodes per axis,
// the sform's columns halved and its origin moved to the new node (1, 1, 1) of the halved map.
void refine_grid_self(const nrh_geom& old, const std::vector<float>& oldData, nrh_geom& grid,
                      std::vector<float>& data) {
    grid = old;
    grid.nx = (old.nx - 3) * 2 + 3;
    grid.ny = (old.ny - 3) * 2 + 3;
    grid.nz = (old.nz - 3) * 2 + 3;
    grid.dx = old.dx / 2.0f;
    grid.dy = old.dy / 2.0f;
    grid.dz = old.dz / 2.0f;
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r)
            if (r < 2 || old.nz > 1) grid.xyz[r * 4 + c] = old.xyz[r * 4 + c] / 2.f;
    for (int r = 0; r < 3; ++r) {
        const double o = (double)grid.xyz[r * 4] + (double)grid.xyz[r * 4 + 1] + (double)grid.xyz[r * 4 + 2] +
                         (double)grid.xyz[r * 4 + 3];
        if (r < 2 || old.nz > 1) grid.xyz[r * 4 + 3] = (float)o;
    }
    mat44_inverse(grid.xyz, grid.ijk);
    subdivide(old, oldData, grid, data);
}

// The cubic B-spline subdivision of reg_spline_refineControlPointGrid3D's data loops onto the
// new grid `grid` (reg-lib/cpu/_reg_localTrans.cpp:1838-2110).
static void subdivide(const nrh_geom& old, const std::vector<float>& oldData, const nrh_geom& grid,
                      std::vector<float>& data) {
    const int on[3] = { old.nx, old.ny, old.nz };
    const size_t OG = voxels(old), NG = voxels(grid);
    data.assign(3 * NG, 0.f);
    auto get = [&](const float *a, int x, int y, int z) -> float {
        if (x < 0 || x >= on[0] || y < 0 || y >= on[1] || z < 0 || z >= on[2]) return 0.f;
        return a[((size_t)z * on[1] + y) * on[0] + x];
    };
    // per-axis taps: even child (offsets -1, 0, 1; weights 1, 6, 1), odd child (offsets 0, 1; 1, 1)
    static const int offE[3] = { -1, 0, 1 }, wE[3] = { 1, 6, 1 };
    static const int offO[2] = { 0, 1 }, wO[2] = { 1, 1 };
    for (int c = 0; c < 3; ++c) {
#pragma clang fp contract(off)
        const float *src = oldData.data() + c * OG;
        float *dst = data.data() + c * NG;
        for (int z = 0; z < on[2]; ++z)
            for (int y = 0; y < on[1]; ++y)
                for (int x = 0; x < on[0]; ++x)
                    for (int cz = 0; cz < 2; ++cz)
                        for (int cy = 0; cy < 2; ++cy)
                            for (int cx = 0; cx < 2; ++cx) {
                                const int X = 2 * x - 1 + cx, Y = 2 * y - 1 + cy, Z = 2 * z - 1 + cz;
                                if (X < 0 || Y < 0 || Z < 0 || X >= grid.nx || Y >= grid.ny || Z >= grid.nz) continue;
                                const int nx_ = cx ? 2 : 3, ny_ = cy ? 2 : 3, nz_ = cz ? 2 : 3;
                                const int *ox = cx ? offO : offE, *oy = cy ? offO : offE, *oz = cz ? offO : offE;
                                const int *wx = cx ? wO : wE, *wy = cy ? wO : wE, *wz = cz ? wO : wE;
                                // group the taps by weight as the reference writes them (1, 6, 36, 216)
                                float sums[4] = { 0.f, 0.f, 0.f, 0.f };
                                for (int k = 0; k < nz_; ++k)
                                    for (int j = 0; j < ny_; ++j)
                                        for (int i = 0; i < nx_; ++i) {
                                            const int w = wx[i] * wy[j] * wz[k];
                                            const int cls = w == 1 ? 0 : w == 6 ? 1 : w == 36 ? 2 : 3;
                                            sums[cls] += get(src, x + ox[i], y + oy[j], z + oz[k]);
                                        }
                                const float den = (float)((cx ? 2 : 8) * (cy ? 2 : 8) * (cz ? 2 : 8));
                                dst[((size_t)Z * grid.ny + Y) * grid.nx + X] =
                                    (sums[0] + 6.0f * sums[1] + 36.0f * sums[2] + 216.0f * sums[3]) / den;
                            }
    }
}

// reg_createImagePyramid<float> (reg-lib/cpu/_reg_tools.cpp:1705-1733): the finest level is the
// input; each coarser level halves the axes whose half-size is still >= 32 voxels.
std::vector<Level> pyramid(const nrh_geom& g, const float *data, int levels, int toPerform) {
    std::vector<Level> p(toPerform);
    Level cur{ g, std::vector<float>(data, data + voxels(g)) };
    auto axes_of = [](const nrh_geom& q, bool ax[3]) {
        ax[0] = q.nx / 2 >= 32;
        ax[1] = q.ny / 2 >= 32;
        ax[2] = q.nz / 2 >= 32;
    };
    for (int l = toPerform; l < levels; ++l) {
        bool ax[3];
        axes_of(cur.geom, ax);
        Level nxt;
        downsample(cur.geom, cur.data, ax, nxt.geom, nxt.data);
        cur = std::move(nxt);
    }
    p[toPerform - 1] = cur;
    for (int l = toPerform - 2; l >= 0; --l) {
        bool ax[3];
        axes_of(p[l + 1].geom, ax);
        downsample(p[l + 1].geom, p[l + 1].data, ax, p[l].geom, p[l].data);
    }
    return p;
}

namespace {

struct Handle {
    nrh_f3d *h = nullptr;
    ~Handle() { if (h) nrh_f3d_destroy(h); }
};

void check(int rc) {
    if (rc != NRH_OK) throw Error(rc, nrh_last_error());
}

}  // namespace

// The host side of one reg_f3d registration that does not depend on the device: weights,
// pyramids, the control-point grid of every level, and the iteration budget.
struct PairRun {
    nrh_f3d_config cfg;
    int levels = 1, toPerform = 1;
    float be = 0, le = 0, simW = 0;
    std::vector<Level> rp, fp;
    nrh_geom grid;
    std::vector<float> cpp;
    size_t maxIter = 0;
    double finalNmi = 0;
    int totalIterations = 0;
};

void pair_begin(PairRun& r, const nrh_geom& refG, const float *refData, const nrh_geom& floG, const float *floData,
                const nrh_f3d_config& cfg) {
    NRH_REQUIRE(cfg.levels >= 1 && cfg.bins >= 1 && cfg.max_iterations >= 0, NRH_ERR_ARG, "invalid configuration");
    r.cfg = cfg;
    const int levels = cfg.levels;
    int toPerform = cfg.levels_to_perform > 0 ? std::min(cfg.levels_to_perform, levels) : levels;
    if (!cfg.use_pyramid) toPerform = levels;
    r.levels = levels;
    r.toPerform = toPerform;
    // reg_f3d::CheckParameters: weights of the objective (reg-lib/_reg_f3d.cpp:142-159)
    float be = cfg.be_weight, le = cfg.le_weight;
    const float penalty = be + le;
    if (penalty >= 1) {
        r.simW = 0;
        be /= penalty;
        le /= penalty;
    } else {
        r.simW = 1 - penalty;
    }
    r.be = be;
    r.le = le;
    // pyramids (or the full-resolution images at every level with -nopy)
    if (cfg.use_pyramid) {
        r.rp = pyramid(refG, refData, levels, toPerform);
        r.fp = pyramid(floG, floData, levels, toPerform);
    } else {
        r.rp.assign(1, Level{ refG, std::vector<float>(refData, refData + voxels(refG)) });
        r.fp.assign(1, Level{ floG, std::vector<float>(floData, floData + voxels(floG)) });
    }
    // control-point grid at the coarsest level (reg-lib/_reg_f3d.cpp:161-195)
    float spmm = cfg.spacing < 0 ? -cfg.spacing * refG.dx : cfg.spacing;
    float spy = cfg.spacing < 0 ? -cfg.spacing * refG.dy : cfg.spacing;
    float spz = cfg.spacing < 0 ? -cfg.spacing * refG.dz : cfg.spacing;
    const float scale = std::pow(2.f, (float)(levels - 1));
    const float gsp[3] = { spmm * scale, spy * scale, spz * scale };
    const nrh_geom& g0 = r.rp[0].geom;
    create_grid(g0, gsp, r.grid, nullptr);
    r.cpp.resize(3 * voxels(r.grid));
    create_grid(g0, gsp, r.grid, r.cpp.data());
    if (cfg.affine) {
        // reg_affine_getDeformationField(affine, grid) (reg-lib/_reg_f3d.cpp:190-191,
        // reg-lib/cpu/_reg_globalTrans.cpp:74-133): node positions A S i in double
        float T[16];
        mat44_product(cfg.affine, r.grid.xyz, T);
        const size_t G = voxels(r.grid);
        size_t i = 0;
        for (int z = 0; z < r.grid.nz; ++z)
            for (int y = 0; y < r.grid.ny; ++y)
                for (int x = 0; x < r.grid.nx; ++x, ++i)
                    for (int c = 0; c < 3; ++c) {
                        volatile double v = (double)T[c * 4] * (double)x;
                        v = v + (double)T[c * 4 + 1] * (double)y;
                        v = v + (double)T[c * 4 + 2] * (double)z;
                        v = v + (double)T[c * 4 + 3];
                        r.cpp[i + c * G] = (float)v;
                    }
    }
    r.maxIter = (size_t)cfg.max_iterations * (size_t)std::pow(2.0, toPerform - 1);
    r.totalIterations = 0;
    r.finalNmi = 0;
}

// reg_f3d::InitCurrentLevel (reg-lib/_reg_f3d.cpp:92-133): the level's penalty weights and grid
// (refined after the first level), and the handle parameters and largest step of the level.
void pair_level(PairRun& r, int level, nrh_f3d_params& p, float& maxStep) {
    const int idx = r.cfg.use_pyramid ? level : 0;
    const nrh_geom& lr = r.rp[idx].geom;
    maxStep = std::max(lr.dx, lr.dy);
    maxStep = std::max(maxStep, lr.dz);
    if (level == 0) {
        r.be = r.be / std::pow(16.f, (float)(r.levels - 1));
        r.le = r.le / std::pow(3.f, (float)(r.levels - 1));
    } else {
        r.be = r.be * 16;
        r.le = r.le * 3;
        nrh_geom ng;
        std::vector<float> nd;
        refine_grid(lr, r.grid, r.cpp, ng, nd);
        r.grid = ng;
        r.cpp.swap(nd);
    }
    p = nrh_f3d_params{};
    p.ref_bins = r.cfg.bins + 4;
    p.flo_bins = r.cfg.bins + 4;
    p.padding = r.cfg.padding;
    p.similarity_weight = r.simW;
    p.be_weight = r.be;
    p.le_weight = r.le;
    p.image_storage = r.cfg.image_storage;
    p.measure = r.cfg.measure;
    p.lncc_sigma = r.cfg.lncc_sigma;
    p.field_order = r.cfg.field_order;
}

// A level's handle: the level's images, the grid's points, and the measure's intensity rescaling
// (reg_nmi::InitialiseMeasure to [2, bins - 3], reg-lib/cpu/_reg_nmi.cpp:115-129; reg_lncc to [0, 1],
// reg-lib/cpu/_reg_lncc.cpp:99-104).
nrh_f3d *pair_handle(PairRun& r, int level, const nrh_f3d_params& p) {
    const int idx = r.cfg.use_pyramid ? level : 0;
    nrh_f3d *h = nullptr;
    check(nrh_f3d_create(&r.rp[idx].geom, r.rp[idx].data.data(), &r.fp[idx].geom, r.fp[idx].data.data(), nullptr,
                         &r.grid, &p, &h));
    Handle H;
    H.h = h;
    check(nrh_f3d_set_control_points(h, r.cpp.data()));
    if (p.measure == NRH_MEASURE_LNCC) check(nrh_f3d_rescale_intensities(h, 0.f, 1.f, 0.f, 1.f));
    else check(nrh_f3d_rescale_intensities(h, 2.f, p.ref_bins - 3.f, 2.f, p.flo_bins - 3.f));
    H.h = nullptr;
    return h;
}

// End of a level: the final points and the similarity of the level's last objective.
void pair_level_end(PairRun& r, double terms0, int iterations) {
    r.totalIterations += iterations;
    r.finalNmi = r.simW > 0 ? terms0 / r.simW : 0.0;
    r.maxIter /= 2;
}

void register_pair(const nrh_geom& refG, const float *refData, const nrh_geom& floG, const float *floData,
                   const nrh_f3d_config& cfg, nrh_geom& gridOut, std::vector<float>& cppOut, double& finalNmi,
                   int& totalIterations) {
    PairRun r;
    pair_begin(r, refG, refData, floG, floData, cfg);
    for (int level = 0; level < r.toPerform; ++level) {
        nrh_f3d_params p;
        float maxStep;
        pair_level(r, level, p, maxStep);
        Handle H;
        H.h = pair_handle(r, level, p);
        nrh_f3d *h = H.h;
        // Optimiser::Initialise (reg-lib/Optimiser.cpp:48-90)
        check(nrh_f3d_store_best(h));
        double best = 0, terms[3];
        check(nrh_f3d_objective(h, &best, terms));
        size_t iter = 0;
        const size_t maxIter = r.maxIter;
        const float smallest = maxStep / 100.f;
        float currentSize = maxStep;
        bool firstCall = true, justRestarted = false;
        while (currentSize) {
            if (iter >= maxIter) break;
            double maxLen = 0;
            check(nrh_f3d_gradient(h, &maxLen));  // GetObjectiveFunctionGradient + NormaliseGradient
            ++iter;
            currentSize = std::min(currentSize, maxStep);
            check(nrh_f3d_conjugate_gradient(h, firstCall ? 1 : 0));
            firstCall = false;
            // Optimiser::Optimise line search (reg-lib/Optimiser.cpp:129-173)
            size_t lineIter = 0;
            float added = 0;
            float curLen = currentSize;
            while (curLen > smallest && lineIter < 12 && iter < maxIter) {
                check(nrh_f3d_update_control_points(h, -curLen));
                double value = 0;
                check(nrh_f3d_objective(h, &value, terms));
                if (value > best) {
                    best = value;
                    added += curLen;
                    curLen *= 1.1f;
                    curLen = std::min(curLen, maxStep);
                    check(nrh_f3d_store_best(h));
                } else {
                    curLen *= 0.5;
                }
                ++iter;
                ++lineIter;
            }
            currentSize = added;
            check(nrh_f3d_restore_best(h));
            // restart to steepest descent once when the line search stalls (reg-lib/_reg_base.cpp:781-793)
            if (currentSize == 0) {
                if (!justRestarted && iter < maxIter) {
                    firstCall = true;
                    currentSize = maxStep;
                    justRestarted = true;
                }
            } else {
                justRestarted = false;
            }
        }
        check(nrh_f3d_get_control_points(h, r.cpp.data()));
        double value = 0;
        check(nrh_f3d_objective(h, &value, terms));
        pair_level_end(r, terms[0], (int)iter);
    }
    gridOut = r.grid;
    cppOut = r.cpp;
    finalNmi = r.finalNmi;
    totalIterations = r.totalIterations;
}


// Many reg_f3d registrations at once: every pair runs register_pair's levels and decisions, but
// all pairs' evaluations share the batch launches of f3d_batch.hip, and each level's iterations
// replay one captured HIP graph (gradient, conjugate gradient, twelve lockstep line-search steps,
// restore and stopping rule) until every pair has stopped.  The pyramids are built on host threads.
void register_batch(int n, const nrh_geom *refs, const float *const *refData, const nrh_geom *flos,
                    const float *const *floData, const nrh_f3d_config& cfg, std::vector<PairRun>& runs,
                    nrh_batch_stats *stats) {
    const auto t0 = std::chrono::steady_clock::now();
    runs.assign(n, PairRun{});
    {
        std::vector<std::thread> pool;
        std::vector<std::string> errs(n);
        const int T = std::max(1, std::min(n, (int)std::thread::hardware_concurrency()));
        for (int t = 0; t < T; ++t)
            pool.emplace_back([&, t] {
                for (int q = t; q < n; q += T) {
                    try {
                        pair_begin(runs[q], refs[q], refData[q], flos[q], floData[q], cfg);
                    } catch (const std::exception& e) {
                        errs[q] = e.what();
                    }
                }
            });
        for (auto& th : pool) th.join();
        for (int q = 0; q < n; ++q) NRH_REQUIRE(errs[q].empty(), NRH_ERR_ARG, "pair " + std::to_string(q) + ": " + errs[q]);
    }
    const int toPerform = runs[0].toPerform;
    double deviceSeconds = 0.0;
    int replays = 0;
    for (int level = 0; level < toPerform; ++level) {
        std::vector<nrh_f3d *> handles(n, nullptr);
        std::vector<float> maxStep(n);
        std::vector<int> maxIter(n);
        struct Owner {
            std::vector<nrh_f3d *> *hs;
            nrh_batch *b = nullptr;
            ~Owner() {
                if (b) batch_free(b);
                else
                    for (nrh_f3d *h : *hs)
                        if (h) nrh_f3d_destroy(h);
            }
        } own{ &handles };
        for (int q = 0; q < n; ++q) {
            nrh_f3d_params p;
            pair_level(runs[q], level, p, maxStep[q]);
            handles[q] = pair_handle(runs[q], level, p);
            maxIter[q] = (int)std::min<size_t>(runs[q].maxIter, 1u << 30);
        }
        nrh_batch *b = batch_new(handles.data(), n);
        b->ownsPairs = true;
        own.b = b;
        batch_autotune(b);
        if (stats && level < 8) stats->blocks_per_xcd[level] = b->tunedBlocksPerXcd;
        batch_ls_alloc(b);
        batch_ls_init(b, maxStep, maxIter);
        NRH_HIP(hipStreamSynchronize(b->stream));
        // one iteration of every pair, captured once and replayed
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        NRH_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        try {
            batch_ls_iteration(b);
        } catch (...) {
            (void)hipStreamEndCapture(b->stream, &graph);
            if (graph) (void)hipGraphDestroy(graph);
            throw;
        }
        NRH_HIP(hipStreamEndCapture(b->stream, &graph));
        const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        NRH_HIP(ie);
        NRH_HIP(hipStreamSynchronize(b->stream));
        const auto d0 = std::chrono::steady_clock::now();
        const int maxReplays = *std::max_element(maxIter.begin(), maxIter.end()) + 1;
        hipError_t ge = hipSuccess;
        for (int r = 0; r < maxReplays; ++r) {
            ge = hipGraphLaunch(exec, b->stream);
            if (ge != hipSuccess) break;
            ge = hipStreamSynchronize(b->stream);
            ++replays;
            if (ge != hipSuccess || *b->remaining == 0) break;
        }
        (void)hipGraphExecDestroy(exec);
        NRH_HIP(ge);
        deviceSeconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - d0).count();
        std::vector<LsState> st(n);
        NRH_HIP(hipMemcpy(st.data(), b->lsState.p, n * sizeof(LsState), hipMemcpyDeviceToHost));
        batch_touch(b, true);
        // the final points and their objective (reg_f3d's last objective of the level)
        batch_objective(b, nullptr);
        NRH_HIP(hipStreamSynchronize(b->stream));
        for (int q = 0; q < n; ++q) {
            check(nrh_f3d_get_control_points(handles[q], runs[q].cpp.data()));
            const double terms0 = runs[q].simW > 0 ? (double)runs[q].simW * b->published[q * kBatchResults] : 0.0;
            pair_level_end(runs[q], terms0, st[q].iter);
        }
    }
    if (stats) {
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        stats->device_seconds = deviceSeconds;
        stats->graph_replays = replays;
        stats->tile[0] = 64;
        stats->tile[1] = 8;
        stats->tile[2] = 8;
    }
}


// reg_f3d2's run (reg_base<T>::Run with the reg_f3d2 overrides, reg-lib/_reg_f3d2.cpp): symmetric
// velocity grids over both images at the coarsest level, refined by subdivision at each level,
// the conjugate-gradient / line-search loop over both grids with the velocity update and the
// symmetrisation after every step; every evaluation runs on the device through the pair.
void register_pair_sym(const nrh_geom& refG, const float *refData, const nrh_geom& floG, const float *floData,
                       const nrh_f3d_config& cfg, nrh_geom& gridOut, std::vector<float>& fwOut,
                       std::vector<float>& bwOut, double& finalNmi, int& totalIterations) {
    NRH_REQUIRE(cfg.levels >= 1 && cfg.bins >= 1 && cfg.max_iterations >= 0, NRH_ERR_ARG, "invalid configuration");
    NRH_REQUIRE(refG.nz > 1 && floG.nz > 1, NRH_ERR_UNSUPPORTED, "the symmetric registration is 3-D only");
    const int levels = cfg.levels;
    int toPerform = cfg.levels_to_perform > 0 ? std::min(cfg.levels_to_perform, levels) : levels;
    if (!cfg.use_pyramid) toPerform = levels;
    // reg_f3d2::CheckParameters (reg-lib/_reg_f3d2.cpp:118-136); no inverse-consistency term
    float be = cfg.be_weight, le = cfg.le_weight;
    const float penalty = be + le;
    float simW;
    if (penalty >= 1) {
        simW = 0;
        be /= penalty;
        le /= penalty;
    } else {
        simW = 1 - penalty;
    }
    std::vector<Level> rp, fp;
    if (cfg.use_pyramid) {
        rp = pyramid(refG, refData, levels, toPerform);
        fp = pyramid(floG, floData, levels, toPerform);
    } else {
        rp.assign(1, Level{ refG, std::vector<float>(refData, refData + voxels(refG)) });
        fp.assign(1, Level{ floG, std::vector<float>(floData, floData + voxels(floG)) });
    }
    // reg_f3d2::Initialise (reg-lib/_reg_f3d2.cpp:536-590): spacing in voxels of the mean voxel size
    const float pd[3][2] = { { refG.dx, floG.dx }, { refG.dy, floG.dy }, { refG.dz, floG.dz } };
    float gsp[3];
    for (int a = 0; a < 3; ++a) {
        gsp[a] = cfg.spacing;
        if (gsp[a] < 0) gsp[a] *= -(pd[a][0] + pd[a][1]) / 2.f;
        gsp[a] *= std::pow(2.f, (float)(levels - 1));
    }
    // reg_createSymmetricControlPointGrids with the half affines (reg-lib/_reg_f3d2.cpp:560-565);
    // the backward content's transformation is the affine's inverse (:602-603)
    float hf[16], hb[16], affInv[16];
    int ext = 0;
    const float *halfF = nullptr, *halfB = nullptr;
    if (cfg.affine) {
        ext = half_affines(cfg.affine, hf, hb);
        halfF = hf;
        halfB = hb;
        mat44_inverse(cfg.affine, affInv);
    }
    nrh_geom grid;
    std::vector<float> fw, bw;
    symmetric_grids(rp[0].geom, fp[0].geom, gsp, grid, nullptr, nullptr, halfF, halfB);
    fw.resize(3 * voxels(grid));
    bw.resize(3 * voxels(grid));
    symmetric_grids(rp[0].geom, fp[0].geom, gsp, grid, fw.data(), bw.data(), halfF, halfB);
    int steps = 6;  // intent_p2 of both grids
    size_t maxIter = (size_t)cfg.max_iterations * (size_t)std::pow(2.0, toPerform - 1);
    totalIterations = 0;
    finalNmi = 0;
    for (int level = 0; level < toPerform; ++level) {
        const int idx = cfg.use_pyramid ? level : 0;
        const nrh_geom& lr = rp[idx].geom;
        const nrh_geom& lf = fp[idx].geom;
        // reg_f3d2::InitCurrentLevel (reg-lib/_reg_f3d2.cpp:55-100): the largest voxel size of both
        float maxStep = std::max(lr.dx, lr.dy);
        maxStep = std::max(maxStep, lf.dx);
        maxStep = std::max(maxStep, lf.dy);
        maxStep = std::max(maxStep, lr.dz);
        maxStep = std::max(maxStep, lf.dz);
        if (level == 0) {
            be = be / std::pow(16.f, (float)(levels - 1));
            le = le / std::pow(3.f, (float)(levels - 1));
        } else {
            be = be * 16;
            le = le * 3;
            nrh_geom ng;
            std::vector<float> nf, nb;
            refine_grid_self(grid, fw, ng, nf);
            refine_grid_self(grid, bw, ng, nb);
            grid = ng;
            fw.swap(nf);
            bw.swap(nb);
        }
        nrh_f3d_params p{};
        p.ref_bins = cfg.bins + 4;
        p.flo_bins = cfg.bins + 4;
        p.padding = cfg.padding;
        p.similarity_weight = simW;
        p.be_weight = be;
        p.le_weight = le;
        p.image_storage = 0;
        p.measure = cfg.measure;
        p.lncc_sigma = cfg.lncc_sigma;
        p.field_order = cfg.field_order;
        Handle F, B;
        check(nrh_f3d_create(&lr, rp[idx].data.data(), &lf, fp[idx].data.data(), nullptr, &grid, &p, &F.h));
        check(nrh_f3d_create(&lf, fp[idx].data.data(), &lr, rp[idx].data.data(), nullptr, &grid, &p, &B.h));
        nrh_f3d *hf = F.h, *hb = B.h;
        check(nrh_f3d_set_control_points(hf, fw.data()));
        check(nrh_f3d_set_control_points(hb, bw.data()));
        check(nrh_f3d_set_velocity_grid(hf, steps));
        check(nrh_f3d_set_velocity_grid(hb, steps));
        if (cfg.affine) {
            check(nrh_f3d_set_velocity_affines(hf, ext ? halfF : nullptr, ext ? halfF : nullptr, cfg.affine));
            check(nrh_f3d_set_velocity_affines(hb, ext ? halfB : nullptr, ext ? halfB : nullptr, affInv));
        }
        if (p.measure == NRH_MEASURE_LNCC) {
            check(nrh_f3d_rescale_intensities(hf, 0.f, 1.f, 0.f, 1.f));
            check(nrh_f3d_rescale_intensities(hb, 0.f, 1.f, 0.f, 1.f));
        } else {
            check(nrh_f3d_rescale_intensities(hf, 2.f, p.ref_bins - 3.f, 2.f, p.flo_bins - 3.f));
            check(nrh_f3d_rescale_intensities(hb, 2.f, p.flo_bins - 3.f, 2.f, p.ref_bins - 3.f));
        }
        check(nrh_f3d_store_best(hf));
        check(nrh_f3d_store_best(hb));
        double best = 0, terms[3];
        check(nrh_f3d2_objective(hf, hb, &best, terms));
        size_t iter = 0;
        const float smallest = maxStep / 100.f;
        float currentSize = maxStep;
        bool firstCall = true, justRestarted = false;
        // reg_f3d2::UpdateParameters (reg-lib/_reg_f3d2.cpp:612-648): restore the best grids, add
        // the scaled gradients, symmetrise
        auto update = [&](float scale) {
            check(nrh_f3d_restore_best(hf));
            check(nrh_f3d_restore_best(hb));
            check(nrh_f3d_update_velocity_field(hf, scale, 1, 1, 1));
            check(nrh_f3d_update_velocity_field(hb, scale, 1, 1, 1));
            check(nrh_f3d_symmetrise_velocity_fields(hf, hb));
        };
        while (currentSize) {
            if (iter >= maxIter) break;
            double maxLen = 0;
            check(nrh_f3d2_gradient(hf, hb, &maxLen));
            ++iter;
            currentSize = std::min(currentSize, maxStep);
            check(nrh_f3d2_conjugate_gradient(hf, hb, firstCall ? 1 : 0));
            firstCall = false;
            size_t lineIter = 0;
            float added = 0;
            float curLen = currentSize;
            while (curLen > smallest && lineIter < 12 && iter < maxIter) {
                update(-curLen);
                double value = 0;
                check(nrh_f3d2_objective(hf, hb, &value, terms));
                if (value > best) {
                    best = value;
                    added += curLen;
                    curLen *= 1.1f;
                    curLen = std::min(curLen, maxStep);
                    check(nrh_f3d_store_best(hf));
                    check(nrh_f3d_store_best(hb));
                } else {
                    curLen *= 0.5;
                }
                ++iter;
                ++lineIter;
            }
            currentSize = added;
            check(nrh_f3d_restore_best(hf));
            check(nrh_f3d_restore_best(hb));
            if (currentSize == 0) {
                if (!justRestarted && iter < maxIter) {
                    firstCall = true;
                    currentSize = maxStep;
                    justRestarted = true;
                }
            } else {
                justRestarted = false;
            }
        }
        totalIterations += (int)iter;
        check(nrh_f3d_get_control_points(hf, fw.data()));
        check(nrh_f3d_get_control_points(hb, bw.data()));
        double value = 0;
        check(nrh_f3d2_objective(hf, hb, &value, terms));
        check(nrh_f3d_get_velocity_steps(hf, &steps));
        finalNmi = simW > 0 ? terms[0] / simW : 0.0;
        maxIter /= 2;
    }
    gridOut = grid;
    fwOut = fw;
    bwOut = bw;
}
}  // namespace nrh

using namespace nrh;

extern "C" {

int nrh_f3d_register(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                     const nrh_f3d_config *cfg, nrh_geom *grid_out, float *cpp_out, size_t cpp_capacity,
                     double *final_nmi, int *iterations) {
    return guard([&] {
        check_geom(ref, "reference");
        check_geom(flo, "floating");
        NRH_REQUIRE(ref_data && flo_data && cfg && grid_out, NRH_ERR_ARG, "null argument");
        nrh_geom g;
        std::vector<float> cpp;
        double nmi = 0;
        int it = 0;
        register_pair(*ref, ref_data, *flo, flo_data, *cfg, g, cpp, nmi, it);
        *grid_out = g;
        if (cpp_out) {
            NRH_REQUIRE(cpp.size() <= cpp_capacity, NRH_ERR_ARG, "cpp_out too small for the final grid");
            std::memcpy(cpp_out, cpp.data(), cpp.size() * sizeof(float));
        }
        if (final_nmi) *final_nmi = nmi;
        if (iterations) *iterations = it;
    });
}

int nrh_f3d_register_batch(int n, const nrh_geom *refs, const float *const *ref_data, const nrh_geom *flos,
                           const float *const *flo_data, const nrh_f3d_config *cfg, nrh_geom *grids_out,
                           float *const *cpp_out, const size_t *capacity, double *final_nmi, int *iterations,
                           nrh_batch_stats *stats) {
    return guard([&] {
        NRH_REQUIRE(n >= 1 && refs && ref_data && flos && flo_data && cfg && grids_out, NRH_ERR_ARG, "null argument");
        NRH_REQUIRE(cfg->measure == NRH_MEASURE_NMI && cfg->le_weight == 0.f && cfg->be_weight > 0.f, NRH_ERR_UNSUPPORTED,
                    "batched registration: NMI with a bending-energy term and no linear energy");
        for (int q = 0; q < n; ++q) {
            check_geom(&refs[q], "reference");
            check_geom(&flos[q], "floating");
            NRH_REQUIRE(ref_data[q] && flo_data[q], NRH_ERR_ARG, "null image");
        }
        if (stats) *stats = nrh_batch_stats{};
        std::vector<PairRun> runs;
        register_batch(n, refs, ref_data, flos, flo_data, *cfg, runs, stats);
        for (int q = 0; q < n; ++q) {
            grids_out[q] = runs[q].grid;
            if (cpp_out && cpp_out[q]) {
                NRH_REQUIRE(capacity && runs[q].cpp.size() <= capacity[q], NRH_ERR_ARG,
                            "cpp_out too small for the final grid of pair " + std::to_string(q));
                std::memcpy(cpp_out[q], runs[q].cpp.data(), runs[q].cpp.size() * sizeof(float));
            }
            if (final_nmi) final_nmi[q] = runs[q].finalNmi;
            if (iterations) iterations[q] = runs[q].totalIterations;
        }
    });
}

int nrh_f3d2_register(const nrh_geom *ref, const float *ref_data, const nrh_geom *flo, const float *flo_data,
                      const nrh_f3d_config *cfg, nrh_geom *grid_out, float *fw_out, float *bw_out, size_t capacity,
                      double *final_nmi, int *iterations) {
    return guard([&] {
        check_geom(ref, "reference");
        check_geom(flo, "floating");
        NRH_REQUIRE(ref_data && flo_data && cfg && grid_out, NRH_ERR_ARG, "null argument");
        nrh_geom g;
        std::vector<float> fw, bw;
        double nmi = 0;
        int it = 0;
        register_pair_sym(*ref, ref_data, *flo, flo_data, *cfg, g, fw, bw, nmi, it);
        *grid_out = g;
        if (fw_out || bw_out) NRH_REQUIRE(fw.size() <= capacity, NRH_ERR_ARG, "output arrays too small for the final grid");
        if (fw_out) std::memcpy(fw_out, fw.data(), fw.size() * sizeof(float));
        if (bw_out) std::memcpy(bw_out, bw.data(), bw.size() * sizeof(float));
        if (final_nmi) *final_nmi = nmi;
        if (iterations) *iterations = it;
    });
}

int nrh_downsample_image(const nrh_geom *in, const float *data, const int axes[3], nrh_geom *out, float *out_data) {
    return guard([&] {
        check_geom(in, "input");
        NRH_REQUIRE(data && axes && out, NRH_ERR_ARG, "null argument");
        const bool ax[3] = { axes[0] != 0, axes[1] != 0, axes[2] != 0 };
        std::vector<float> src(data, data + voxels(*in)), res;
        nrh_geom g;
        downsample(*in, src, ax, g, res);
        *out = g;
        if (out_data) std::memcpy(out_data, res.data(), res.size() * sizeof(float));
    });
}

int nrh_refine_control_point_grid(const nrh_geom *ref, const nrh_geom *grid, const float *data, nrh_geom *out,
                                  float *out_data) {
    return guard([&] {
        check_geom(ref, "reference");
        check_geom(grid, "grid");
        NRH_REQUIRE(data && out, NRH_ERR_ARG, "null argument");
        std::vector<float> src(data, data + 3 * voxels(*grid)), res;
        nrh_geom g;
        refine_grid(*ref, *grid, src, g, res);
        *out = g;
        if (out_data) std::memcpy(out_data, res.data(), res.size() * sizeof(float));
    });
}

}  // extern "C"
