/*
 * nr_oracle.c -- CPU restatement of the NiftyReg F3D hot path (test infrastructure only).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this library, and
 * only as the checker.  Each function cites the reference function it restates; the
 * arithmetic follows the reference's scalar (non-SSE) path: float products and sums where the
 * reference uses float, double where it promotes to double, no fused multiply-add (this file
 * is built with -ffp-contract=off, as the reference is on x86-64 without FMA).
 */
#include "nr_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#define M(m, i, j) ((m)[(i) * 4 + (j)])

/* Floor<int>(float) (reg-lib/cpu/Maths.hpp:63-66): truncate, then step down for negatives. */
static int floor_int(float x) {
    int i = (int)x;
    return i - (x < (float)i);
}

/* Mat44Mul<float> (reg-lib/cpu/Maths.hpp:272-285): double accumulation, float result. */
/* 3x3 float matrix (niftilib mat33) */
typedef struct { float m[3][3]; } orc_m33;

static void mat44_mul_f(const float *m, const float in[3], float out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = (float)((double)M(m, r, 0) * (double)in[0] + (double)M(m, r, 1) * (double)in[1] +
                         (double)M(m, r, 2) * (double)in[2] + (double)M(m, r, 3));
}

/* Mat44Mul<double> */
static void mat44_mul_d(const float *m, const double in[3], double out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = (double)M(m, r, 0) * in[0] + (double)M(m, r, 1) * in[1] + (double)M(m, r, 2) * in[2] +
                 (double)M(m, r, 3);
}

/* mat44 product A*B in double, stored in float (reg-lib/cpu/Maths.hpp:235-245) */
static void mat44_mm(const float *a, const float *b, float *out) {
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
            out[i * 4 + j] = (float)((double)M(a, i, 0) * (double)M(b, 0, j) + (double)M(a, i, 1) * (double)M(b, 1, j) +
                                     (double)M(a, i, 2) * (double)M(b, 2, j) + (double)M(a, i, 3) * (double)M(b, 3, j));
}

void orc_mat44_inverse(const float in[16], float out[16]) {
    /* Inverse of an affine 4x4 by cofactors in double (the textbook adjugate formula). */
    double a[3][3], t[3];
    for (int i = 0; i < 3; ++i) {
        for (int j = 0; j < 3; ++j) a[i][j] = M(in, i, j);
        t[i] = M(in, i, 3);
    }
    double c[3][3];
    c[0][0] = a[1][1] * a[2][2] - a[2][1] * a[1][2];
    c[0][1] = -(a[0][1] * a[2][2] - a[2][1] * a[0][2]);
    c[0][2] = a[0][1] * a[1][2] - a[1][1] * a[0][2];
    c[1][0] = -(a[1][0] * a[2][2] - a[2][0] * a[1][2]);
    c[1][1] = a[0][0] * a[2][2] - a[2][0] * a[0][2];
    c[1][2] = -(a[0][0] * a[1][2] - a[1][0] * a[0][2]);
    c[2][0] = a[1][0] * a[2][1] - a[2][0] * a[1][1];
    c[2][1] = -(a[0][0] * a[2][1] - a[2][0] * a[0][1]);
    c[2][2] = a[0][0] * a[1][1] - a[1][0] * a[0][1];
    double det = a[0][0] * c[0][0] + a[0][1] * c[1][0] + a[0][2] * c[2][0];
    double inv = det != 0.0 ? 1.0 / det : 0.0;
    for (int i = 0; i < 3; ++i) {
        double s = 0;
        for (int j = 0; j < 3; ++j) {
            out[i * 4 + j] = (float)(c[i][j] * inv);
            s += c[i][j] * t[j];
        }
        out[i * 4 + 3] = (float)(-s * inv);
    }
    out[12] = out[13] = out[14] = 0.f;
    out[15] = det != 0.0 ? 1.f : 0.f;
}

/* ------------------------------------------------------------------ spline basis */
void orc_bspline_basis(float t, float v[4]) {
    float ff = t * t;
    float fff = ff * t;
    float mf = (float)(1.0 - t);
    v[0] = (float)((double)(mf * mf * mf) / 6.0);
    v[1] = (float)((3.0 * fff - 6.0 * ff + 4.0) / 6.0);
    v[2] = (float)((-3.0 * fff + 3.0 * ff + 3.0 * t + 1.0) / 6.0);
    v[3] = (float)(fff / 6.0);
}

void orc_bspline_basis_d1(float t, float v[4], float d[4]) {
    orc_bspline_basis(t, v);
    d[3] = (float)(t * t / 2.0);
    d[0] = (float)(t - 1.0 / 2.0 - d[3]);
    d[2] = (float)(1.0 + d[0] - 2.0 * d[3]);
    d[1] = -d[0] - d[2] - d[3];
}

/* ------------------------------------------------------------------ deformation field */
static void grid_values(const orc_geom *cpp, const float *cppData, int x0, int y0, int z0,
                        float *gx, float *gy, float *gz) {
    /* get_GridValues (reg-lib/cpu/_reg_splineBasis.cpp:603-665) for an in-range window. */
    const size_t g = (size_t)cpp->nx * cpp->ny * cpp->nz;
    int coord = 0;
    for (int Z = z0; Z < z0 + 4; ++Z)
        for (int Y = y0; Y < y0 + 4; ++Y)
            for (int X = x0; X < x0 + 4; ++X, ++coord) {
                size_t idx = ((size_t)Z * cpp->ny + Y) * cpp->nx + X;
                gx[coord] = cppData[idx];
                gy[coord] = cppData[idx + g];
                gz[coord] = cppData[idx + 2 * g];
            }
}

void orc_spline_field(const orc_geom *ref, const orc_geom *cpp, const float *cppData,
                      const int *mask, float *field) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    float *fx = field, *fy = field + n, *fz = field + 2 * n;
    const float gvs[3] = { cpp->dx / ref->dx, cpp->dy / ref->dy, cpp->dz / ref->dz };

    if (gvs[0] == 5.f && gvs[1] == 5.f && gvs[2] == 5.f) {
        /* Look-up table of the 125 x 64 products x*(y*z) (reg-lib/cpu/_reg_localTrans.cpp:1112-1175) */
        float *coef = (float *)malloc(125 * 64 * sizeof(float));
        for (int z = 0; z < 5; ++z) {
            float zb[4], yb[4], xb[4], yz[16];
            orc_bspline_basis((float)z / 5.f, zb);
            for (int y = 0; y < 5; ++y) {
                orc_bspline_basis((float)y / 5.f, yb);
                int c = 0;
                for (int a = 0; a < 4; ++a)
                    for (int j = 0; j < 4; ++j) yz[c++] = yb[j] * zb[a];
                for (int x = 0; x < 5; ++x) {
                    orc_bspline_basis((float)x / 5.f, xb);
                    float *cf = coef + ((z * 5 + y) * 5 + x) * 64;
                    int k = 0;
                    for (int a = 0; a < 16; ++a)
                        for (int i = 0; i < 4; ++i) cf[k++] = xb[i] * yz[a];
                }
            }
        }
        /* Loop over 5x5x5 voxel blocks (1178-1300) */
#pragma omp parallel for collapse(2) schedule(static)
        for (int zp = 0; zp < cpp->nz - 3; ++zp)
            for (int yp = 0; yp < cpp->ny - 3; ++yp) {
                float gx[64], gy[64], gz[64];
                for (int xp = 0; xp < cpp->nx - 3; ++xp) {
                    grid_values(cpp, cppData, xp, yp, zp, gx, gy, gz);
                    for (int c = 0; c < 5; ++c) {
                        int z = zp * 5 + c;
                        if (z >= ref->nz) break;
                        for (int b = 0; b < 5; ++b) {
                            int y = yp * 5 + b;
                            if (y >= ref->ny) break;
                            for (int a = 0; a < 5; ++a) {
                                int x = xp * 5 + a;
                                size_t idx = ((size_t)z * ref->ny + y) * ref->nx + x;
                                if (x >= ref->nx || (mask && mask[idx] < 0)) continue;
                                const float *cf = coef + ((c * 5 + b) * 5 + a) * 64;
                                float r0 = 0, r1 = 0, r2 = 0;
                                for (int k = 0; k < 64; ++k) {
                                    r0 += gx[k] * cf[k];
                                    r1 += gy[k] * cf[k];
                                    r2 += gz[k] * cf[k];
                                }
                                fx[idx] = r0; fy[idx] = r1; fz[idx] = r2;
                            }
                        }
                    }
                }
            }
        free(coef);
        return;
    }

    /* General spacing (reg-lib/cpu/_reg_localTrans.cpp:1305-1450) */
#pragma omp parallel for schedule(static)
    for (int z = 0; z < ref->nz; ++z) {
        float zb[4], yb[4], xb[4], yz[16], xyz[64], gx[64], gy[64], gz[64];
        int zPre = (int)((float)z / gvs[2]);
        float basis = (float)z / gvs[2] - (float)zPre;
        if (basis < 0) basis = 0;
        orc_bspline_basis(basis, zb);
        for (int y = 0; y < ref->ny; ++y) {
            int yPre = (int)((float)y / gvs[1]);
            basis = (float)y / gvs[1] - (float)yPre;
            if (basis < 0) basis = 0;
            orc_bspline_basis(basis, yb);
            int c = 0;
            for (int a = 0; a < 4; ++a)
                for (int j = 0; j < 4; ++j) yz[c++] = yb[j] * zb[a];
            for (int x = 0; x < ref->nx; ++x) {
                int xPre = (int)((float)x / gvs[0]);
                basis = (float)x / gvs[0] - (float)xPre;
                if (basis < 0) basis = 0;
                orc_bspline_basis(basis, xb);
                c = 0;
                for (int a = 0; a < 16; ++a)
                    for (int i = 0; i < 4; ++i) xyz[c++] = xb[i] * yz[a];
                grid_values(cpp, cppData, xPre, yPre, zPre, gx, gy, gz);
                size_t idx = ((size_t)z * ref->ny + y) * ref->nx + x;
                float r0 = 0, r1 = 0, r2 = 0;
                if (!mask || mask[idx] > -1) {
                    for (int a = 0; a < 64; ++a) {
                        r0 += gx[a] * xyz[a];
                        r1 += gy[a] * xyz[a];
                        r2 += gz[a] * xyz[a];
                    }
                }
                fx[idx] = r0; fy[idx] = r1; fz[idx] = r2;
            }
        }
    }
}

/* ------------------------------------------------------------------ resampling */
void orc_resample_linear(const orc_geom *flo, const float *floData, const orc_geom *ref,
                         const float *field, const int *mask, float pad, float *warped) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < n; ++idx) {
        if (mask && mask[idx] < 0) continue;
        float world[3] = { field[idx], field[idx + n], field[idx + 2 * n] }, pos[3];
        mat44_mul_f(flo->ijk, world, pos);
        int prev[3];
        float xb[2], yb[2], zb[2];
        float *bases[3] = { xb, yb, zb };
        for (int i = 0; i < 3; ++i) {
            prev[i] = floor_int(pos[i]);
            float rel = pos[i] - (float)prev[i];
            if (rel < 0) rel = 0;
            bases[i][1] = rel;
            bases[i][0] = 1.f - rel;
        }
        float intensity = 0;
        int inside = prev[0] > -1 && prev[0] + 1 < flo->nx && prev[1] > -1 && prev[1] + 1 < flo->ny &&
                     prev[2] > -1 && prev[2] + 1 < flo->nz;
        for (int c = 0; c < 2; ++c) {
            int Z = prev[2] + c;
            float yTemp = 0;
            for (int b = 0; b < 2; ++b) {
                int Y = prev[1] + b;
                float xTemp = 0;
                for (int a = 0; a < 2; ++a) {
                    int X = prev[0] + a;
                    float v;
                    if (inside || (X > -1 && X < flo->nx && Y > -1 && Y < flo->ny && Z > -1 && Z < flo->nz))
                        v = floData[((size_t)Z * flo->ny + Y) * flo->nx + X];
                    else v = pad;
                    xTemp += v * xb[a];
                }
                yTemp += xTemp * yb[b];
            }
            intensity += yTemp * zb[c];
        }
        warped[idx] = intensity;
    }
}

void orc_image_gradient_linear(const orc_geom *flo, const float *floData, const orc_geom *ref,
                               const float *field, const int *mask, float pad, float *grad) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    const float deriv[2] = { -1.f, 1.f };
#pragma omp parallel for schedule(static)
    for (size_t idx = 0; idx < n; ++idx) {
        float g[3] = { 0, 0, 0 };
        if (!mask || mask[idx] > -1) {
            float world[3] = { field[idx], field[idx + n], field[idx + 2 * n] }, pos[3];
            mat44_mul_f(flo->ijk, world, pos);
            int prev[3] = { floor_int(pos[0]), floor_int(pos[1]), floor_int(pos[2]) };
            float rel, xb[2], yb[2], zb[2];
            rel = pos[0] - (float)prev[0]; xb[0] = (float)(1.0 - rel); xb[1] = rel;
            rel = pos[1] - (float)prev[1]; yb[0] = (float)(1.0 - rel); yb[1] = rel;
            rel = pos[2] - (float)prev[2]; zb[0] = (float)(1.0 - rel); zb[1] = rel;
            for (int c = 0; c < 2; ++c) {
                int Z = prev[2] + c;
                int zIn = Z > -1 && Z < flo->nz;
                float xx = 0, yy = 0, zz = 0;
                for (int b = 0; b < 2; ++b) {
                    int Y = prev[1] + b;
                    int yzIn = zIn && Y > -1 && Y < flo->ny;
                    float xT = 0, yT = 0;
                    for (int a = 0; a < 2; ++a) {
                        int X = prev[0] + a;
                        float coeff = (yzIn && X > -1 && X < flo->nx) ?
                            floData[((size_t)Z * flo->ny + Y) * flo->nx + X] : pad;
                        xT += coeff * deriv[a];
                        yT += coeff * xb[a];
                    }
                    xx += xT * yb[b];
                    yy += yT * deriv[b];
                    zz += yT * yb[b];
                }
                g[0] += xx * zb[c];
                g[1] += yy * zb[c];
                g[2] += zz * deriv[c];
            }
            for (int i = 0; i < 3; ++i)
                if (g[i] != g[i]) g[i] = 0;
        }
        grad[idx] = g[0];
        grad[idx + n] = g[1];
        grad[idx + 2 * n] = g[2];
    }
}

/* ------------------------------------------------------------------ NMI */
void orc_intensity_rescale(float *data, size_t n, float newMin, float newMax) {
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (size_t i = 0; i < n; ++i) {
        float v = data[i];
        if (v == v) {
            mn = v < mn ? v : mn;
            mx = v > mx ? v : mx;
        }
    }
    double diff = (double)(mx - mn);
    double newDiff = (double)(newMax - newMin);
    for (size_t i = 0; i < n; ++i) {
        double v = (double)data[i];
        if (v == v) {
            v = (v - (double)mn) / diff;
            v = v * newDiff + newMin;
        }
        data[i] = (float)v;
    }
}

/* GetBasisSplineValue<double> (reg-lib/cpu/_reg_nmi.h:116-129); note the float constant 2.f/3.f. */
static double bspline_d(double x) {
    x = x < 0 ? -x : x;
    double v = 0;
    if (x < 2.f) {
        if (x < 1.f) v = (double)(2.f / 3.f) + (0.5f * x - 1.f) * x * x;
        else {
            x -= 2.f;
            v = -x * x * x / 6.f;
        }
    }
    return v;
}

/* GetBasisSplineDerivativeValue<double> (reg-lib/cpu/_reg_nmi.h:131-145) */
static double bspline_deriv_d(double ox) {
    double x = ox < 0 ? -ox : ox;
    double v = 0;
    if (x < 2.f) {
        if (x < 1.f) v = (1.5f * x - 2.f) * ox;
        else {
            x -= 2.f;
            v = -0.5f * x * x;
            if (ox < 0) v = -v;
        }
    }
    return v;
}

double orc_nmi(const float *ref, const float *war, const int *mask, size_t n, int rb, int fb,
               double *pro, double *logh, double *ent) {
    const int total = rb * fb + rb + fb;
    memset(pro, 0, total * sizeof(double));
    /* integer counts (the reference counts in per-thread double arrays; counts are exact) */
    unsigned long long *counts = (unsigned long long *)calloc((size_t)rb * fb, sizeof(unsigned long long));
    for (size_t i = 0; i < n; ++i) {
        if (mask && mask[i] < 0) continue;
        float r = ref[i], w = war[i];
        if (r == r && w == w && 0 <= r && r < rb && 0 <= w && w < fb)
            counts[(int)r + (int)w * rb]++;
    }
    for (int i = 0; i < rb * fb; ++i) pro[i] = (double)counts[i];
    free(counts);
    const double kernel[3] = { bspline_d(-1.0), bspline_d(0.0), bspline_d(-1.0) };
    memset(logh, 0, total * sizeof(double));
    for (int f = 0; f < fb; ++f)
        for (int r = 0; r < rb; ++r) {
            double v = 0;
            for (int it = 0; it < 3; ++it) {
                int idx = r - 1 + it;
                if (-1 < idx && idx < rb) v += pro[idx + rb * f] * kernel[it];
            }
            logh[r + rb * f] = v;
        }
    for (int r = 0; r < rb; ++r)
        for (int f = 0; f < fb; ++f) {
            double v = 0;
            for (int it = 0; it < 3; ++it) {
                int idx = f - 1 + it;
                if (-1 < idx && idx < fb) v += logh[r + rb * idx] * kernel[it];
            }
            pro[r + rb * f] = v;
        }
    double active = 0;
    for (int i = 0; i < total; ++i) active += pro[i];
    ent[3] = active;
    for (int i = 0; i < total; ++i) pro[i] /= active;
    for (int r = 0; r < rb; ++r) {
        double s = 0;
        for (int f = 0; f < fb; ++f) s += pro[r + rb * f];
        pro[rb * fb + r] = s;
    }
    for (int f = 0; f < fb; ++f) {
        double s = 0;
        for (int r = 0; r < rb; ++r) s += pro[rb * f + r];
        pro[rb * fb + rb + f] = s;
    }
    memset(logh, 0, total * sizeof(double));
    double hr = 0, hf = 0, hj = 0;
    for (int r = 0; r < rb; ++r) {
        double p = pro[rb * fb + r];
        if (p > 0) { double l = log(p); hr -= p * l; logh[rb * fb + r] = l; }
    }
    for (int f = 0; f < fb; ++f) {
        double p = pro[rb * fb + rb + f];
        if (p > 0) { double l = log(p); hf -= p * l; logh[rb * fb + rb + f] = l; }
    }
    for (int i = 0; i < rb * fb; ++i) {
        double p = pro[i];
        if (p > 0) { double l = log(p); hj -= p * l; logh[i] = l; }
    }
    ent[0] = hr; ent[1] = hf; ent[2] = hj;
    return (hr + hf) / hj;
}

void orc_nmi_voxel_gradient(const float *ref, const float *war, const float *warGrad,
                            const int *mask, size_t n, int rb, int fb, const double *logh,
                            const double *ent, double weight, float *grad) {
    const double nmi = (ent[0] + ent[1]) / ent[2];
    const size_t refOff = (size_t)rb * fb, floOff = refOff + rb;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        if (mask && mask[i] < 0) continue;
        float rv = ref[i], wv = war[i];
        if (!(rv == rv && wv == wv)) continue;
        float g[3] = { warGrad[i], warGrad[i + n], warGrad[i + 2 * n] };
        double jd[3] = { 0, 0, 0 }, rd[3] = { 0, 0, 0 }, wd[3] = { 0, 0, 0 };
        for (int r = (int)(rv - 1.f); r < (int)(rv + 3.f); ++r) {
            if (!(-1 < r && r < rb)) continue;
            for (int w = (int)(wv - 1.f); w < (int)(wv + 3.f); ++w) {
                if (!(-1 < w && w < fb)) continue;
                double common = bspline_d((double)(rv - (float)r)) * bspline_deriv_d((double)(wv - (float)w));
                double jl = logh[r + w * rb], rl = logh[r + refOff], wl = logh[w + floOff];
                for (int d = 0; d < 3; ++d) {
                    if (g[d] == g[d]) {
                        rd[d] += common * g[d] * rl;
                        wd[d] += common * g[d] * wl;
                        jd[d] += common * g[d] * jl;
                    }
                }
            }
        }
        for (int d = 0; d < 3; ++d)
            grad[i + d * n] += (float)(weight * (rd[d] + wd[d] - nmi * jd[d]) / (ent[2] * ent[3]));
    }
}

/* ------------------------------------------------------------------ convolution */
void orc_kernel_convolution_axis(const orc_geom *g, float *data, float sigma, int kernelType,
                                 const int *mask, int axis) {
    const int dims[3] = { g->nx, g->ny, g->nz };
    const float pix[3] = { g->dx, g->dy, g->dz };
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    unsigned char *isnan_ = (unsigned char *)malloc(n);
    float *dens = (float *)malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        dens[i] = ((!mask || mask[i] >= 0) && data[i] == data[i]) ? 1.f : 0.f;
        isnan_[i] = dens[i] == 0.f;
        if (isnan_[i]) data[i] = 0;
    }
    if (dims[axis] > 1) {
        double temp;
        if (sigma > 0) temp = (double)(sigma / pix[axis]);
        else temp = fabs((double)sigma);
        int radius = 0;
        if (kernelType == 2) radius = (int)(temp * 3.0f);
        else if (kernelType == 3) radius = (int)(temp * 2.0f);
        if (radius > 0) {
            float *kernel = (float *)malloc((2 * radius + 1) * sizeof(float));
            double ksum = 0;
            for (int i = -radius; i <= radius; ++i) {
                if (kernelType == 3) {
                    double rel = fabs(i / temp);
                    if (rel < 1.0) kernel[i + radius] = (float)(2.0 / 3.0 - rel * rel + 0.5 * rel * rel * rel);
                    else if (rel < 2.0) kernel[i + radius] = (float)(-((rel - 2.0) * (rel - 2.0) * (rel - 2.0)) / 6.0);
                    else kernel[i + radius] = 0;
                } else {
                    kernel[i + radius] = (float)(exp(-(double)(i * i) / (2.0 * temp * temp)) / (temp * 2.506628274631));
                }
                ksum += kernel[i + radius];
            }
            (void)ksum;
            int planes, lineOff;
            if (axis == 0) { planes = dims[1] * dims[2]; lineOff = 1; }
            else if (axis == 1) { planes = dims[0] * dims[2]; lineOff = dims[0]; }
            else { planes = dims[0] * dims[1]; lineOff = planes; }
#pragma omp parallel for schedule(static)
            for (int p = 0; p < planes; ++p) {
                size_t start;
                if (axis == 0) start = (size_t)p * dims[0];
                else if (axis == 1) start = (size_t)(p / dims[0]) * dims[0] * dims[1] + p % dims[0];
                else start = p;
                int L = dims[axis];
                float *bi = (float *)malloc(L * sizeof(float));
                float *bd = (float *)malloc(L * sizeof(float));
                for (int l = 0; l < L; ++l) {
                    bi[l] = data[start + (size_t)l * lineOff];
                    bd[l] = dens[start + (size_t)l * lineOff];
                }
                for (int l = 0; l < L; ++l) {
                    int pre = l - radius, pst = l + radius + 1;
                    const float *kp = kernel;
                    if (pre < 0) { kp = kernel - pre; pre = 0; }
                    if (pst > L) pst = L;
                    double is = 0, ds = 0;
                    for (int k = pre; k < pst; ++k) {
                        float kv = *kp++;
                        is += kv * bi[k];
                        ds += kv * bd[k];
                    }
                    data[start + (size_t)l * lineOff] = (float)is;
                    dens[start + (size_t)l * lineOff] = (float)ds;
                }
                free(bi);
                free(bd);
            }
            free(kernel);
        }
    }
    for (size_t i = 0; i < n; ++i) {
        if (isnan_[i]) data[i] = NAN;
        else data[i] = data[i] / dens[i];
    }
    free(isnan_);
    free(dens);
}

void orc_convolve_voxel_gradient(const orc_geom *ref, const orc_geom *cpp, float *voxelGrad) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    const float sp[3] = { cpp->dx, cpp->dy, cpp->dz };
    for (int axis = 0; axis < 3; ++axis)
        for (int comp = 0; comp < 3; ++comp)
            orc_kernel_convolution_axis(ref, voxelGrad + comp * n, sp[axis], 3, NULL, axis);
}

static void voxel_to_node(const orc_geom *cpp, const orc_geom *ref, const float *voxelGrad,
                          const float reorient_in[16], const float *ext0, float weight, float *nodeGrad);
void orc_voxel_to_node(const orc_geom *cpp, const orc_geom *ref, const float *voxelGrad,
                       const float reorient_ijk[16], float weight, float *nodeGrad) {
    voxel_to_node(cpp, ref, voxelGrad, reorient_ijk, NULL, weight, nodeGrad);
}
void orc_voxel_to_node_ext(const orc_geom *cpp, const orc_geom *ref, const float *voxelGrad,
                           const float reorient_ijk[16], const float ext0[16], float weight, float *nodeGrad) {
    voxel_to_node(cpp, ref, voxelGrad, reorient_ijk, ext0, weight, nodeGrad);
}
/* reg-lib/cpu/_reg_localTrans.cpp:1568-1592: a grid extension 0 (A) puts inverse(A) between the
 * grid and the image, and inverse(A)'s 3x3 (nifti_mat33_inverse) before the reorientation
 * (nifti_mat33_mul). */
static orc_m33 m33_inverse(orc_m33 R);
static orc_m33 m33_mul(orc_m33 A, orc_m33 B);
static void voxel_to_node(const orc_geom *cpp, const orc_geom *ref, const float *voxelGrad,
                          const float reorient_in[16], const float *ext0, float weight, float *nodeGrad) {
    const size_t nv = (size_t)ref->nx * ref->ny * ref->nz;
    const size_t ng = (size_t)cpp->nx * cpp->ny * cpp->nz;
    float T[16], reorient_ijk[16];
    memcpy(reorient_ijk, reorient_in, sizeof reorient_ijk);
    if (ext0) {
        float inv[16], g2r[16];
        orc_nifti_mat44_inverse(ext0, inv);
        mat44_mm(inv, cpp->xyz, g2r);
        mat44_mm(ref->ijk, g2r, T);
        orc_m33 a, o;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                a.m[i][j] = M(ext0, i, j);
                o.m[i][j] = M(reorient_in, i, j);
            }
        const orc_m33 r = m33_mul(m33_inverse(a), o);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) M(reorient_ijk, i, j) = r.m[i][j];
    } else {
        mat44_mm(ref->ijk, cpp->xyz, T);
    }
    /* ratio = node spacing (sform column norm) / voxel pixdim, folded into the weight */
    for (int i = 0; i < 3; ++i) {
        float r = sqrtf(M(cpp->xyz, i, 0) * M(cpp->xyz, i, 0) + M(cpp->xyz, i, 1) * M(cpp->xyz, i, 1) +
                        M(cpp->xyz, i, 2) * M(cpp->xyz, i, 2));
        const float pd[3] = { ref->dx, ref->dy, ref->dz };
        r /= pd[i];
        weight *= r;
    }
    const float *vx = voxelGrad, *vy = voxelGrad + nv, *vz = voxelGrad + 2 * nv;
    for (int z = 0; z < cpp->nz; ++z)
        for (int y = 0; y < cpp->ny; ++y)
            for (int x = 0; x < cpp->nx; ++x) {
                float nc[3] = { (float)x, (float)y, (float)z }, vc[3];
                mat44_mul_f(T, nc, vc);
                int pre[3] = { floor_int(vc[0]), floor_int(vc[1]), floor_int(vc[2]) };
                float bx[2], by[2], bz[2];
                bx[1] = vc[0] - (float)pre[0]; bx[0] = 1.f - bx[1];
                by[1] = vc[1] - (float)pre[1]; by[0] = 1.f - by[1];
                bz[1] = vc[2] - (float)pre[2]; bz[0] = 1.f - bz[1];
                float iv[3] = { 0, 0, 0 };
                for (int c = 0; c < 2; ++c) {
                    int Z = pre[2] + c;
                    if (Z < 0 || Z >= ref->nz) continue;
                    for (int b = 0; b < 2; ++b) {
                        int Y = pre[1] + b;
                        if (Y < 0 || Y >= ref->ny) continue;
                        for (int a = 0; a < 2; ++a) {
                            int X = pre[0] + a;
                            if (X < 0 || X >= ref->nx) continue;
                            size_t idx = ((size_t)Z * ref->ny + Y) * ref->nx + X;
                            float w = bx[a] * by[b];
                            w *= bz[c];
                            iv[0] += w * vx[idx];
                            iv[1] += w * vy[idx];
                            iv[2] += w * vz[idx];
                        }
                    }
                }
                /* reorientation applied transposed (1654-1666) */
                float ro[3];
                for (int j = 0; j < 3; ++j)
                    ro[j] = M(reorient_ijk, 0, j) * iv[0] + M(reorient_ijk, 1, j) * iv[1] + M(reorient_ijk, 2, j) * iv[2];
                size_t nidx = ((size_t)z * cpp->ny + y) * cpp->nx + x;
                nodeGrad[nidx] = ro[0] * weight;
                nodeGrad[nidx + ng] = ro[1] * weight;
                nodeGrad[nidx + 2 * ng] = ro[2] * weight;
            }
}

/* ------------------------------------------------------------------ bending energy */
static float round6(double v) {
    /* The reference tabulates these stencils as 6-decimal literals; reproduce the literal. */
    double r = floor(fabs(v) * 1e6 + 0.5) / 1e6;
    return (float)(v < 0 ? -r : r);
}

void orc_second_order_basis(float xx[27], float yy[27], float zz[27], float xy[27], float yz[27], float xz[27]) {
    /* Tensor products of the cubic B-spline at a node: value {1/6, 2/3, 1/6},
     * first derivative {-1/2, 0, 1/2}, second derivative {1, -2, 1}. */
    const double v[3] = { 1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0 };
    const double d1[3] = { -0.5, 0.0, 0.5 };
    const double d2[3] = { 1.0, -2.0, 1.0 };
    int i = 0;
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a, ++i) {
                xx[i] = round6(d2[a] * v[b] * v[c]);
                yy[i] = round6(v[a] * d2[b] * v[c]);
                zz[i] = round6(v[a] * v[b] * d2[c]);
                xy[i] = round6(d1[a] * d1[b] * v[c]);
                yz[i] = round6(v[a] * d1[b] * d1[c]);
                xz[i] = round6(d1[a] * v[b] * d1[c]);
            }
}

double orc_bending_energy(const orc_geom *cpp, const float *cppData) {
    const size_t g = (size_t)cpp->nx * cpp->ny * cpp->nz;
    float bxx[27], byy[27], bzz[27], bxy[27], byz[27], bxz[27];
    orc_second_order_basis(bxx, byy, bzz, bxy, byz, bxz);
    /* per-node terms in parallel, then one sequential sum in node order: the reference's value
     * at one OpenMP thread (its reduction order, and so its last bits, follow the thread count) */
    double *terms = (double *)calloc(g, sizeof(double));
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 1; z < cpp->nz - 1; ++z)
        for (int y = 1; y < cpp->ny - 1; ++y)
            for (int x = 1; x < cpp->nx - 1; ++x) {
                float d[3][6];
                memset(d, 0, sizeof d);
                int i = 0;
                for (int c = -1; c < 2; ++c)
                    for (int b = -1; b < 2; ++b)
                        for (int a = -1; a < 2; ++a, ++i) {
                            size_t idx = ((size_t)(z + c) * cpp->ny + y + b) * cpp->nx + x + a;
                            for (int k = 0; k < 3; ++k) {
                                float cv = cppData[idx + k * g];
                                d[k][0] += bxx[i] * cv;
                                d[k][1] += byy[i] * cv;
                                d[k][2] += bzz[i] * cv;
                                d[k][3] += bxy[i] * cv;
                                d[k][4] += byz[i] * cv;
                                d[k][5] += bxz[i] * cv;
                            }
                        }
                /* same association as the reference expression (regul.cpp:146-149) */
                double s = (double)(d[0][0] * d[0][0] + d[0][1] * d[0][1] + d[0][2] * d[0][2]) +
                           2.0 * (double)(d[0][3] * d[0][3] + d[0][4] * d[0][4] + d[0][5] * d[0][5]);
                for (int k = 1; k < 3; ++k) {
                    s += (double)(d[k][0] * d[k][0]);
                    s += (double)(d[k][1] * d[k][1]);
                    s += (double)(d[k][2] * d[k][2]);
                    s += 2.0 * (double)(d[k][3] * d[k][3] + d[k][4] * d[k][4] + d[k][5] * d[k][5]);
                }
                terms[((size_t)z * cpp->ny + y) * cpp->nx + x] = s;
            }
    double total = 0;
    for (int z = 1; z < cpp->nz - 1; ++z)
        for (int y = 1; y < cpp->ny - 1; ++y)
            for (int x = 1; x < cpp->nx - 1; ++x) total += terms[((size_t)z * cpp->ny + y) * cpp->nx + x];
    free(terms);
    return total / (double)(3 * g);
}

/* Identity node position in float, as reg_getDisplacementFromDeformation_3D (reg-lib/cpu/_reg_tools.cpp:1988-2031) */
static void node_position(const orc_geom *cpp, int x, int y, int z, float out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = M(cpp->xyz, r, 0) * (float)x + M(cpp->xyz, r, 1) * (float)y + M(cpp->xyz, r, 2) * (float)z +
                 M(cpp->xyz, r, 3);
}

/* The in-place reg_getDisplacementFromDeformation / reg_getDeformationFromDisplacement pair that
 * reg_spline_approxBendingEnergyGradient3D wraps around its work (reg-lib/cpu/_reg_localTrans_regul.cpp:
 * 312, 450): every point becomes (p - node) + node in float. */
void orc_displacement_round_trip(const orc_geom *cpp, float *cppData) {
    const size_t g = (size_t)cpp->nx * cpp->ny * cpp->nz;
    for (int z = 0; z < cpp->nz; ++z)
        for (int y = 0; y < cpp->ny; ++y)
            for (int x = 0; x < cpp->nx; ++x) {
                const size_t n = ((size_t)z * cpp->ny + y) * cpp->nx + x;
                float p[3];
                node_position(cpp, x, y, z, p);
                for (int k = 0; k < 3; ++k) {
                    const float d = cppData[n + k * g] - p[k];
                    cppData[n + k * g] = d + p[k];
                }
            }
}

void orc_bending_energy_gradient(const orc_geom *cpp, const float *cppData, float weight, float *grad) {
    const int nx = cpp->nx, ny = cpp->ny, nz = cpp->nz;
    const size_t g = (size_t)nx * ny * nz;
    float bxx[27], byy[27], bzz[27], bxy[27], byz[27], bxz[27];
    orc_second_order_basis(bxx, byy, bzz, bxy, byz, bxz);
    float *disp = (float *)malloc(3 * g * sizeof(float));
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                size_t idx = ((size_t)z * ny + y) * nx + x;
                float p[3];
                node_position(cpp, x, y, z, p);
                for (int k = 0; k < 3; ++k) disp[idx + k * g] = cppData[idx + k * g] - p[k];
            }
    float *der = (float *)calloc(18 * g, sizeof(float));
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                float d[3][6];
                memset(d, 0, sizeof d);
                int i = 0;
                for (int c = -1; c < 2; ++c)
                    for (int b = -1; b < 2; ++b)
                        for (int a = -1; a < 2; ++a, ++i) {
                            if (x + a < 0 || y + b < 0 || z + c < 0 || x + a >= nx || y + b >= ny || z + c >= nz) continue;
                            size_t idx = ((size_t)(z + c) * ny + y + b) * nx + x + a;
                            for (int k = 0; k < 3; ++k) {
                                float cv = disp[idx + k * g];
                                d[k][0] += bxx[i] * cv;
                                d[k][1] += byy[i] * cv;
                                d[k][2] += bzz[i] * cv;
                                d[k][3] += bxy[i] * cv;
                                d[k][4] += byz[i] * cv;
                                d[k][5] += bxz[i] * cv;
                            }
                        }
                float *o = der + 18 * (((size_t)z * ny + y) * nx + x);
                for (int t = 0; t < 3; ++t)
                    for (int k = 0; k < 3; ++k) o[t * 3 + k] = d[k][t];
                for (int t = 3; t < 6; ++t)
                    for (int k = 0; k < 3; ++k) o[t * 3 + k] = (float)(2.0 * d[k][t]);
            }
    const float ratio = weight / (float)g;
    const float *basis[6] = { bxx, byy, bzz, bxy, byz, bxz };
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                float gv[3] = { 0, 0, 0 };
                int a = 0;
                for (int Z = z - 1; Z < z + 2; ++Z)
                    for (int Y = y - 1; Y < y + 2; ++Y)
                        for (int X = x - 1; X < x + 2; ++X, ++a) {
                            if (X < 0 || Y < 0 || Z < 0 || X >= nx || Y >= ny || Z >= nz) continue;
                            const float *dp = der + 18 * (((size_t)Z * ny + Y) * nx + X);
                            for (int t = 0; t < 6; ++t)
                                for (int k = 0; k < 3; ++k) gv[k] += dp[t * 3 + k] * basis[t][a];
                        }
                size_t idx = ((size_t)z * ny + y) * nx + x;
                for (int k = 0; k < 3; ++k) grad[idx + k * g] += ratio * gv[k];
            }
    free(der);
    free(disp);
}

/* ------------------------------------------------------------------ optimiser helpers */
float orc_max_length(const float *grad, size_t nodes) {
    float mx = 0;
    for (size_t i = 0; i < nodes; ++i) {
        float x = grad[i], y = grad[i + nodes], z = grad[i + 2 * nodes];
        float l = (float)sqrt(x * x + y * y + z * z);
        mx = l > mx ? l : mx;
    }
    return mx;
}

void orc_normalise_gradient(float *grad, size_t nodes, double maxLength) {
    if (maxLength == 0) return;
    double inv = 1.0 / maxLength;
    for (size_t i = 0; i < 3 * nodes; ++i) grad[i] = (float)((double)grad[i] * inv);
}

double orc_cg_update(float *grad, float *g, float *h, size_t n, int first) {
    if (first) {
        for (size_t i = 0; i < n; ++i) h[i] = g[i] = -grad[i];
        return 0;
    }
    double dgg = 0, gg = 0;
    for (size_t i = 0; i < n; ++i) {
        gg += h[i] * g[i];
        dgg += (grad[i] + g[i]) * grad[i];
    }
    double gam = dgg / gg;
    for (size_t i = 0; i < n; ++i) {
        g[i] = -grad[i];
        h[i] = (float)(g[i] + gam * h[i]);
        grad[i] = -h[i];
    }
    return gam;
}

void orc_update_dof(float *cur, const float *best, const float *grad, float scale, size_t n) {
    for (size_t i = 0; i < n; ++i) cur[i] = best[i] + scale * grad[i];
}

/* ------------------------------------------------------------------ affine field */
void orc_affine_field(const orc_geom *ref, const float affine[16], const int *mask, float *field) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    float T[16];
    mat44_mm(affine, ref->xyz, T);
    for (int z = 0; z < ref->nz; ++z)
        for (int y = 0; y < ref->ny; ++y)
            for (int x = 0; x < ref->nx; ++x) {
                size_t idx = ((size_t)z * ref->ny + y) * ref->nx + x;
                if (mask && mask[idx] < 0) continue;
                double v[3] = { (double)x, (double)y, (double)z }, p[3];
                mat44_mul_d(T, v, p);
                field[idx] = (float)p[0];
                field[idx + n] = (float)p[1];
                field[idx + 2 * n] = (float)p[2];
            }
}

/* ------------------------------------------------------------------ Aladin: active blocks */
/* Numerical-Recipes heap sort of values with a companion index array, ascending
 * (reg-lib/cpu/Maths.cpp:114-153).  Tie order is whatever this sift-down sequence produces,
 * so it is restated step for step (1-based heap over v[0..n-1]). */
static void heap_sort_indexed(float *v, int *idx, int n) {
    float *a = v - 1;
    int *ix = idx - 1;
    int l = (n >> 1) + 1, ir = n;
    if (n < 2) return;
    for (;;) {
        float val;
        int iv;
        if (l > 1) {
            --l;
            val = a[l];
            iv = ix[l];
        } else {
            val = a[ir];
            iv = ix[ir];
            a[ir] = a[1];
            ix[ir] = ix[1];
            if (--ir == 1) {
                a[1] = val;
                ix[1] = iv;
                return;
            }
        }
        int i = l, j = l + l;
        while (j <= ir) {
            if (j < ir && a[j] < a[j + 1]) ++j;
            if (val < a[j]) {
                a[i] = a[j];
                ix[i] = ix[j];
                i = j;
                j <<= 1;
            } else break;
        }
        a[i] = val;
        ix[i] = iv;
    }
}

int orc_block_count(const orc_geom *ref, int nb[3]) {
    nb[0] = (ref->nx + 3) / 4;
    nb[1] = (ref->ny + 3) / 4;
    nb[2] = (ref->nz + 3) / 4;
    return nb[0] * nb[1] * nb[2];
}

/* Variance of one 4x4x4 block (reg-lib/cpu/_reg_blockMatching.cpp:70-128): the mean runs over
 * finite voxels with mask > -1, the variance over every finite voxel of the block, both as
 * float sums in z, y, x order.  Returns -1 for an unusable block. */
float orc_block_variance(const orc_geom *ref, const float *img, const int *mask, int bi, int bj, int bk) {
    float mean = 0.f, count = 0.f, var = 0.f;
    for (int z = 4 * bk; z < 4 * bk + 4 && z < ref->nz; ++z)
        for (int y = 4 * bj; y < 4 * bj + 4 && y < ref->ny; ++y)
            for (int x = 4 * bi; x < 4 * bi + 4 && x < ref->nx; ++x) {
                size_t i = ((size_t)z * ref->ny + y) * ref->nx + x;
                float v = img[i];
                if (v == v && (!mask || mask[i] > -1)) {
                    mean += v;
                    count += 1.f;
                }
            }
    mean /= count;
    for (int z = 4 * bk; z < 4 * bk + 4 && z < ref->nz; ++z)
        for (int y = 4 * bj; y < 4 * bj + 4 && y < ref->ny; ++y)
            for (int x = 4 * bi; x < 4 * bi + 4 && x < ref->nx; ++x) {
                float v = img[((size_t)z * ref->ny + y) * ref->nx + x];
                if (v == v) var += (mean - v) * (mean - v);
            }
    var /= count;
    return (count > 32.f && var > 0.f) ? var : -1.f;
}

/* initialise_block_matching_method + _reg_set_active_blocks<float>
 * (reg-lib/cpu/_reg_blockMatching.cpp:213-273, 49-206).  totalBlock[b] = rank of block b among the
 * kept blocks (0 = largest variance) or -1.  Returns the active block number. */
int orc_active_blocks(const orc_geom *ref, const float *img, const int *mask, int percent, int *totalBlock) {
    int nb[3];
    const int total = orc_block_count(ref, nb);
    float *var = (float *)malloc(sizeof(float) * total);
    int *idx = (int *)malloc(sizeof(int) * total);
    int unusable = 0, b = 0;
    for (int k = 0; k < nb[2]; ++k)
        for (int j = 0; j < nb[1]; ++j)
            for (int i = 0; i < nb[0]; ++i, ++b) {
                var[b] = orc_block_variance(ref, img, mask, i, j, k);
                if (var[b] < 0.f) ++unusable;
                idx[b] = b;
            }
    int active = (int)((double)total * ((double)percent / 100.0));
    if (active > total - unusable) active = total - unusable;
    heap_sort_indexed(var, idx, total);
    for (int r = 0; r < total; ++r) totalBlock[idx[total - 1 - r]] = r < active ? r : -1;
    free(var);
    free(idx);
    return active;
}

/* block_matching_method3D<float> (reg-lib/cpu/_reg_blockMatching.cpp:455-690) with a capture
 * range of 3 voxels and step 1.  Positions are written at 3*rank; a block without any offset of
 * more than 32 overlapping voxels gets a NaN warped position.  Returns the defined count. */
int orc_block_matching(const orc_geom *ref, const float *img, const float *warped, const int *mask,
                       const int *totalBlock, float *refPos, float *warpPos) {
    int nb[3];
    orc_block_count(ref, nb);
    int defined = 0, b = 0;
    for (int k = 0; k < nb[2]; ++k)
        for (int j = 0; j < nb[1]; ++j)
            for (int i = 0; i < nb[0]; ++i, ++b) {
                if (totalBlock[b] < 0) continue;
                float rv[64];
                unsigned char rok[64];
                for (int a = 0; a < 64; ++a) {
                    int x = 4 * i + (a & 3), y = 4 * j + ((a >> 2) & 3), z = 4 * k + (a >> 4);
                    rok[a] = 0;
                    rv[a] = 0.f;
                    if (x < ref->nx && y < ref->ny && z < ref->nz) {
                        size_t q = ((size_t)z * ref->ny + y) * ref->nx + x;
                        float v = img[q];
                        if (v == v && (!mask || mask[q] > -1)) { rv[a] = v; rok[a] = 1; }
                    }
                }
                float best = 0.f, disp[3] = { NAN, 0.f, 0.f };
                for (int n = -3; n <= 3; ++n)
                    for (int m = -3; m <= 3; ++m)
                        for (int l = -3; l <= 3; ++l) {
                            float wv[64];
                            unsigned char wok[64];
                            for (int a = 0; a < 64; ++a) {
                                int x = 4 * i + l + (a & 3), y = 4 * j + m + ((a >> 2) & 3), z = 4 * k + n + (a >> 4);
                                wok[a] = 0;
                                wv[a] = 0.f;
                                if (x > -1 && x < ref->nx && y > -1 && y < ref->ny && z > -1 && z < ref->nz) {
                                    float v = warped[((size_t)z * ref->ny + y) * ref->nx + x];
                                    if (v == v) { wv[a] = v; wok[a] = 1; }
                                }
                            }
                            float rm = 0.f, wm = 0.f, cnt = 0.f;
                            for (int a = 0; a < 64; ++a)
                                if (rok[a] && wok[a]) { rm += rv[a]; wm += wv[a]; cnt += 1.f; }
                            if (!(cnt > 32.f)) continue;
                            rm /= cnt;
                            wm /= cnt;
                            float rvar = 0.f, wvar = 0.f, cc = 0.f;
                            for (int a = 0; a < 64; ++a)
                                if (rok[a] && wok[a]) {
                                    float rt = rv[a] - rm, wt = wv[a] - wm;
                                    rvar += rt * rt;
                                    wvar += wt * wt;
                                    cc += rt * wt;
                                }
                            cc = (rvar * wvar) > 0.f ? fabsf(cc / sqrtf(rvar * wvar)) : 0.f;
                            if (cc > best) {
                                best = cc + 1.0e-7f;
                                disp[0] = (float)l;
                                disp[1] = (float)m;
                                disp[2] = (float)n;
                            }
                        }
                float corner[3] = { (float)(4 * i), (float)(4 * j), (float)(4 * k) };
                disp[0] += corner[0];
                disp[1] += corner[1];
                disp[2] += corner[2];
                const int r = 3 * totalBlock[b];
                mat44_mul_f(ref->xyz, corner, refPos + r);
                mat44_mul_f(ref->xyz, disp, warpPos + r);
                if (disp[0] == disp[0]) ++defined;
            }
    return defined;
}

/* ------------------------------------------------------------------ LNCC */
/* reg_tools_kernelConvolution<float> over the three axes of one volume (reg-lib/cpu/_reg_tools.cpp:
 * 832-1112, scalar build): density = (mask >= 0 && finite); intensity and density convolved line by
 * line along x, then y, then z (float products, double line sums); intensity / density at the end,
 * NaN where the density started at 0. */
void orc_kernel_convolution(const orc_geom *g, float *data, float sigma, int kernelType, const int *mask) {
    const int dims[3] = { g->nx, g->ny, g->nz };
    const float pix[3] = { g->dx, g->dy, g->dz };
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    unsigned char *isnan_ = (unsigned char *)malloc(n);
    float *dens = (float *)malloc(n * sizeof(float));
    for (size_t i = 0; i < n; ++i) {
        dens[i] = ((!mask || mask[i] >= 0) && data[i] == data[i]) ? 1.f : 0.f;
        isnan_[i] = dens[i] == 0.f;
        if (isnan_[i]) data[i] = 0;
    }
    for (int axis = 0; axis < 3; ++axis) {
        if (dims[axis] <= 1) continue;
        double temp = sigma > 0 ? (double)(sigma / pix[axis]) : fabs((double)sigma);
        int radius = kernelType == 2 ? (int)(temp * 3.0f) : (int)(temp * 2.0f);
        if (radius <= 0) continue;
        float *kernel = (float *)malloc((2 * radius + 1) * sizeof(float));
        for (int i = -radius; i <= radius; ++i) {
            if (kernelType == 3) {
                double rel = fabs(i / temp);
                if (rel < 1.0) kernel[i + radius] = (float)(2.0 / 3.0 - rel * rel + 0.5 * rel * rel * rel);
                else if (rel < 2.0) kernel[i + radius] = (float)(-((rel - 2.0) * (rel - 2.0) * (rel - 2.0)) / 6.0);
                else kernel[i + radius] = 0;
            } else {
                kernel[i + radius] = (float)(exp(-(double)(i * i) / (2.0 * (temp * temp))) / (temp * 2.506628274631));
            }
        }
        int planes, lineOff;
        if (axis == 0) { planes = dims[1] * dims[2]; lineOff = 1; }
        else if (axis == 1) { planes = dims[0] * dims[2]; lineOff = dims[0]; }
        else { planes = dims[0] * dims[1]; lineOff = planes; }
#pragma omp parallel for schedule(static)
        for (int p = 0; p < planes; ++p) {
            size_t start;
            if (axis == 0) start = (size_t)p * dims[0];
            else if (axis == 1) start = (size_t)(p / dims[0]) * dims[0] * dims[1] + p % dims[0];
            else start = p;
            const int L = dims[axis];
            float *bi = (float *)malloc(L * sizeof(float));
            float *bd = (float *)malloc(L * sizeof(float));
            for (int l = 0; l < L; ++l) {
                bi[l] = data[start + (size_t)l * lineOff];
                bd[l] = dens[start + (size_t)l * lineOff];
            }
            for (int l = 0; l < L; ++l) {
                int pre = l - radius, pst = l + radius + 1;
                const float *kp = kernel;
                if (pre < 0) { kp = kernel - pre; pre = 0; }
                if (pst > L) pst = L;
                double is = 0, ds = 0;
                for (int k = pre; k < pst; ++k) {
                    const float kv = *kp++;
                    is += kv * bi[k];
                    ds += kv * bd[k];
                }
                data[start + (size_t)l * lineOff] = (float)is;
                dens[start + (size_t)l * lineOff] = (float)ds;
            }
            free(bi);
            free(bd);
        }
        free(kernel);
    }
    for (size_t i = 0; i < n; ++i) data[i] = isnan_[i] ? NAN : data[i] / dens[i];
    free(isnan_);
    free(dens);
}

/* UpdateLocalStatImages<float> (reg-lib/cpu/_reg_lncc.cpp:195-245): combined mask (reference mask
 * with the NaN voxels of both images removed), Gaussian-weighted local means and standard
 * deviations of the reference and warped images. */
static void lncc_local_stats(const orc_geom *g, const float *ref, const float *war, const int *mask, float sigma,
                             float *mean, float *sdev, float *wmean, float *wsdev, int *combined) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    for (size_t i = 0; i < n; ++i) {
        combined[i] = mask ? mask[i] : 0;
        if (ref[i] != ref[i]) combined[i] = -1;
        if (war[i] != war[i]) combined[i] = -1;
        mean[i] = ref[i];
        sdev[i] = ref[i] * ref[i];
        wmean[i] = war[i];
        wsdev[i] = war[i] * war[i];
    }
    orc_kernel_convolution(g, mean, sigma, 2, combined);
    orc_kernel_convolution(g, sdev, sigma, 2, combined);
    orc_kernel_convolution(g, wmean, sigma, 2, combined);
    orc_kernel_convolution(g, wsdev, sigma, 2, combined);
    for (size_t i = 0; i < n; ++i) {
        sdev[i] = sqrtf(sdev[i] - mean[i] * mean[i]);
        wsdev[i] = sqrtf(wsdev[i] - wmean[i] * wmean[i]);
        if (sdev[i] < 1.e-06) sdev[i] = 0;
        if (wsdev[i] < 1.e-06) wsdev[i] = 0;
    }
}

/* reg_lncc value for one time point (reg-lib/cpu/_reg_lncc.cpp:247-300): mean |local correlation|
 * over the combined mask, the sum accumulated in voxel order. */
double orc_lncc_value(const orc_geom *g, const float *ref, const float *war, const int *mask, float sigma) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    float *mean = (float *)malloc(n * sizeof(float)), *sdev = (float *)malloc(n * sizeof(float));
    float *wmean = (float *)malloc(n * sizeof(float)), *wsdev = (float *)malloc(n * sizeof(float));
    float *corr = (float *)malloc(n * sizeof(float));
    int *comb = (int *)malloc(n * sizeof(int));
    lncc_local_stats(g, ref, war, mask, sigma, mean, sdev, wmean, wsdev, comb);
    for (size_t i = 0; i < n; ++i) corr[i] = ref[i] * war[i];
    orc_kernel_convolution(g, corr, sigma, 2, comb);
    double sum = 0;
    size_t active = 0;
    for (size_t i = 0; i < n; ++i) {
        if (comb[i] > -1) {
            const double l = (corr[i] - (mean[i] * wmean[i])) / (sdev[i] * wsdev[i]);
            if (l == l && !isinf(l)) {
                sum += fabs(l);
                ++active;
            }
        }
    }
    free(mean); free(sdev); free(wmean); free(wsdev); free(corr); free(comb);
    return sum / (double)active;
}

/* reg_getVoxelBasedLnccGradient<float> (reg-lib/cpu/_reg_lncc.cpp:393-512); grad -= , then
 * non-finite gradient entries are zeroed. */
void orc_lncc_gradient(const orc_geom *g, const float *ref, const float *war, const float *warGrad, const int *mask,
                       float sigma, double weight, float *grad) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    float *mean = (float *)malloc(n * sizeof(float)), *sdev = (float *)malloc(n * sizeof(float));
    float *wmean = (float *)malloc(n * sizeof(float)), *wsdev = (float *)malloc(n * sizeof(float));
    float *corr = (float *)malloc(n * sizeof(float));
    int *comb = (int *)malloc(n * sizeof(int));
    lncc_local_stats(g, ref, war, mask, sigma, mean, sdev, wmean, wsdev, comb);
    for (size_t i = 0; i < n; ++i) corr[i] = ref[i] * war[i];
    orc_kernel_convolution(g, corr, sigma, 2, comb);
    size_t active = 0;
    for (size_t i = 0; i < n; ++i) {
        if (comb[i] > -1) {
            const double rm = mean[i], wm = wmean[i], rs = sdev[i], ws = wsdev[i];
            const double c = corr[i] - (rm * wm);
            double t1 = 1.0 / (rs * ws);
            double t2 = c / (rs * ws * ws * ws);
            double t3 = (c * wm) / (rs * ws * ws * ws) - rm / (rs * ws);
            if (t1 == t1 && !isinf(t1) && t2 == t2 && !isinf(t2) && t3 == t3 && !isinf(t3)) {
                if (c < 0) { t1 = -t1; t2 = -t2; t3 = -t3; }
                wmean[i] = (float)t1;
                wsdev[i] = (float)t2;
                corr[i] = (float)t3;
                ++active;
            } else wmean[i] = wsdev[i] = corr[i] = 0;
        } else wmean[i] = wsdev[i] = corr[i] = 0;
    }
    const double adjusted = weight / (double)active;
    orc_kernel_convolution(g, wmean, sigma, 2, comb);
    orc_kernel_convolution(g, wsdev, sigma, 2, comb);
    orc_kernel_convolution(g, corr, sigma, 2, comb);
    for (size_t i = 0; i < n; ++i) {
        if (comb[i] > -1) {
            const double common = (wmean[i] * ref[i] - wsdev[i] * war[i] + corr[i]) * adjusted;
            for (int c = 0; c < 3; ++c) grad[i + c * n] -= (float)(warGrad[i + c * n] * common);
        }
    }
    for (size_t i = 0; i < 3 * n; ++i)
        if (grad[i] != grad[i] || isinf(grad[i])) grad[i] = 0;
    free(mean); free(sdev); free(wmean); free(wsdev); free(corr); free(comb);
}

/* ------------------------------------------------------------------ linear energy */
/* 3x3 float matrices with the arithmetic of niftilib's nifti_mat33_* helpers
 * (reg-io/niftilib/nifti1_io.c:1807-1960): double intermediates where those use double
 * (including the long double reciprocal of the determinant), float products and sums elsewhere. */

static orc_m33 m33_inverse(orc_m33 R) {
    double r11 = R.m[0][0], r12 = R.m[0][1], r13 = R.m[0][2];
    double r21 = R.m[1][0], r22 = R.m[1][1], r23 = R.m[1][2];
    double r31 = R.m[2][0], r32 = R.m[2][1], r33 = R.m[2][2];
    double deti = r11 * r22 * r33 - r11 * r32 * r23 - r21 * r12 * r33 + r21 * r32 * r13 + r31 * r12 * r23 - r31 * r22 * r13;
    if (deti != 0.0L) deti = 1.0L / deti;
    orc_m33 Q;
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

static float m33_determ(orc_m33 R) {
    double r11 = R.m[0][0], r12 = R.m[0][1], r13 = R.m[0][2];
    double r21 = R.m[1][0], r22 = R.m[1][1], r23 = R.m[1][2];
    double r31 = R.m[2][0], r32 = R.m[2][1], r33 = R.m[2][2];
    return (float)(r11 * r22 * r33 - r11 * r32 * r23 - r21 * r12 * r33 + r21 * r32 * r13 + r31 * r12 * r23 - r31 * r22 * r13);
}

static float m33_rownorm(orc_m33 A) {
    float r1 = (float)(fabs(A.m[0][0]) + fabs(A.m[0][1]) + fabs(A.m[0][2]));
    float r2 = (float)(fabs(A.m[1][0]) + fabs(A.m[1][1]) + fabs(A.m[1][2]));
    float r3 = (float)(fabs(A.m[2][0]) + fabs(A.m[2][1]) + fabs(A.m[2][2]));
    if (r1 < r2) r1 = r2;
    if (r1 < r3) r1 = r3;
    return r1;
}

static float m33_colnorm(orc_m33 A) {
    float r1 = (float)(fabs(A.m[0][0]) + fabs(A.m[1][0]) + fabs(A.m[2][0]));
    float r2 = (float)(fabs(A.m[0][1]) + fabs(A.m[1][1]) + fabs(A.m[2][1]));
    float r3 = (float)(fabs(A.m[0][2]) + fabs(A.m[1][2]) + fabs(A.m[2][2]));
    if (r1 < r2) r1 = r2;
    if (r1 < r3) r1 = r3;
    return r1;
}

static orc_m33 m33_mul(orc_m33 A, orc_m33 B) {
    orc_m33 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
    return C;
}

/* Closest orthogonal matrix by Higham's scaled Newton iteration, as nifti_mat33_polar. */
static orc_m33 m33_polar(orc_m33 A) {
    orc_m33 X = A, Y, Z;
    float alp, bet, gam, gmi, dif = 1.0f;
    int k = 0;
    gam = m33_determ(X);
    while (gam == 0.0) {
        gam = (float)(0.00001 * (0.001 + m33_rownorm(X)));
        X.m[0][0] += gam;
        X.m[1][1] += gam;
        X.m[2][2] += gam;
        gam = m33_determ(X);
    }
    for (;;) {
        Y = m33_inverse(X);
        if (dif > 0.3) {
            alp = (float)(sqrt(m33_rownorm(X) * m33_colnorm(X)));
            bet = (float)(sqrt(m33_rownorm(Y) * m33_colnorm(Y)));
            gam = (float)(sqrt(bet / alp));
            gmi = (float)(1.0 / gam);
        } else {
            gam = gmi = 1.0f;
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Z.m[i][j] = (float)(0.5 * (gam * X.m[i][j] + gmi * Y.m[j][i]));
        double d = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) d += fabs(Z.m[i][j] - X.m[i][j]);
        dif = (float)d;
        k = k + 1;
        if (k > 100 || dif < 3.e-6) break;
        X = Z;
    }
    return Z;
}

/* Six-significant-digit literal of a double (the reference's tables of first-order basis
 * products are written that way). */
static float sig6(double v) {
    if (v == 0) return 0.f;
    const double a = fabs(v);
    const int e = (int)floor(log10(a));
    const double scale = pow(10.0, 5 - e);
    const double r = floor(a * scale + 0.5) / scale;
    return (float)(v < 0 ? -r : r);
}

/* set_first_order_basis_values 3-D (reg-lib/cpu/_reg_splineBasis.cpp:182-290): tensor products of
 * the cubic B-spline first derivative {-1/2, 0, 1/2} and value {1/6, 2/3, 1/6} at a node. */
void orc_first_order_basis(float bx[27], float by[27], float bz[27]) {
    const double v[3] = { 1.0 / 6.0, 2.0 / 3.0, 1.0 / 6.0 }, d1[3] = { -0.5, 0.0, 0.5 };
    int i = 0;
    for (int c = 0; c < 3; ++c)
        for (int b = 0; b < 3; ++b)
            for (int a = 0; a < 3; ++a, ++i) {
                bx[i] = sig6(d1[a] * v[b] * v[c]);
                by[i] = sig6(v[a] * d1[b] * v[c]);
                bz[i] = sig6(v[a] * v[b] * d1[c]);
            }
}

/* Rotation-free displacement Jacobian at an interior node (reg-lib/cpu/_reg_localTrans_regul.cpp:
 * 596-634): first-order basis contraction of the 27 positions, mm -> grid voxel, polar rotation
 * removed, identity subtracted. */
static orc_m33 le_matrix(const orc_geom *cpp, const float *d, int x, int y, int z, const float *bx, const float *by,
                         const float *bz, orc_m33 reo) {
    const size_t g = (size_t)cpp->nx * cpp->ny * cpp->nz;
    orc_m33 M;
    memset(&M, 0, sizeof M);
    int i = 0;
    for (int c = -1; c < 2; ++c)
        for (int b = -1; b < 2; ++b)
            for (int a = -1; a < 2; ++a, ++i) {
                const size_t idx = ((size_t)(z + c) * cpp->ny + y + b) * cpp->nx + x + a;
                const float cx = d[idx], cy = d[idx + g], cz = d[idx + 2 * g];
                M.m[0][0] += bx[i] * cx;
                M.m[1][0] += by[i] * cx;
                M.m[2][0] += bz[i] * cx;
                M.m[0][1] += bx[i] * cy;
                M.m[1][1] += by[i] * cy;
                M.m[2][1] += bz[i] * cy;
                M.m[0][2] += bx[i] * cz;
                M.m[1][2] += by[i] * cz;
                M.m[2][2] += bz[i] * cz;
            }
    M = m33_mul(reo, M);
    const orc_m33 R = m33_inverse(m33_polar(M));
    M = m33_mul(R, M);
    M.m[0][0]--;
    M.m[1][1]--;
    M.m[2][2]--;
    return M;
}

static orc_m33 grid_reorientation(const orc_geom *cpp) {
    orc_m33 r;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) r.m[i][j] = cpp->ijk[i * 4 + j];
    return r;
}

/* reg_spline_approxLinearEnergyValue3D (reg-lib/cpu/_reg_localTrans_regul.cpp:559-647); sums in
 * node order. */
double orc_linear_energy(const orc_geom *cpp, const float *cppData) {
    float bx[27], by[27], bz[27];
    orc_first_order_basis(bx, by, bz);
    const orc_m33 reo = grid_reorientation(cpp);
    double total = 0;
    for (int z = 1; z < cpp->nz - 1; ++z)
        for (int y = 1; y < cpp->ny - 1; ++y)
            for (int x = 1; x < cpp->nx - 1; ++x) {
                const orc_m33 M = le_matrix(cpp, cppData, x, y, z, bx, by, bz, reo);
                double v = 0;
                for (int b = 0; b < 3; ++b)
                    for (int a = 0; a < 3; ++a) {
                        const double s = 0.5 * (M.m[a][b] + M.m[b][a]);
                        v += s * s;
                    }
                total += v;
            }
    return total / (double)(3 * (size_t)cpp->nx * cpp->ny * cpp->nz);
}

/* reg_spline_approxLinearEnergyGradient3D (reg-lib/cpu/_reg_localTrans_regul.cpp:739-824); grad += */
void orc_linear_energy_gradient(const orc_geom *cpp, const float *cppData, float weight, float *grad) {
    float bx[27], by[27], bz[27];
    orc_first_order_basis(bx, by, bz);
    const orc_m33 reo = grid_reorientation(cpp);
    const orc_m33 inv = m33_inverse(reo);
    const size_t g = (size_t)cpp->nx * cpp->ny * cpp->nz;
    const float ratio = weight / (float)g;
    for (int z = 1; z < cpp->nz - 1; ++z)
        for (int y = 1; y < cpp->ny - 1; ++y)
            for (int x = 1; x < cpp->nx - 1; ++x) {
                const orc_m33 M = le_matrix(cpp, cppData, x, y, z, bx, by, bz, reo);
                int i = 26;
                for (int c = -1; c < 2; ++c)
                    for (int b = -1; b < 2; ++b)
                        for (int a = -1; a < 2; ++a, --i) {
                            const size_t idx = ((size_t)(z + c) * cpp->ny + y + b) * cpp->nx + x + a;
                            const float gv0 = -2.f * M.m[0][0] * bx[i];
                            const float gv1 = -2.f * M.m[1][1] * by[i];
                            const float gv2 = -2.f * M.m[2][2] * bz[i];
                            grad[idx] += ratio * (inv.m[0][0] * gv0 + inv.m[0][1] * gv1 + inv.m[0][2] * gv2);
                            grad[idx + g] += ratio * (inv.m[1][0] * gv0 + inv.m[1][1] * gv1 + inv.m[1][2] * gv2);
                            grad[idx + 2 * g] += ratio * (inv.m[2][0] * gv0 + inv.m[2][1] * gv1 + inv.m[2][2] * gv2);
                        }
            }
}

/* ---- symmetric Aladin: matrix inverse and log-Euclidean average ------------------------- */

/* nifti_mat44_inverse (reg-io/niftilib/nifti1_io.c:1680-1717): cofactors in double, the
 * determinant's reciprocal in long double, float results, bottom row (0, 0, 0, 1). */
void orc_nifti_mat44_inverse(const float A[16], float out[16]) {
    const double r11 = A[0], r12 = A[1], r13 = A[2], v1 = A[3];
    const double r21 = A[4], r22 = A[5], r23 = A[6], v2 = A[7];
    const double r31 = A[8], r32 = A[9], r33 = A[10], v3 = A[11];
    double deti = r11 * r22 * r33 - r11 * r32 * r23 - r21 * r12 * r33 + r21 * r32 * r13 + r31 * r12 * r23 -
                  r31 * r22 * r13;
    if (deti != 0.0L) deti = (double)(1.0L / (long double)deti);
    out[0] = (float)(deti * (r22 * r33 - r32 * r23));
    out[1] = (float)(deti * (-r12 * r33 + r32 * r13));
    out[2] = (float)(deti * (r12 * r23 - r22 * r13));
    out[3] = (float)(deti * (-r12 * r23 * v3 + r12 * v2 * r33 + r22 * r13 * v3 - r22 * v1 * r33 - r32 * r13 * v2 +
                             r32 * v1 * r23));
    out[4] = (float)(deti * (-r21 * r33 + r31 * r23));
    out[5] = (float)(deti * (r11 * r33 - r31 * r13));
    out[6] = (float)(deti * (-r11 * r23 + r21 * r13));
    out[7] = (float)(deti * (r11 * r23 * v3 - r11 * v2 * r33 - r21 * r13 * v3 + r21 * v1 * r33 + r31 * r13 * v2 -
                             r31 * v1 * r23));
    out[8] = (float)(deti * (r21 * r32 - r31 * r22));
    out[9] = (float)(deti * (-r11 * r32 + r31 * r12));
    out[10] = (float)(deti * (r11 * r22 - r21 * r12));
    out[11] = (float)(deti * (-r11 * r22 * v3 + r11 * r32 * v2 + r21 * r12 * v3 - r21 * r32 * v1 - r31 * r12 * v2 +
                              r31 * r22 * v1));
    out[12] = out[13] = out[14] = 0.f;
    out[15] = deti == 0.0 ? 0.f : 1.f;
}

/* 4x4 double helpers (row-major). */
static void d4_mul(const double *A, const double *B, double *C) {
    double T[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = 0;
            for (int k = 0; k < 4; ++k) s += A[i * 4 + k] * B[k * 4 + j];
            T[i * 4 + j] = s;
        }
    memcpy(C, T, sizeof(T));
}

static double d4_norm1(const double *A) {
    double best = 0;
    for (int j = 0; j < 4; ++j) {
        double s = 0;
        for (int i = 0; i < 4; ++i) s += fabs(A[i * 4 + j]);
        if (s > best) best = s;
    }
    return best;
}

/* Gauss-Jordan inverse with partial pivoting; returns 0 when singular. */
static int d4_inverse(const double *A, double *X) {
    double M[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            M[i][j] = A[i * 4 + j];
            M[i][j + 4] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (fabs(M[r][c]) > fabs(M[p][c])) p = r;
        if (M[p][c] == 0.0) return 0;
        if (p != c)
            for (int j = 0; j < 8; ++j) {
                const double t = M[c][j];
                M[c][j] = M[p][j];
                M[p][j] = t;
            }
        const double inv = 1.0 / M[c][c];
        for (int j = 0; j < 8; ++j) M[c][j] *= inv;
        for (int r = 0; r < 4; ++r) {
            if (r == c) continue;
            const double f = M[r][c];
            if (f != 0.0)
                for (int j = 0; j < 8; ++j) M[r][j] -= f * M[c][j];
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) X[i * 4 + j] = M[i][j + 4];
    return 1;
}

/* Gauss-Legendre nodes and weights on [0, 1] (Newton on the Legendre polynomial). */
static void gauss_legendre01(int m, double *x, double *w) {
    for (int i = 0; i < m; ++i) {
        double t = cos(3.14159265358979323846 * (i + 0.75) / (m + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = t;
            for (int k = 2; k <= m; ++k) {
                const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = m * (t * p1 - p0) / (t * t - 1.0);
            const double dt = p1 / dp;
            t -= dt;
            if (fabs(dt) < 1e-17) break;
        }
        {
            double p0 = 1.0, p1 = t;
            for (int k = 2; k <= m; ++k) {
                const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = m * (t * p1 - p0) / (t * t - 1.0);
        }
        x[i] = 0.5 * (t + 1.0);
        w[i] = 1.0 / ((1.0 - t * t) * dp * dp);  /* 2/((1-t^2)P'^2) on [-1,1], halved on [0,1] */
    }
}

/* Principal matrix logarithm by inverse scaling and squaring: Denman-Beavers square roots until
 * ||A - I||_1 <= 1/4, then the [8/8] Pade approximant of log(I + X) in its partial-fraction
 * (8-point Gauss-Legendre) form, scaled back by 2^k.  Returns 0 when a square root fails. */
static int d4_logm(const double *A, double *L) {
    double Y[16], Z[16], Yi[16], Zi[16], X[16];
    memcpy(Y, A, sizeof(Y));
    int k = 0;
    for (;;) {
        for (int i = 0; i < 16; ++i) X[i] = Y[i] - ((i % 5) == 0 ? 1.0 : 0.0);
        if (d4_norm1(X) <= 0.25 || k >= 60) break;
        for (int i = 0; i < 16; ++i) Z[i] = (i % 5) == 0 ? 1.0 : 0.0;
        for (int it = 0; it < 100; ++it) {
            if (!d4_inverse(Y, Yi) || !d4_inverse(Z, Zi)) return 0;
            double diff = 0;
            for (int i = 0; i < 16; ++i) {
                const double yn = 0.5 * (Y[i] + Zi[i]);
                diff += fabs(yn - Y[i]);
                Z[i] = 0.5 * (Z[i] + Yi[i]);
                Y[i] = yn;
            }
            if (diff <= 1e-15 * d4_norm1(Y)) break;
        }
        ++k;
    }
    double xg[8], wg[8], S[16] = { 0 }, T[16], Ti[16], P[16];
    gauss_legendre01(8, xg, wg);
    for (int j = 0; j < 8; ++j) {
        for (int i = 0; i < 16; ++i) T[i] = ((i % 5) == 0 ? 1.0 : 0.0) + xg[j] * X[i];
        if (!d4_inverse(T, Ti)) return 0;
        d4_mul(X, Ti, P);
        for (int i = 0; i < 16; ++i) S[i] += wg[j] * P[i];
    }
    const double scale = ldexp(1.0, k);
    for (int i = 0; i < 16; ++i) L[i] = scale * S[i];
    return 1;
}

/* Matrix exponential by scaling and squaring with a degree-20 Taylor polynomial
 * (||A / 2^s||_1 <= 1/2). */
static void d4_expm(const double *A, double *E) {
    const double n = d4_norm1(A);
    int s = 0;
    if (n > 0.5) s = (int)ceil(log2(n / 0.5));
    const double f = ldexp(1.0, -s);
    double B[16], term[16], sum[16];
    for (int i = 0; i < 16; ++i) {
        B[i] = A[i] * f;
        term[i] = (i % 5) == 0 ? 1.0 : 0.0;
        sum[i] = term[i];
    }
    for (int k = 1; k <= 20; ++k) {
        d4_mul(term, B, term);
        for (int i = 0; i < 16; ++i) {
            term[i] /= k;
            sum[i] += term[i];
        }
    }
    for (int i = 0; i < s; ++i) d4_mul(sum, sum, sum);
    memcpy(E, sum, sizeof(sum));
}

/* NiftyReg::Mat44Avg2 (reg-lib/cpu/Maths.cpp:479-491) with Mat44Logm / Mat44Expm (:414-477):
 * the logarithms rounded to float, their bottom rows zeroed, summed in double and rounded to
 * float, halved, exponentiated in double and rounded to float.  Returns 0 if a logarithm fails. */
int orc_mat44_avg2(const float A[16], const float B[16], float out[16]) {
    double a[16], b[16], la[16], lb[16], h[16], e[16];
    for (int i = 0; i < 16; ++i) {
        a[i] = A[i];
        b[i] = B[i];
    }
    if (!d4_logm(a, la) || !d4_logm(b, lb)) return 0;
    for (int i = 0; i < 16; ++i) {
        float fa = (float)la[i], fb = (float)lb[i];
        if (i >= 12) fa = fb = 0.f;
        const float sum = (float)((double)fa + (double)fb);
        h[i] = (double)(float)(sum * 0.5);
    }
    d4_expm(h, e);
    for (int i = 0; i < 16; ++i) out[i] = (float)e[i];
    return 1;
}

/* reg_aladin_sym::UpdateTransformationMatrix (reg-lib/_reg_aladin_sym.cpp:128-157) after both
 * LTS fits: forward <- avg2(forward, inverse(backward)), backward <- avg2(inverse(forward),
 * backward), both bottom rows reset to (0, 0, 0, 1). */
int orc_symmetric_update(float fw[16], float bw[16]) {
    float fi[16], bi[16], nf[16], nb[16];
    orc_nifti_mat44_inverse(fw, fi);
    orc_nifti_mat44_inverse(bw, bi);
    if (!orc_mat44_avg2(fw, bi, nf) || !orc_mat44_avg2(fi, bw, nb)) return 0;
    for (int i = 12; i < 15; ++i) nf[i] = nb[i] = 0.f;
    nf[15] = nb[15] = 1.f;
    memcpy(fw, nf, sizeof(nf));
    memcpy(bw, nb, sizeof(nb));
    return 1;
}

/* ------------------------------------------------------------------ velocity fields (F3D2) */
/* The scaling-and-squaring exponentiation of a cubic B-spline velocity grid and the operations
 * reg_f3d2 builds on it.  Grids and fields hold positions (deformations) in mm. */

/* Voxel -> mm position in float, as reg_getDeformationFromDisplacement_3D adds it
 * (reg-lib/cpu/_reg_tools.cpp:2107-2150): float products and sums, left to right. */
static void identity_pos(const float *m, int x, int y, int z, float out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = M(m, r, 0) * (float)x + M(m, r, 1) * (float)y + M(m, r, 2) * (float)z + M(m, r, 3);
}

/* mm -> voxel in float (the real-to-voxel products of reg-lib/cpu/_reg_localTrans.cpp:2384-2396
 * and 940-952) */
static void to_voxel(const float *m, const float p[3], float out[3]) {
    for (int r = 0; r < 3; ++r) out[r] = M(m, r, 0) * p[0] + M(m, r, 1) * p[1] + M(m, r, 2) * p[2] + M(m, r, 3);
}

static int clampi(int v, int n) { return v < 0 ? 0 : (v >= n ? n - 1 : v); }

/* get_SlidedValues 3-D (reg-lib/cpu/_reg_splineBasis.cpp:496-551) for a deformation (not a
 * displacement): the value at the clamped index plus the mm offset of the clamped steps. */
static void slided(const float *d, size_t n, int nx, int ny, int nz, const float *xyz, int x, int y, int z,
                   float out[3]) {
    const int cx = clampi(x, nx), cy = clampi(y, ny), cz = clampi(z, nz);
    const int sx = x - cx, sy = y - cy, sz = z - cz;
    const size_t idx = ((size_t)cz * ny + cy) * nx + cx;
    for (int r = 0; r < 3; ++r) {
        const float shift = (float)sx * M(xyz, r, 0) + (float)sy * M(xyz, r, 1) + (float)sz * M(xyz, r, 2);
        out[r] = d[idx + r * n] + shift;
    }
}

/* get_GridValues 3-D (reg-lib/cpu/_reg_splineBasis.cpp:602-665) with sliding outside the grid,
 * including its flag that, once a row falls outside, reads the rest of that slice through the
 * sliding path (which returns the same values inside). */
static void grid_values_slided(const orc_geom *cpp, const float *d, int x0, int y0, int z0, float *gx, float *gy,
                               float *gz) {
    const size_t g = (size_t)cpp->nx * cpp->ny * cpp->nz;
    int coord = 0;
    for (int Z = z0; Z < z0 + 4; ++Z) {
        int out = !(Z > -1 && Z < cpp->nz);
        for (int Y = y0; Y < y0 + 4; ++Y) {
            if (!(Y > -1 && Y < cpp->ny && !out)) out = 1;
            for (int X = x0; X < x0 + 4; ++X, ++coord) {
                if (X > -1 && X < cpp->nx && !out) {
                    const size_t idx = ((size_t)Z * cpp->ny + Y) * cpp->nx + X;
                    gx[coord] = d[idx];
                    gy[coord] = d[idx + g];
                    gz[coord] = d[idx + 2 * g];
                } else {
                    float v[3];
                    slided(d, g, cpp->nx, cpp->ny, cpp->nz, cpp->xyz, X, Y, Z, v);
                    gx[coord] = v[0];
                    gy[coord] = v[1];
                    gz[coord] = v[2];
                }
            }
        }
    }
}

/* reg_spline_getFlowFieldFromVelocityGrid (reg-lib/cpu/_reg_localTrans.cpp:3465-3491) for a grid
 * without affine extensions: the identity field of the reference geometry, then the composition
 * branch of reg_cubic_spline_getDeformationField3D (:903-1079, scalar order) evaluates the grid at
 * those positions.  flow: planar [3][N] positions. */
/* Mat44Mul<double> (reg-lib/cpu/Maths.hpp:272-286) of a position, rounded to float. */
static void mul_d(const float *m, double x, double y, double z, float out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = (float)((double)M(m, r, 0) * x + (double)M(m, r, 1) * y + (double)M(m, r, 2) * z + (double)M(m, r, 3));
}
static void vel_flow(const orc_geom *ref, const orc_geom *cpp, const float *vel, const float *ext0, float *flow);
void orc_vel_flow(const orc_geom *ref, const orc_geom *cpp, const float *vel, float *flow) {
    vel_flow(ref, cpp, vel, NULL, flow);
}
/* With a grid extension 0 the identity is first composed with it (reg_affine_deformationField3D
 * in composition mode, reg-lib/cpu/_reg_localTrans.cpp:1478-1487). */
static void vel_flow(const orc_geom *ref, const orc_geom *cpp, const float *vel, const float *ext0, float *flow) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    size_t i = 0;
    for (int z = 0; z < ref->nz; ++z)
        for (int y = 0; y < ref->ny; ++y)
            for (int x = 0; x < ref->nx; ++x, ++i) {
                float real[3], vox[3], xb[4], yb[4], zb[4], gx[64], gy[64], gz[64];
                identity_pos(ref->xyz, x, y, z, real);
                if (ext0) {
                    const float id[3] = { real[0], real[1], real[2] };
                    mul_d(ext0, id[0], id[1], id[2], real);
                }
                to_voxel(cpp->ijk, real, vox);
                int pre[3];
                float *bases[3] = { xb, yb, zb };
                for (int r = 0; r < 3; ++r) {
                    pre[r] = floor_int(vox[r]);
                    float t = vox[r] - (float)pre[r];
                    pre[r]--;
                    if (t < 0) t = 0;
                    orc_bspline_basis(t, bases[r]);
                }
                grid_values_slided(cpp, vel, pre[0], pre[1], pre[2], gx, gy, gz);
                float r0 = 0, r1 = 0, r2 = 0;
                int coord = 0;
                for (int c = 0; c < 4; ++c)
                    for (int b = 0; b < 4; ++b)
                        for (int a = 0; a < 4; ++a, ++coord) {
                            const float w = xb[a] * yb[b] * zb[c];
                            r0 += gx[coord] * w;
                            r1 += gy[coord] * w;
                            r2 += gz[coord] * w;
                        }
                flow[i] = r0;
                flow[i + n] = r1;
                flow[i + 2 * n] = r2;
            }
}

/* reg_defField_compose3D without a mask (reg-lib/cpu/_reg_localTrans.cpp:2333-2457): every
 * position of `upd` is replaced by the trilinear interpolation of `def` (same geometry g) at that
 * position, with sliding outside the field.  upd may alias nothing else. */
void orc_def_field_compose(const orc_geom *g, const float *def, float *upd) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    for (size_t i = 0; i < n; ++i) {
        const float real[3] = { upd[i], upd[i + n], upd[i + 2 * n] };
        float vox[3];
        to_voxel(g->ijk, real, vox);
        int pre[3];
        float rel[3][2];
        for (int r = 0; r < 3; ++r) {
            pre[r] = floor_int(vox[r]);
            rel[r][1] = vox[r] - (float)pre[r];
            rel[r][0] = 1.f - rel[r][1];
        }
        float acc[3] = { 0, 0, 0 };
        for (int c = 0; c < 2; ++c) {
            const int Z = pre[2] + c;
            const int inZ = Z > -1 && Z < g->nz;
            for (int b = 0; b < 2; ++b) {
                const int Y = pre[1] + b;
                const float tb = rel[1][b] * rel[2][c];
                const int inY = Y > -1 && Y < g->ny;
                for (int a = 0; a < 2; ++a) {
                    const int X = pre[0] + a;
                    float v[3];
                    if (X > -1 && X < g->nx && inY && inZ) {
                        const size_t idx = ((size_t)Z * g->ny + Y) * g->nx + X;
                        v[0] = def[idx];
                        v[1] = def[idx + n];
                        v[2] = def[idx + 2 * n];
                    } else {
                        slided(def, n, g->nx, g->ny, g->nz, g->xyz, X, Y, Z, v);
                    }
                    const float w = rel[0][a] * tb;
                    acc[0] += v[0] * w;
                    acc[1] += v[1] * w;
                    acc[2] += v[2] * w;
                }
            }
        }
        upd[i] = acc[0];
        upd[i + n] = acc[1];
        upd[i + 2 * n] = acc[2];
    }
}

/* The flow as a displacement scaled by 2^-steps (sign of *steps: direction) back in deformation
 * form: reg_getDisplacementFromDeformation, reg_tools_divideValueToImage and
 * reg_getDeformationFromDisplacement (reg-lib/cpu/_reg_localTrans.cpp:3513-3561). */
static void scaled_flow(const orc_geom *ref, float *flow, int steps, int update, int *stepsOut,
                        const float *affOnly) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    size_t i = 0;
    for (int z = 0; z < ref->nz; ++z)
        for (int y = 0; y < ref->ny; ++y)
            for (int x = 0; x < ref->nx; ++x, ++i) {
                float p[3];
                identity_pos(ref->xyz, x, y, z, p);
                for (int r = 0; r < 3; ++r) {
                    /* reg_tools_subtractImageFromImage (double) of the affine-only field, or
                     * reg_getDisplacementFromDeformation (float) */
                    if (affOnly) flow[i + r * n] = (float)((double)flow[i + r * n] - (double)affOnly[i + r * n]);
                    else flow[i + r * n] -= p[r];
                }
            }
    int sq = 1;
    if (update || steps == 0) {
        /* reg-lib/cpu/_reg_localTrans.cpp:3520-3544: the largest |displacement| sets the number of
         * squarings (each step below 0.28 mm in 3-D, 0.35 mm in 2-D), at least 6 */
        float mn = flow[0], mx = flow[0];
        for (size_t k = 1; k < 3 * n; ++k) {
            if (flow[k] < mn) mn = flow[k];
            if (mx < flow[k]) mx = flow[k];
        }
        float extrema = fabsf(mn) < mx ? mx : fabsf(mn);
        const float maxLength = ref->nz > 1 ? 0.28f : 0.35f;
        while (extrema / (float)pow(2, sq) >= maxLength) sq++;
        if (sq < 6) sq = 6;
        steps = steps >= 0 ? sq : -sq;
    } else {
        sq = steps < 0 ? -steps : steps;
    }
    const double scaling = steps < 0 ? -(double)(float)pow(2, sq) : (double)(float)pow(2, sq);
    i = 0;
    for (int z = 0; z < ref->nz; ++z)
        for (int y = 0; y < ref->ny; ++y)
            for (int x = 0; x < ref->nx; ++x, ++i) {
                float p[3];
                identity_pos(ref->xyz, x, y, z, p);
                for (int r = 0; r < 3; ++r) flow[i + r * n] = (float)((double)flow[i + r * n] / scaling) + p[r];
            }
    *stepsOut = steps;
}

/* reg_spline_getDefFieldFromVelocityGrid for a velocity grid without affine extensions
 * (reg-lib/cpu/_reg_localTrans.cpp:3583-3613, 3493-3581): flow field, displacement scaled by
 * 2^-steps, then `steps` self-compositions.  *steps is the grid's intent_p2: read, and updated
 * when `update` is set or it is 0. */
void orc_vel_def_field(const orc_geom *ref, const orc_geom *cpp, const float *vel, int *steps, int update,
                       float *field) {
    orc_vel_def_field_aff(ref, cpp, vel, NULL, NULL, steps, update, field);
}

/* The affine-only field of reg_affine_deformationField3D without composition
 * (reg-lib/cpu/_reg_globalTrans.cpp:74-133): (A * sform) i in double, rounded to float. */
static float *affine_only(const orc_geom *g, const float *A) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    float *out = (float *)malloc(3 * n * sizeof(float));
    float T[16];
    mat44_mm(A, g->xyz, T);
    size_t i = 0;
    for (int z = 0; z < g->nz; ++z)
        for (int y = 0; y < g->ny; ++y)
            for (int x = 0; x < g->nx; ++x, ++i) {
                float p[3];
                mul_d(T, x, y, z, p);
                for (int r = 0; r < 3; ++r) out[i + r * n] = p[r];
            }
    return out;
}

/* After the squaring (reg-lib/cpu/_reg_localTrans.cpp:3569-3577, 3662-3683): with an extension 0
 * the displacement (float) plus the affine-only field (double); then extension 1 composed
 * (Mat44Mul<double>). */
static void affine_parts(const orc_geom *ref, float *fields, int count, const float *affOnly, const float *ext1) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    for (int k = 0; k < count; ++k) {
        float *F = fields + (size_t)k * 3 * n;
        size_t i = 0;
        for (int z = 0; z < ref->nz; ++z)
            for (int y = 0; y < ref->ny; ++y)
                for (int x = 0; x < ref->nx; ++x, ++i) {
                    float v[3] = { F[i], F[i + n], F[i + 2 * n] };
                    if (affOnly) {
                        float p[3];
                        identity_pos(ref->xyz, x, y, z, p);
                        for (int r = 0; r < 3; ++r) {
                            const float d = v[r] - p[r];
                            v[r] = (float)((double)d + (double)affOnly[i + r * n]);
                        }
                    }
                    if (ext1) {
                        const float w[3] = { v[0], v[1], v[2] };
                        mul_d(ext1, w[0], w[1], w[2], v);
                    }
                    F[i] = v[0];
                    F[i + n] = v[1];
                    F[i + 2 * n] = v[2];
                }
    }
}

void orc_vel_def_field_aff(const orc_geom *ref, const orc_geom *cpp, const float *vel, const float *ext0,
                           const float *ext1, int *steps, int update, float *field) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    float *flow = (float *)malloc(3 * n * sizeof(float));
    float *affOnly = ext0 ? affine_only(ref, ext0) : NULL;
    vel_flow(ref, cpp, vel, ext0, flow);
    int s = *steps;
    scaled_flow(ref, flow, s, update, &s, affOnly);
    *steps = s;
    const int sq = s < 0 ? -s : s;
    memcpy(field, flow, 3 * n * sizeof(float));
    for (int k = 0; k < sq; ++k) {
        orc_def_field_compose(ref, field, flow);
        memcpy(field, flow, 3 * n * sizeof(float));
    }
    affine_parts(ref, field, 1, affOnly, ext1);
    free(affOnly);
    free(flow);
}

/* reg_spline_getIntermediateDefFieldFromVelGrid without affine extensions
 * (reg-lib/cpu/_reg_localTrans.cpp:3616-3683): fields[0] is the scaled flow, fields[k+1] the
 * composition of fields[k] with itself; (|steps| + 1) planar [3][N] fields. */
void orc_vel_intermediate(const orc_geom *ref, const orc_geom *cpp, const float *vel, int steps, float *fields) {
    orc_vel_intermediate_aff(ref, cpp, vel, NULL, NULL, steps, fields);
}

void orc_vel_intermediate_aff(const orc_geom *ref, const orc_geom *cpp, const float *vel, const float *ext0,
                              const float *ext1, int steps, float *fields) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    float *affOnly = ext0 ? affine_only(ref, ext0) : NULL;
    vel_flow(ref, cpp, vel, ext0, fields);
    int s;
    scaled_flow(ref, fields, steps, 0, &s, affOnly);
    const int sq = steps < 0 ? -steps : steps;
    for (int k = 0; k < sq; ++k) {
        memcpy(fields + (size_t)(k + 1) * 3 * n, fields + (size_t)k * 3 * n, 3 * n * sizeof(float));
        orc_def_field_compose(ref, fields + (size_t)k * 3 * n, fields + (size_t)(k + 1) * 3 * n);
    }
    affine_parts(ref, fields, sq + 1, affOnly, ext1);
    free(affOnly);
}

/* reg_getRealImageSpacing (reg-lib/cpu/_reg_tools.cpp:253-274): column lengths of the sform from
 * Mat44Mul<float> images of voxels 0 and e_k. */
static void real_spacing(const float *xyz, float sp[3]) {
    const float o[3] = { 0, 0, 0 };
    float r1[3];
    mat44_mul_f(xyz, o, r1);
    for (int k = 0; k < 3; ++k) {
        float e[3] = { 0, 0, 0 }, r2[3];
        e[k] = 1;
        mat44_mul_f(xyz, e, r2);
        const float a = r1[0] - r2[0], b = r1[1] - r2[1], c = r1[2] - r2[2];
        sp[k] = sqrtf(a * a + b * b + c * c);
    }
}

/* reg_trilinearResampleGradient (reg-lib/cpu/_reg_resampling.cpp:1547-1766): the vector image
 * `grad` (geometry g) sampled trilinearly at the positions `def` (same geometry), each sample
 * multiplied by the field's forward-difference Jacobian (rotation of the sform removed, divided
 * by the real spacing).  Padding for out-of-image taps. */
void orc_resample_gradient(const orc_geom *g, const float *grad, const float *def, float pad, float *out) {
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    orc_m33 s;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) s.m[i][j] = g->xyz[i * 4 + j];
    const orc_m33 reo = m33_inverse(m33_polar(s));
    float sp[3];
    real_spacing(g->xyz, sp);
    const float basis[2] = { 1, 0 }, deriv[2] = { -1, 1 };
    size_t i = 0;
    for (int z = 0; z < g->nz; ++z)
        for (int y = 0; y < g->ny; ++y)
            for (int x = 0; x < g->nx; ++x, ++i) {
                const float d[3] = { def[i], def[i + n], def[i + 2 * n] };
                float fc[3];
                to_voxel(g->ijk, d, fc);
                int lo[3], hi[3];
                float bw[3][2];
                for (int r = 0; r < 3; ++r) {
                    lo[r] = floor_int(fc[r]);
                    const int t = (int)fc[r];
                    hi[r] = t + (fc[r] > (float)t);
                    bw[r][1] = fabsf(fc[r] - (float)lo[r]);
                    bw[r][0] = 1 - bw[r][1];
                }
                float vx = 0, vy = 0, vz = 0;
                for (int c = 0; c < 2; ++c) {
                    const int Z = c ? hi[2] : lo[2];
                    if (Z > -1 && Z < g->nz) {
                        for (int b = 0; b < 2; ++b) {
                            const int Y = b ? hi[1] : lo[1];
                            if (Y > -1 && Y < g->ny) {
                                for (int a = 0; a < 2; ++a) {
                                    const int X = a ? hi[0] : lo[0];
                                    const float w = bw[0][a] * bw[1][b] * bw[2][c];
                                    if (X > -1 && X < g->nx) {
                                        const size_t idx = ((size_t)Z * g->ny + Y) * g->nx + X;
                                        vx += grad[idx] * w;
                                        vy += grad[idx + n] * w;
                                        vz += grad[idx + 2 * n] * w;
                                    } else {
                                        vx += pad * w;
                                        vy += pad * w;
                                        vz += pad * w;
                                    }
                                }
                            } else {
                                vx += pad * bw[1][b] * bw[2][c];
                                vy += pad * bw[1][b] * bw[2][c];
                                vz += pad * bw[1][b] * bw[2][c];
                            }
                        }
                    } else {
                        vx += pad * bw[2][c];
                        vy += pad * bw[2][c];
                        vz += pad * bw[2][c];
                    }
                }
                /* forward differences of the field; at the last index of an axis the pair steps
                 * back by one (and the other two derivatives are then read one voxel back too) */
                orc_m33 J;
                memset(&J, 0, sizeof J);
                for (int c = 0; c < 2; ++c) {
                    int Z = z + c;
                    float bz0 = basis[c], bz1 = deriv[c];
                    if (z == g->nz - 1) {
                        if (c == 1) Z -= 2;
                        bz0 = fabsf(bz0 - 1);
                        bz1 *= -1;
                    }
                    for (int b = 0; b < 2; ++b) {
                        int Y = y + b;
                        float by0 = basis[b], by1 = deriv[b];
                        if (y == g->ny - 1) {
                            if (b == 1) Y -= 2;
                            by0 = fabsf(by0 - 1);
                            by1 *= -1;
                        }
                        for (int a = 0; a < 2; ++a) {
                            int X = x + a;
                            float bx0 = basis[a], bx1 = deriv[a];
                            if (x == g->nx - 1) {
                                if (a == 1) X -= 2;
                                bx0 = fabsf(bx0 - 1);
                                bx1 *= -1;
                            }
                            const float w0 = bx1 * by0 * bz0, w1 = bx0 * by1 * bz0, w2 = bx0 * by0 * bz1;
                            const size_t idx = ((size_t)Z * g->ny + Y) * g->nx + X;
                            const float dx = def[idx], dy = def[idx + n], dz = def[idx + 2 * n];
                            J.m[0][0] += w0 * dx;
                            J.m[0][1] += w1 * dx;
                            J.m[0][2] += w2 * dx;
                            J.m[1][0] += w0 * dy;
                            J.m[1][1] += w1 * dy;
                            J.m[1][2] += w2 * dy;
                            J.m[2][0] += w0 * dz;
                            J.m[2][1] += w1 * dz;
                            J.m[2][2] += w2 * dz;
                        }
                    }
                }
                J = m33_mul(reo, J);
                for (int r = 0; r < 3; ++r)
                    for (int k = 0; k < 3; ++k) J.m[r][k] /= sp[k];
                out[i] = J.m[0][0] * vx + J.m[0][1] * vy + J.m[0][2] * vz;
                out[i + n] = J.m[1][0] * vx + J.m[1][1] * vy + J.m[1][2] * vz;
                out[i + 2 * n] = J.m[2][0] * vx + J.m[2][1] * vy + J.m[2][2] * vz;
            }
}

/* Compute::ExponentiateGradient without affine components (reg-lib/Compute.cpp:283-326): the
 * backward grid's intermediate fields at the forward reference geometry; for k < |steps|,
 * grad <- resample(grad, fields[k]) + grad; finally grad / 2^|steps|. */
void orc_exponentiate_gradient(const orc_geom *ref, const orc_geom *cppBw, const float *velBw, int steps,
                               float *grad) {
    orc_exponentiate_gradient_aff(ref, cppBw, velBw, NULL, NULL, NULL, steps, grad);
}

/* With the backward content's transformation T, its displacement ((T * sform) i in double,
 * minus x in float) is subtracted (double) from each intermediate field first
 * (reg-lib/Compute.cpp:303-316). */
void orc_exponentiate_gradient_aff(const orc_geom *ref, const orc_geom *cppBw, const float *velBw,
                                   const float *ext0, const float *ext1, const float *transBw, int steps,
                                   float *grad) {
    const size_t n = (size_t)ref->nx * ref->ny * ref->nz;
    const int sq = steps < 0 ? -steps : steps;
    float *fields = (float *)malloc((size_t)(sq + 1) * 3 * n * sizeof(float));
    float *tmp = (float *)malloc(3 * n * sizeof(float));
    orc_vel_intermediate_aff(ref, cppBw, velBw, ext0, ext1, steps, fields);
    if (transBw) {
        float *disp = affine_only(ref, transBw);
        size_t i = 0;
        for (int z = 0; z < ref->nz; ++z)
            for (int y = 0; y < ref->ny; ++y)
                for (int x = 0; x < ref->nx; ++x, ++i) {
                    float p[3];
                    identity_pos(ref->xyz, x, y, z, p);
                    for (int r = 0; r < 3; ++r) disp[i + r * n] -= p[r];
                }
        for (int k = 0; k < sq; ++k)
            for (size_t j = 0; j < 3 * n; ++j) {
                float *F = fields + (size_t)k * 3 * n;
                F[j] = (float)((double)F[j] - (double)disp[j]);
            }
        free(disp);
    }
    for (int k = 0; k < sq; ++k) {
        orc_resample_gradient(ref, grad, fields + (size_t)k * 3 * n, 0.f, tmp);
        for (size_t i = 0; i < 3 * n; ++i) grad[i] = (float)((double)tmp[i] + (double)grad[i]);
    }
    const double div = pow(2, sq);
    for (size_t i = 0; i < 3 * n; ++i) grad[i] = (float)((double)grad[i] / div);
    free(fields);
    free(tmp);
}

/* Compute::UpdateVelocityField (reg-lib/Compute.cpp:336-348): scaled = gradient * scale (zeroed
 * on the axes not optimised), grid = grid + scaled; reg_tools arithmetic in double. */
void orc_update_velocity(float *grid, const float *grad, float scale, size_t nodes, int ox, int oy, int oz) {
    const int on[3] = { ox, oy, oz };
    for (int c = 0; c < 3; ++c)
        for (size_t i = 0; i < nodes; ++i) {
            const float s = on[c] ? (float)((double)grad[i + c * nodes] * (double)scale) : 0.f;
            grid[i + c * nodes] = (float)((double)grid[i + c * nodes] + (double)s);
        }
}

/* Compute::SymmetriseVelocityFields (reg-lib/Compute.cpp:357-391): both grids (same geometry)
 * to displacements, fw <- (fw - bw) / 2 and bw <- (bw - fw) / 2 from the values before the
 * update, back to positions. */
void orc_symmetrise_velocity(const orc_geom *cpp, float *fw, float *bw) {
    const size_t g = (size_t)cpp->nx * cpp->ny * cpp->nz;
    size_t i = 0;
    for (int z = 0; z < cpp->nz; ++z)
        for (int y = 0; y < cpp->ny; ++y)
            for (int x = 0; x < cpp->nx; ++x, ++i) {
                float p[3];
                identity_pos(cpp->xyz, x, y, z, p);
                for (int r = 0; r < 3; ++r) {
                    const float f = fw[i + r * g] - p[r], b = bw[i + r * g] - p[r];
                    const float nf = (float)((double)f - (double)b), nb = (float)((double)b - (double)f);
                    fw[i + r * g] = (float)((double)nf * 0.5) + p[r];
                    bw[i + r * g] = (float)((double)nb * 0.5) + p[r];
                }
            }
}

/* reg_createSymmetricControlPointGrids without an affine input (reg-lib/cpu/_reg_localTrans.cpp:
 * 132-358): the grid covers both images' corners in mm, axis-aligned, sform only.  Writes the
 * grid geometry; fw and bw (may be NULL) receive the identity positions. */
void orc_symmetric_grids(const orc_geom *ref, const orc_geom *flo, const float spacing[3], orc_geom *grid,
                         float *fw, float *bw) {
    orc_symmetric_grids_aff(ref, flo, spacing, NULL, grid, fw, bw);
}

/* The half affines of reg_createSymmetricControlPointGrids (reg-lib/cpu/_reg_localTrans.cpp:
 * 150-166): Mat44Logm (float -> double -> float), times 0.5, Mat44Expm; the backward half from
 * the inverse.  Returns 1 when both differ from the identity (the grids then carry them as
 * extensions 0 and 1, :320-353), 0 otherwise, -1 when a logarithm fails. */
int orc_half_affines(const float A[16], float hf[16], float hb[16]) {
    float inv[16];
    orc_nifti_mat44_inverse(A, inv);
    const float *src[2] = { A, inv };
    float *dst[2] = { hf, hb };
    for (int t = 0; t < 2; ++t) {
        double m[16], l[16], h[16], e[16];
        for (int i = 0; i < 16; ++i) m[i] = src[t][i];
        if (!d4_logm(m, l)) return -1;
        for (int i = 0; i < 16; ++i) h[i] = (double)(float)((double)(float)l[i] * 0.5);
        d4_expm(h, e);
        for (int i = 0; i < 16; ++i) dst[t][i] = (float)e[i];
    }
    int idf = 1, idb = 1;
    for (int i = 0; i < 16; ++i) {
        const float one = (i % 5) == 0 ? 1.f : 0.f;
        if (hf[i] != one) idf = 0;
        if (hb[i] != one) idb = 0;
    }
    return !idf && !idb;
}

/* With an affine the image spaces are first multiplied by the half affines (operator*(mat44,
 * mat44), reg-lib/cpu/Maths.hpp:237-246). */
void orc_symmetric_grids_aff(const orc_geom *ref, const orc_geom *flo, const float spacing[3], const float *affine,
                             orc_geom *grid, float *fw, float *bw) {
    float mn[3], mx[3], refSpace[16], floSpace[16];
    memcpy(refSpace, ref->xyz, sizeof refSpace);
    memcpy(floSpace, flo->xyz, sizeof floSpace);
    if (affine) {
        float hf[16], hb[16];
        orc_half_affines(affine, hf, hb);
        mat44_mm(hf, ref->xyz, refSpace);
        mat44_mm(hb, flo->xyz, floSpace);
    }
    for (int c = 0; c < 8; ++c) {
        const float rc[3] = { (c & 1) ? (float)ref->nx : 0.f, (c & 2) ? (float)ref->ny : 0.f,
                              (c & 4) ? (float)ref->nz : 0.f };
        const float fc[3] = { (c & 1) ? (float)flo->nx : 0.f, (c & 2) ? (float)flo->ny : 0.f,
                              (c & 4) ? (float)flo->nz : 0.f };
        float ro[3], fo[3];
        mat44_mul_f(refSpace, rc, ro);
        mat44_mul_f(floSpace, fc, fo);
        for (int r = 0; r < 3; ++r) {
            if (c == 0) {
                mn[r] = ro[r] < fo[r] ? ro[r] : fo[r];
                mx[r] = ro[r] < fo[r] ? fo[r] : ro[r];
            } else {
                if (ro[r] < mn[r]) mn[r] = ro[r];
                if (fo[r] < mn[r]) mn[r] = fo[r];
                if (mx[r] < ro[r]) mx[r] = ro[r];
                if (mx[r] < fo[r]) mx[r] = fo[r];
            }
        }
    }
    int dims[3];
    for (int r = 0; r < 3; ++r) {
        const float v = (mx[r] - mn[r]) / spacing[r] + 3.f;
        const int t = (int)v;
        dims[r] = t + (v > (float)t);
    }
    memset(grid, 0, sizeof *grid);
    grid->nx = dims[0];
    grid->ny = dims[1];
    grid->nz = dims[2];
    grid->dx = spacing[0];
    grid->dy = spacing[1];
    grid->dz = spacing[2];
    for (int r = 0; r < 3; ++r) {
        M(grid->xyz, r, r) = spacing[r];
        M(grid->xyz, r, 3) = mn[r] - spacing[r];
    }
    M(grid->xyz, 3, 3) = 1.f;
    orc_nifti_mat44_inverse(grid->xyz, grid->ijk);
    const size_t g = (size_t)dims[0] * dims[1] * dims[2];
    size_t i = 0;
    for (int z = 0; z < dims[2]; ++z)
        for (int y = 0; y < dims[1]; ++y)
            for (int x = 0; x < dims[0]; ++x, ++i) {
                float p[3];
                identity_pos(grid->xyz, x, y, z, p);
                for (int r = 0; r < 3; ++r) {
                    if (fw) fw[i + r * g] = p[r];
                    if (bw) bw[i + r * g] = p[r];
                }
            }
}
