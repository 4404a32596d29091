// 4x4 matrix functions in double shared by device code and host code of the library: products,
// Gauss-Jordan inverse, logarithm (inverse scaling and squaring: Denman-Beavers square roots,
// then the [8/8] Pade approximant of log(I + X) in Gauss-Legendre form), exponential (Taylor
// scaling and squaring), and niftilib's float mat44 inverse.  They stand in for the Eigen
// logm / expm that NiftyReg::Mat44Logm / Mat44Expm call (reg-lib/cpu/Maths.cpp:414-478).
#pragma once

#include <cmath>

#include "recip_x87.hpp"

namespace nrh {

struct GaussLegendre8 {
    double x[8], w[8];  // nodes and weights on [0, 1]
};

NRH_HOST_DEVICE inline void d4_mul(const double *A, const double *B, double *C) {
    double T[16];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double acc = 0;
            for (int k = 0; k < 4; ++k) acc = std::fma(A[i * 4 + k], B[k * 4 + j], acc);
            T[i * 4 + j] = acc;
        }
    for (int i = 0; i < 16; ++i) C[i] = T[i];
}

NRH_HOST_DEVICE inline double d4_norm1(const double *A) {
    double best = 0;
    for (int j = 0; j < 4; ++j) {
        const double c = std::fabs(A[j]) + std::fabs(A[4 + j]) + std::fabs(A[8 + j]) + std::fabs(A[12 + j]);
        best = c > best ? c : best;
    }
    return best;
}

// Gauss-Jordan with partial pivoting; false when singular.
NRH_HOST_DEVICE inline bool d4_inverse(const double *A, double *X) {
    double M[4][8];
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            M[i][j] = A[i * 4 + j];
            M[i][j + 4] = i == j ? 1.0 : 0.0;
        }
    for (int c = 0; c < 4; ++c) {
        int p = c;
        for (int r = c + 1; r < 4; ++r)
            if (std::fabs(M[r][c]) > std::fabs(M[p][c])) p = r;
        if (M[p][c] == 0.0) return false;
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
            for (int j = 0; j < 8; ++j) M[r][j] = std::fma(-f, M[c][j], M[r][j]);
        }
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) X[i * 4 + j] = M[i][j + 4];
    return true;
}

NRH_HOST_DEVICE inline bool d4_logm(const double *A, const GaussLegendre8& gl, double *L) {
    double Y[16], Z[16], Yi[16], Zi[16], X[16];
    for (int i = 0; i < 16; ++i) Y[i] = A[i];
    int k = 0;
    for (;;) {
        for (int i = 0; i < 16; ++i) X[i] = Y[i] - ((i % 5) == 0 ? 1.0 : 0.0);
        if (d4_norm1(X) <= 0.25) break;
        if (k >= 60) return false;
        for (int i = 0; i < 16; ++i) Z[i] = (i % 5) == 0 ? 1.0 : 0.0;
        for (int it = 0; it < 100; ++it) {
            if (!d4_inverse(Y, Yi) || !d4_inverse(Z, Zi)) return false;
            double diff = 0;
            for (int i = 0; i < 16; ++i) {
                const double yn = 0.5 * (Y[i] + Zi[i]);
                diff += std::fabs(yn - Y[i]);
                Z[i] = 0.5 * (Z[i] + Yi[i]);
                Y[i] = yn;
            }
            if (diff <= 1e-15 * d4_norm1(Y)) break;
        }
        ++k;
    }
    double S[16] = {}, T[16], Ti[16], P[16];
    for (int j = 0; j < 8; ++j) {
        for (int i = 0; i < 16; ++i) T[i] = std::fma(gl.x[j], X[i], (i % 5) == 0 ? 1.0 : 0.0);
        if (!d4_inverse(T, Ti)) return false;
        d4_mul(X, Ti, P);
        for (int i = 0; i < 16; ++i) S[i] = std::fma(gl.w[j], P[i], S[i]);
    }
    const double scale = std::ldexp(1.0, k);
    for (int i = 0; i < 16; ++i) L[i] = scale * S[i];
    return true;
}

NRH_HOST_DEVICE inline void d4_expm(const double *A, double *E) {
    const double n = d4_norm1(A);
    int s = 0;
    if (n > 0.5) s = (int)std::ceil(std::log2(n / 0.5));
    const double f = std::ldexp(1.0, -s);
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
    for (int i = 0; i < 16; ++i) E[i] = sum[i];
}

// nifti_mat44_inverse (reg-io/niftilib/nifti1_io.c:1680-1717): cofactors in double, the
// determinant's reciprocal rounded as x87 long double then double, float results.
NRH_HOST_DEVICE inline void nifti_inverse(const float *R, float *Q) {
#pragma clang fp contract(off)
    const double r11 = R[0], r12 = R[1], r13 = R[2], v1 = R[3];
    const double r21 = R[4], r22 = R[5], r23 = R[6], v2 = R[7];
    const double r31 = R[8], r32 = R[9], r33 = R[10], v3 = R[11];
    double deti = r11 * r22 * r33 - r11 * r32 * r23 - r21 * r12 * r33 + r21 * r32 * r13 + r31 * r12 * r23 -
                  r31 * r22 * r13;
    if (deti != 0.0) deti = recip_x87(deti);  // niftilib's 1.0l / deti
    Q[0] = (float)(deti * (r22 * r33 - r32 * r23));
    Q[1] = (float)(deti * (-r12 * r33 + r32 * r13));
    Q[2] = (float)(deti * (r12 * r23 - r22 * r13));
    Q[3] = (float)(deti * (-r12 * r23 * v3 + r12 * v2 * r33 + r22 * r13 * v3 - r22 * v1 * r33 - r32 * r13 * v2 +
                           r32 * v1 * r23));
    Q[4] = (float)(deti * (-r21 * r33 + r31 * r23));
    Q[5] = (float)(deti * (r11 * r33 - r31 * r13));
    Q[6] = (float)(deti * (-r11 * r23 + r21 * r13));
    Q[7] = (float)(deti * (r11 * r23 * v3 - r11 * v2 * r33 - r21 * r13 * v3 + r21 * v1 * r33 + r31 * r13 * v2 -
                           r31 * v1 * r23));
    Q[8] = (float)(deti * (r21 * r32 - r31 * r22));
    Q[9] = (float)(deti * (-r11 * r32 + r31 * r12));
    Q[10] = (float)(deti * (r11 * r22 - r21 * r12));
    Q[11] = (float)(deti * (-r11 * r22 * v3 + r11 * r32 * v2 + r21 * r12 * v3 - r21 * r32 * v1 - r31 * r12 * v2 +
                            r31 * r22 * v1));
    Q[12] = Q[13] = Q[14] = 0.f;
    Q[15] = deti == 0.0 ? 0.f : 1.f;
}

// Nodes and weights of the 8-point Gauss-Legendre rule on [0, 1] (Newton on P_8).
inline GaussLegendre8 gauss_legendre8() {
    GaussLegendre8 g;
    const int m = 8;
    auto legendre = [&](double t, double& pm, double& dpm) {
        double p0 = 1.0, p1 = t;
        for (int k = 2; k <= m; ++k) {
            const double p2 = ((2.0 * k - 1.0) * t * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        pm = p1;
        dpm = m * (t * p1 - p0) / (t * t - 1.0);
    };
    for (int i = 0; i < m; ++i) {
        double t = std::cos(3.14159265358979323846 * (i + 0.75) / (m + 0.5)), p, dp;
        for (int it = 0; it < 100; ++it) {
            legendre(t, p, dp);
            const double dt = p / dp;
            t -= dt;
            if (std::fabs(dt) < 1e-17) break;
        }
        legendre(t, p, dp);
        g.x[i] = 0.5 * (t + 1.0);
        g.w[i] = 1.0 / ((1.0 - t * t) * dp * dp);
    }
    return g;
}

}  // namespace nrh
