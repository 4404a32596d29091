// nr_oracle_lts.cpp -- CPU restatement of Aladin's least-trimmed-squares update (test
// infrastructure only; part of liboracle.so, see nr_oracle.h).
//
// optimize (reg-lib/cpu/_reg_blockMatching.cpp:722-821) and optimize_3D
// (reg-lib/cpu/_reg_globalTrans.cpp:328-394).  The trimming step uses std::nth_element with the
// same comparator as the reference, so the kept subset and its order follow the same standard
// library routine.  The fits are restated from the mathematics: the affine fit is the FP64
// least-squares solution of [x y z 1] P = w by Householder QR (the reference uses Eigen's
// column-pivoting QR, reg-lib/cpu/Maths.cpp:325-349, so the two agree to FP64 rounding); the
// rigid fit is Kabsch with a one-sided Jacobi SVD (reg-lib/cpu/Maths.cpp:351-390).
#include "nr_oracle.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <utility>
#include <vector>

namespace {

struct Pt { float r[3], w[3]; };

void mat44_mul_f(const float *m, const float in[3], float out[3]) {
    for (int r = 0; r < 3; ++r)
        out[r] = (float)((double)m[r * 4] * in[0] + (double)m[r * 4 + 1] * in[1] + (double)m[r * 4 + 2] * in[2] +
                         (double)m[r * 4 + 3]);
}

void eye(float m[16]) {
    std::memset(m, 0, 16 * sizeof(float));
    m[0] = m[5] = m[10] = m[15] = 1.f;
}

// Least squares [r 1] P = w, P 4x3, by Householder QR in double.
void fit_affine(const std::vector<Pt>& p, float m[16]) {
    const size_t n = p.size();
    std::vector<double> A(n * 4), B(n * 3);
    for (size_t i = 0; i < n; ++i) {
        for (int c = 0; c < 3; ++c) {
            A[i * 4 + c] = p[i].r[c];
            B[i * 3 + c] = p[i].w[c];
        }
        A[i * 4 + 3] = 1.0;
    }
    double R[4][4] = {};
    for (int c = 0; c < 4; ++c) {
        double norm = 0;
        for (size_t i = c; i < n; ++i) norm += A[i * 4 + c] * A[i * 4 + c];
        norm = std::sqrt(norm);
        const double alpha = A[c * 4 + c] > 0 ? -norm : norm;
        // v = x - alpha e1, stored in place
        A[c * 4 + c] -= alpha;
        double vv = 0;
        for (size_t i = c; i < n; ++i) vv += A[i * 4 + c] * A[i * 4 + c];
        if (vv > 0) {
            for (int k = c + 1; k < 4; ++k) {
                double d = 0;
                for (size_t i = c; i < n; ++i) d += A[i * 4 + c] * A[i * 4 + k];
                const double f = 2 * d / vv;
                for (size_t i = c; i < n; ++i) A[i * 4 + k] -= f * A[i * 4 + c];
            }
            for (int k = 0; k < 3; ++k) {
                double d = 0;
                for (size_t i = c; i < n; ++i) d += A[i * 4 + c] * B[i * 3 + k];
                const double f = 2 * d / vv;
                for (size_t i = c; i < n; ++i) B[i * 3 + k] -= f * A[i * 4 + c];
            }
        }
        R[c][c] = alpha;
        for (int k = c + 1; k < 4; ++k) R[c][k] = A[c * 4 + k];
    }
    double P[4][3];
    for (int k = 0; k < 3; ++k)
        for (int c = 3; c >= 0; --c) {
            double s = B[c * 3 + k];
            for (int j = c + 1; j < 4; ++j) s -= R[c][j] * P[j][k];
            P[c][k] = s / R[c][c];
        }
    eye(m);
    for (int r = 0; r < 3; ++r) {
        for (int c = 0; c < 3; ++c) m[r * 4 + c] = (float)P[c][r];
        m[r * 4 + 3] = (float)P[3][r];
    }
}

// H = U S V^T for a 3x3 H by one-sided Jacobi; singular values sorted descending.
void svd3(const double H[3][3], double U[3][3], double S[3], double V[3][3]) {
    double A[3][3];
    std::memcpy(A, H, sizeof(A));
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) V[i][j] = i == j;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                double a = 0, b = 0, c = 0;
                for (int i = 0; i < 3; ++i) {
                    a += A[i][p] * A[i][p];
                    b += A[i][q] * A[i][q];
                    c += A[i][p] * A[i][q];
                }
                if (std::fabs(c) <= 1e-300 || std::fabs(c) <= 1e-15 * std::sqrt(a * b)) continue;
                off = std::max(off, std::fabs(c) / std::sqrt(a * b));
                const double zeta = (b - a) / (2 * c);
                const double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1 + zeta * zeta));
                const double cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
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
    int order[3] = { 0, 1, 2 };
    double s[3];
    for (int j = 0; j < 3; ++j) s[j] = std::sqrt(A[0][j] * A[0][j] + A[1][j] * A[1][j] + A[2][j] * A[2][j]);
    std::sort(order, order + 3, [&](int x, int y) { return s[x] > s[y]; });
    double V2[3][3];
    for (int j = 0; j < 3; ++j) {
        const int o = order[j];
        S[j] = s[o];
        for (int i = 0; i < 3; ++i) {
            U[i][j] = s[o] > 0 ? A[i][o] / s[o] : 0.0;
            V2[i][j] = V[i][o];
        }
    }
    std::memcpy(V, V2, sizeof(V2));
}

// Kabsch: R = V U^T of the cross-covariance of the centred clouds, reflection guarded.
void fit_rigid(const std::vector<Pt>& p, float m[16]) {
    const size_t n = p.size();
    double c1[3] = {}, c2[3] = {};
    for (const Pt& q : p)
        for (int k = 0; k < 3; ++k) { c1[k] += q.r[k]; c2[k] += q.w[k]; }
    for (int k = 0; k < 3; ++k) { c1[k] /= (double)n; c2[k] /= (double)n; }
    double H[3][3] = {};
    for (const Pt& q : p)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) H[a][b] += (q.r[a] - c1[a]) * (q.w[b] - c2[b]);
    double U[3][3], S[3], V[3][3], R[3][3];
    svd3(H, U, S, V);
    auto vut = [&] {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) R[a][b] = V[a][0] * U[b][0] + V[a][1] * U[b][1] + V[a][2] * U[b][2];
    };
    vut();
    const double det = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) -
                       R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
                       R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
    if (det < 0) {
        for (int i = 0; i < 3; ++i) V[i][2] = -V[i][2];
        vut();
    }
    eye(m);
    for (int a = 0; a < 3; ++a) {
        for (int b = 0; b < 3; ++b) m[a * 4 + b] = (float)R[a][b];
        m[a * 4 + 3] = (float)(c2[a] - (R[a][0] * c1[0] + R[a][1] * c1[1] + R[a][2] * c1[2]));
    }
}

} // namespace

extern "C" int orc_lts(const float *refPos, const float *warpPos, int nActive, int percentToKeep, int affine,
                       float matrix[16]) {
    std::vector<Pt> pts;
    for (int i = 0; i < nActive; ++i) {
        const float *w = warpPos + 3 * i;
        if (!(w[0] == w[0])) continue;
        Pt q;
        std::memcpy(q.r, refPos + 3 * i, sizeof(q.r));
        mat44_mul_f(matrix, w, q.w);
        pts.push_back(q);
    }
    if ((int)pts.size() < (affine ? 8 : 4)) return -1;
    auto fit = [&](const std::vector<Pt>& p, float *m) { affine ? fit_affine(p, m) : fit_rigid(p, m); };
    float M[16], last[16];
    fit(pts, M);
    const unsigned num = (unsigned)pts.size();
    const unsigned long keep = (unsigned long)(num * (percentToKeep / 100.0f));
    std::vector<std::pair<double, unsigned>> res(num);
    std::vector<Pt> top;
    double lastDistance = std::numeric_limits<double>::max();
    std::memset(last, 0, sizeof(last));
    for (int it = 0; it < 30; ++it) {
        for (unsigned p = 0; p < num; ++p) {
            float t[3];
            mat44_mul_f(M, pts[p].r, t);
            const float dx = t[0] - pts[p].w[0], dy = t[1] - pts[p].w[1], dz = t[2] - pts[p].w[2];
            res[p] = { (double)std::sqrt(dx * dx + dy * dy + dz * dz), p };
        }
        std::nth_element(res.begin(), res.begin() + keep, res.end(),
                         [](const std::pair<double, unsigned>& a, const std::pair<double, unsigned>& b) {
                             return a.first < b.first;
                         });
        double distance = 0;
        top.clear();
        for (unsigned long k = 0; k < keep; ++k) {
            top.push_back(pts[res[k].second]);
            distance += res[k].first;
        }
        if (distance > lastDistance || (lastDistance - distance) < 1e-3) {
            std::memcpy(M, last, sizeof(M));
            break;
        }
        lastDistance = distance;
        std::memcpy(last, M, sizeof(M));
        fit(top, M);
    }
    std::memcpy(matrix, M, sizeof(M));
    return 0;
}
