// 3x3 float matrices with niftilib's nifti_mat33_* arithmetic (reg-io/niftilib/nifti1_io.c:
// 1807-1960), shared by device code and host code of the library.
#pragma once

#include "recip_x87.hpp"

namespace nrh {

// Double intermediates where niftilib uses double, float products and sums elsewhere.
// niftilib takes the determinant's reciprocal in long double (80-bit on x86-64); recip_x87
// reproduces that rounding in double arithmetic.
struct M33 {
    float m[3][3];
};

NRH_HOST_DEVICE inline M33 m33_inverse(const M33& R) {
#pragma clang fp contract(off)
    const double r11 = R.m[0][0], r12 = R.m[0][1], r13 = R.m[0][2];
    const double r21 = R.m[1][0], r22 = R.m[1][1], r23 = R.m[1][2];
    const double r31 = R.m[2][0], r32 = R.m[2][1], r33 = R.m[2][2];
    double deti = r11 * r22 * r33 - r11 * r32 * r23 - r21 * r12 * r33 + r21 * r32 * r13 + r31 * r12 * r23 -
                  r31 * r22 * r13;
    if (deti != 0.0) deti = recip_x87(deti);  // niftilib's 1.0l / deti
    M33 Q;
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

NRH_HOST_DEVICE inline float m33_determ(const M33& R) {
#pragma clang fp contract(off)
    const double r11 = R.m[0][0], r12 = R.m[0][1], r13 = R.m[0][2];
    const double r21 = R.m[1][0], r22 = R.m[1][1], r23 = R.m[1][2];
    const double r31 = R.m[2][0], r32 = R.m[2][1], r33 = R.m[2][2];
    return (float)(r11 * r22 * r33 - r11 * r32 * r23 - r21 * r12 * r33 + r21 * r32 * r13 + r31 * r12 * r23 -
                   r31 * r22 * r13);
}

NRH_HOST_DEVICE inline float m33_rownorm(const M33& A) {
#pragma clang fp contract(off)
    float r = 0.f;
    for (int i = 0; i < 3; ++i) {
        const float v = (float)(fabs((double)A.m[i][0]) + fabs((double)A.m[i][1]) + fabs((double)A.m[i][2]));
        if (r < v || i == 0) r = v;
    }
    return r;
}

NRH_HOST_DEVICE inline float m33_colnorm(const M33& A) {
#pragma clang fp contract(off)
    float r = 0.f;
    for (int j = 0; j < 3; ++j) {
        const float v = (float)(fabs((double)A.m[0][j]) + fabs((double)A.m[1][j]) + fabs((double)A.m[2][j]));
        if (r < v || j == 0) r = v;
    }
    return r;
}

NRH_HOST_DEVICE inline M33 m33_mul(const M33& A, const M33& B) {
#pragma clang fp contract(off)
    M33 C;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) C.m[i][j] = A.m[i][0] * B.m[0][j] + A.m[i][1] * B.m[1][j] + A.m[i][2] * B.m[2][j];
    return C;
}

// Higham's scaled Newton iteration for the closest orthogonal matrix (nifti_mat33_polar).
NRH_HOST_DEVICE inline M33 m33_polar(const M33& A) {
#pragma clang fp contract(off)
    M33 X = A, Y, Z;
    float gam = m33_determ(X), gmi, dif = 1.0f;
    int k = 0;
    for (int guard = 0; gam == 0.0f && guard < 1000; ++guard) {
        gam = (float)(0.00001 * (0.001 + (double)m33_rownorm(X)));
        X.m[0][0] += gam;
        X.m[1][1] += gam;
        X.m[2][2] += gam;
        gam = m33_determ(X);
    }
    for (;;) {
        Y = m33_inverse(X);
        if (dif > 0.3) {
            const float alp = (float)sqrt((double)(m33_rownorm(X) * m33_colnorm(X)));
            const float bet = (float)sqrt((double)(m33_rownorm(Y) * m33_colnorm(Y)));
            gam = (float)sqrt((double)(bet / alp));
            gmi = (float)(1.0 / (double)gam);
        } else {
            gam = gmi = 1.0f;
        }
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) Z.m[i][j] = (float)(0.5 * (double)(gam * X.m[i][j] + gmi * Y.m[j][i]));
        double d = 0;
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) d += fabs((double)(Z.m[i][j] - X.m[i][j]));
        dif = (float)d;
        k = k + 1;
        if (k > 100 || dif < 3.e-6) break;
        X = Z;
    }
    return Z;
}

}  // namespace nrh
