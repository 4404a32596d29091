// The reciprocal niftilib computes when it inverts a matrix: `deti = 1.0l / deti` with a double
// deti, i.e. the quotient rounded to the 64-bit significand of x87 long double and then to double
// (reg-io/niftilib/nifti1_io.c:1693 and the mat33 inverse).  That double rounding differs from
// the correctly rounded double reciprocal when the 64-bit rounding lands exactly on a midpoint
// between two doubles (about one quotient in 2^11); the tie then goes to the even significand.
// This function reproduces it in double arithmetic: the exact remainder of 1/d from one fused
// multiply-add locates the quotient relative to the midpoint next to the rounded reciprocal.
// Shared by device code (hipcc) and a host unit test (g++).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>

#if defined(__HIPCC__)
#define NRH_HOST_DEVICE __host__ __device__
#else
#define NRH_HOST_DEVICE
#endif

namespace nrh {

NRH_HOST_DEVICE inline double recip_x87(double d) {
    const double r = 1.0 / d;
    if (r == 0.0 || !std::isfinite(r)) return r;
    const double e = std::fma(-r, d, 1.0);  // 1 - r d, exact
    if (e == 0.0) return r;
    const double delta = e / d;  // q - r, where q = 1/d exactly
    int64_t bits;
    std::memcpy(&bits, &r, sizeof(bits));
    // neighbour of r on q's side (the significand grows when delta has r's sign)
    const int64_t nbits = ((delta > 0.0) == (r > 0.0)) ? bits + 1 : bits - 1;
    double n;
    std::memcpy(&n, &nbits, sizeof(n));
    const double half = std::fabs(n - r) * 0.5;  // distance from r to the midpoint
    // q rounds to the midpoint at 64 bits when it lies within half a 64-bit ulp of it
    if (std::fabs(half - std::fabs(delta)) < std::ldexp(half, -11)) return (bits & 1) ? n : r;
    return r;
}

}  // namespace nrh
