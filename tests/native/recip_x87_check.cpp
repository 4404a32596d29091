// Checks nrh::recip_x87 against the x87 long-double reciprocal rounded to double, on random
// doubles of many binades.  Prints "<checked> <mismatches> <cases where plain 1/d differs>".
#include <cstdio>
#include <cstdint>
#include <random>

#include "recip_x87.hpp"

int main(int argc, char **argv) {
    const long n = argc > 1 ? std::atol(argv[1]) : 4000000;
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> mant(1.0, 2.0);
    std::uniform_int_distribution<int> expo(-60, 60);
    long bad = 0, plainDiffers = 0;
    for (long i = 0; i < n; ++i) {
        double d = std::ldexp(mant(rng), expo(rng));
        if (i & 1) d = -d;
        volatile long double q = 1.0L / (long double)d;
        const double want = (double)q;
        const double got = nrh::recip_x87(d);
        bad += got != want;
        plainDiffers += (1.0 / d) != want;
    }
    std::printf("%ld %ld %ld\n", n, bad, plainDiffers);
    return 0;
}
