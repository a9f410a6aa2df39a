// Small math helpers shared by device and host code.
#pragma once
#include <cmath>

#if defined(__HIPCC__)
#define SF_HD __host__ __device__
#else
#define SF_HD
#endif

namespace sf {

// x^p for a small integer exponent p >= 0, rounded once from a double-double
// product (error ~2^-104 relative before the final rounding), so the result is
// the correctly rounded power except in vanishingly rare near-tie cases.  The
// contact force evaluates pow(sigma/r, power) with an integer power
// (membrane/repulsion.cpp:15).  glibc's pow (0.52 ulp) agrees with the
// correctly rounded value for all but about 1 in 1200 arguments of x^4 over
// x in [0.28, 6.7]; this gives the correctly rounded value without depending
// on the device libm's ulp behaviour.
SF_HD inline double pow_int_cr(double x, int p) {
    if (p == 0) return 1.0;
    if (p == 1) return x;
    if (!(x == x) || x == 0.0 || x * 0.0 != 0.0) return std::pow(x, static_cast<double>(p));
    double rh = 1.0, rl = 0.0;    // result  = rh + rl
    double bh = x, bl = 0.0;      // base    = bh + bl
    int e = p;
    while (e > 0) {
        if (e & 1) {
            double h = rh * bh;
            double l = std::fma(rh, bh, -h) + (rh * bl + rl * bh);
            rh = h + l;
            rl = l - (rh - h);
        }
        e >>= 1;
        if (e) {
            double h = bh * bh;
            double l = std::fma(bh, bh, -h) + 2.0 * bh * bl;
            bh = h + l;
            bl = l - (bh - h);
        }
    }
    return rh + rl;
}

}  // namespace sf
