// FP64 ln(u) on (0,1] and (sin, cos)(2 pi u) on [0,1) for the Box-Muller
// thermal field.  Both arguments are uniforms in the unit interval, so the
// general library routines' range handling is not needed:
//   ln u    : u = m 2^e with m in [sqrt(1/2), sqrt(2)), s = (m-1)/(m+1),
//             ln m = 2 atanh(s) = 2s (1 + s^2/3 + s^4/5 + ...) (|s| <= 0.172,
//             series to s^25), ln 2 split hi/lo (the fdlibm constants);
//   sincos  : 8u = octant + f exactly, angle (pi/4) f or (pi/4)(1-f) in
//             [0, pi/4], Taylor series to t^17 / t^18, octant symmetries.
// Accuracy (checked against libm on the host, tests/test_rngmath.py): a few
// ulp.  Plain C++ so the same source compiles for the host test.
#pragma once
#ifndef MJ_HD
#if defined(__HIPCC_RTC__)
#define MJ_HD __host__ __device__ __forceinline__
#elif defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MJ_HD __host__ __device__ __forceinline__
#else
#define MJ_HD static inline
#include <math.h>
#endif
#endif

MJ_HD double mj_log01(double u)
{
    int e;
    double m = frexp(u, &e);                 // u = m 2^e, m in [0.5, 1)
    if (m < 0.70710678118654752440) {
        m = 2.0 * m;
        e = e - 1;
    }
    const double s = (m - 1.0) / (m + 1.0); // m - 1 exact
    const double z = s * s;
    double p = 1.0 / 25.0;
    p = fma(p, z, 1.0 / 23.0);
    p = fma(p, z, 1.0 / 21.0);
    p = fma(p, z, 1.0 / 19.0);
    p = fma(p, z, 1.0 / 17.0);
    p = fma(p, z, 1.0 / 15.0);
    p = fma(p, z, 1.0 / 13.0);
    p = fma(p, z, 1.0 / 11.0);
    p = fma(p, z, 1.0 / 9.0);
    p = fma(p, z, 1.0 / 7.0);
    p = fma(p, z, 1.0 / 5.0);
    p = fma(p, z, 1.0 / 3.0);
    const double s2 = 2.0 * s;
    const double lm = fma(s2, z * p, s2);
    const double de = (double)e;
    return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, lm));
}

MJ_HD void mj_sincos2pi(double u, double *sn, double *cs)
{
    const double x8 = 8.0 * u;               // exact
    const double q = floor(x8);
    const double f = x8 - q;                 // exact, [0, 1)
    const int o = (int)q & 7;
    const double g = (o & 1) ? (1.0 - f) : f;
    const double t = g * 0.78539816339744830962;   // pi/4
    const double t2 = t * t;
    double ps = -1.0 / 355687428096000.0;    // -1/17!
    ps = fma(ps, t2, 1.0 / 1307674368000.0);
    ps = fma(ps, t2, -1.0 / 6227020800.0);
    ps = fma(ps, t2, 1.0 / 39916800.0);
    ps = fma(ps, t2, -1.0 / 362880.0);
    ps = fma(ps, t2, 1.0 / 5040.0);
    ps = fma(ps, t2, -1.0 / 120.0);
    ps = fma(ps, t2, 1.0 / 6.0);
    // sin t = t - t^3 (1/6 - t^2/120 + ...)
    const double st = fma(-(t * t2), ps, t);
    double pc = 1.0 / 6402373705728000.0;    // 1/18!
    pc = fma(pc, t2, -1.0 / 20922789888000.0);
    pc = fma(pc, t2, 1.0 / 87178291200.0);
    pc = fma(pc, t2, -1.0 / 479001600.0);
    pc = fma(pc, t2, 1.0 / 3628800.0);
    pc = fma(pc, t2, -1.0 / 40320.0);
    pc = fma(pc, t2, 1.0 / 720.0);
    pc = fma(pc, t2, -1.0 / 24.0);
    pc = fma(pc, t2, 0.5);
    // cos t = 1 - t^2 (1/2 - t^2/24 + ...)
    const double ct = fma(-t2, pc, 1.0);
    const int swap = ((o + 1) >> 1) & 1;
    const double sa = swap ? ct : st;
    const double ca = swap ? st : ct;
    *sn = (o >= 4) ? -sa : sa;
    *cs = (((o + 2) >> 2) & 1) ? -ca : ca;
}
