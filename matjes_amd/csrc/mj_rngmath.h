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
//
// The thermal normals use the word forms below (mj_ln_u32, mj_sincos2pi_u32,
// mj_cos2pi_u32, mj_sqrt_pos): the uniform is taken from its 32-bit word by
// integer operations (octant and fraction of the angle, exponent offset of the
// logarithm), the logarithm's quotient and the square root start from the
// hardware reciprocal / reciprocal square root with fma corrections on the
// device (IEEE on the host), and the third normal needs a cosine only (one
// sine polynomial on [0, pi/2]).  Accuracy: a few ulp (host: tests of these
// functions against libm; device: the normals against the oracle's libm ones).
#pragma once
#include <stdint.h>
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

// a / b for b in [1.7, 2.5] (the logarithm's quotient): on the device the
// hardware reciprocal, one Newton step and one quotient correction (~0.5 ulp)
MJ_HD double mj_div_log(double a, double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(b);
    const double e = fma(-b, r, 1.0);
    r = fma(r, e, r);
    const double q = a * r;
    const double d = fma(-b, q, a);
    return fma(d, r, q);
#else
    return a / b;
#endif
}

// sqrt(y) for y = 0 or y in [1e-10, 50] (the Box-Muller radius): the hardware
// reciprocal square root, a Goldschmidt step and one correction on the device
MJ_HD double mj_sqrt_pos(double y)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double z = __builtin_amdgcn_rsq(y);
    double g = y * z;
    double h = z * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, y);
    g = fma(d, h, g);
    return y == 0.0 ? y : g;
#else
    return sqrt(y);
#endif
}

// ln((w + 1) 2^-32): x = w + 1 = m 2^e exactly, m in [sqrt(1/2), sqrt(2)),
// ln x = e ln 2 + 2 atanh(s), s = (m - 1)/(m + 1), |s| <= 0.172 (series to s^21)
MJ_HD double mj_ln_u32(uint32_t w)
{
    const double x = (double)w + 1.0;
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m = 2.0 * m;
        e = e - 1;
    }
    const double s = mj_div_log(m - 1.0, m + 1.0);
    const double z = s * s;
    double p = 1.0 / 21.0;
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
    const double de = (double)(e - 32);
    return fma(de, 6.93147180369123816490e-01, fma(de, 1.90821492927058770002e-10, lm));
}

// (sin, cos)(2 pi w 2^-32): octant = top 3 bits, fraction f = the low 29 bits
// 2^-29 (exact); angle (pi/4) f or (pi/4)(1 - f) in [0, pi/4] (series to t^15 / t^16)
MJ_HD void mj_sincos2pi_u32(uint32_t w, double *sn, double *cs)
{
    const int o = (int)(w >> 29);
    const double f = (double)(w & 0x1FFFFFFFu) * 1.862645149230957031250e-09;   // 2^-29
    const double g = (o & 1) ? (1.0 - f) : f;
    const double t = g * 0.78539816339744830962;   // pi/4
    const double t2 = t * t;
    double ps = 1.0 / 1307674368000.0;     // 1/15!
    ps = fma(ps, t2, -1.0 / 6227020800.0);
    ps = fma(ps, t2, 1.0 / 39916800.0);
    ps = fma(ps, t2, -1.0 / 362880.0);
    ps = fma(ps, t2, 1.0 / 5040.0);
    ps = fma(ps, t2, -1.0 / 120.0);
    ps = fma(ps, t2, 1.0 / 6.0);
    const double st = fma(-(t * t2), ps, t);
    double pc = -1.0 / 20922789888000.0;   // -1/16!
    pc = fma(pc, t2, 1.0 / 87178291200.0);
    pc = fma(pc, t2, -1.0 / 479001600.0);
    pc = fma(pc, t2, 1.0 / 3628800.0);
    pc = fma(pc, t2, -1.0 / 40320.0);
    pc = fma(pc, t2, 1.0 / 720.0);
    pc = fma(pc, t2, -1.0 / 24.0);
    pc = fma(pc, t2, 0.5);
    const double ct = fma(-t2, pc, 1.0);
    const int swap = ((o + 1) >> 1) & 1;
    const double sa = swap ? ct : st;
    const double ca = swap ? st : ct;
    *sn = (o >= 4) ? -sa : sa;
    *cs = (((o + 2) >> 2) & 1) ? -ca : ca;
}

// cos(2 pi w 2^-32) alone: quadrant q = top 2 bits, f = the low 30 bits 2^-30;
// cos(q pi/2 + pi f/2) = +-sin(pi g / 2) with g = f (q odd) or 1 - f (q even),
// negative for q = 1, 2; one sine series on [0, pi/2] (to t^21)
MJ_HD double mj_cos2pi_u32(uint32_t w)
{
    const int q = (int)(w >> 30);
    const double f = (double)(w & 0x3FFFFFFFu) * 9.31322574615478515625e-10;   // 2^-30
    const double g = (q & 1) ? f : (1.0 - f);
    const double t = g * 1.57079632679489661923;   // pi/2
    const double t2 = t * t;
    double p = -1.0 / 51090942171709440000.0;   // -1/21!
    p = fma(p, t2, 1.0 / 121645100408832000.0);
    p = fma(p, t2, -1.0 / 355687428096000.0);
    p = fma(p, t2, 1.0 / 1307674368000.0);
    p = fma(p, t2, -1.0 / 6227020800.0);
    p = fma(p, t2, 1.0 / 39916800.0);
    p = fma(p, t2, -1.0 / 362880.0);
    p = fma(p, t2, 1.0 / 5040.0);
    p = fma(p, t2, -1.0 / 120.0);
    p = fma(p, t2, 1.0 / 6.0);
    const double sv = fma(-(t * t2), p, t);
    return (q == 1 || q == 2) ? -sv : sv;
}
