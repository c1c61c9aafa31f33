// Device-side helpers for the LLG hot path on gfx950.
//
// All per-site vector algebra is written out in the reference's operation
// order and the library is compiled without floating-point contraction, so
// LLG / euler / langevin reproduce src/solvers/propagator.f90:13-40,
// src/solvers/solver.f90:157-178 and src/dynamics/BTeff_sd.f90:18-45 op for
// op.  The stencil sum (where the reference's own order depends on its matrix
// backend) uses explicit fma().
#pragma once
#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>

#include "mj_rngmath.h"
#endif
#include <stdint.h>

#define MJ_HBAR 6.58211928e-1  // eV fs (src/fixed-parameters/constant.f90:25)
#define MJ_RHBAR 0x1.84eeb74185935p+0  // RN(1/MJ_HBAR)

// a / b for a divisor b known with its correctly rounded reciprocal r = RN(1/b):
// q0 = RN(a r), e = a - b q0 (exact, fma), RN(q0 + e r) is the correctly rounded
// quotient (Markstein, IBM J. Res. Dev. 34 (1990) 111) outside the
// overflow/underflow range; three instructions instead of the general
// division sequence, bit-identical to a / b.
__device__ __forceinline__ double div_rn(double a, double b, double r)
{
    const double q0 = a * r;
    const double e = fma(-b, q0, a);
    return fma(e, r, q0);
}

struct v3 {
    double x, y, z;
};

__device__ __forceinline__ v3 mk3(double x, double y, double z)
{
    v3 r;
    r.x = x;
    r.y = y;
    r.z = z;
    return r;
}

// cross_NM (src/vector_utils/vector.f90:451-465)
__device__ __forceinline__ v3 cross3(v3 a, v3 b)
{
    return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// Correctly rounded sqrt(x) and 1 / x for the per-site algebra: the
// instruction sequences the compiler emits for sqrt and 1.0 / x in double
// (rsq / rcp seed, then fma corrections: two Newton steps and a final
// correction for 1 / x, a Goldschmidt step and two corrections for sqrt)
// without their range-scaling steps, which change nothing for x in
// [2^-767, 2^766] (sqrt) and [2^-1000, 2^1000] (1 / x): spin lengths, field
// norms and the Cayley denominator lie far inside.  sqrt(0) = 0 is kept.
// Same bits as the library operations there (checked on the device against
// sqrt and 1.0 / x, mj_math_probe); 8 and 3 fewer instructions per call.
__device__ __forceinline__ double mj_sqrt_rn(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    double d = fma(-g, g, x);
    h = fma(h, r, h);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    return x == 0.0 ? x : g;
}

__device__ __forceinline__ double mj_rcp_rn(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    const double f = fma(-x, r, 1.0);
    return fma(f, r, r);
}

__device__ __forceinline__ double dot3(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ double norm3(v3 a) { return mj_sqrt_rn((a.x * a.x + a.y * a.y) + a.z * a.z); }

// normalize with the 1e-8 guard (src/vector_utils/vector.f90:76-84); the three
// quotients m_k / |m| share one correctly rounded reciprocal (div_rn)
__device__ __forceinline__ v3 unit3(v3 m)
{
    double nr = norm3(m);
    if (nr > 1.0e-8) {
        const double r = mj_rcp_rn(nr);
        return mk3(div_rn(m.x, nr, r), div_rn(m.y, nr, r), div_rn(m.z, nr, r));
    }
    return m;
}

// D = -m^ x (B + a m^ x B)/(1 + a^2)   (propagator.f90:32-40, BTeff_sd.f90:38-43)
// den = 1 + a^2 and rden = RN(1/den)
__device__ __forceinline__ v3 llg_rhs(v3 m, v3 B, double alpha, double den, double rden)
{
    v3 mn = unit3(m);
    v3 c = cross3(mn, B);
    v3 l = mk3(B.x + alpha * c.x, B.y + alpha * c.y, B.z + alpha * c.z);
    v3 o = cross3(mn, l);
    return mk3(div_rn(-o.x, den, rden), div_rn(-o.y, den, rden), div_rn(-o.z, den, rden));
}

// the thermal increment DT*dt/hbar of euler (solver.f90:174), rounded as there
__device__ __forceinline__ v3 thermal_increment(v3 dtf, double dt)
{
    return mk3(div_rn(dtf.x * dt, MJ_HBAR, MJ_RHBAR), div_rn(dtf.y * dt, MJ_HBAR, MJ_RHBAR),
               div_rn(dtf.z * dt, MJ_HBAR, MJ_RHBAR));
}

// euler (solver.f90:172-176): t = (m + Dint*dt/hbar) + DT*dt/hbar; m' = t*|m|/|t|,
// with q = DT*dt/hbar given
__device__ __forceinline__ v3 euler_step_q(v3 m, v3 dint, v3 q, double dt)
{
    double mn = norm3(m);
    v3 t = mk3((m.x + div_rn(dint.x * dt, MJ_HBAR, MJ_RHBAR)) + q.x, (m.y + div_rn(dint.y * dt, MJ_HBAR, MJ_RHBAR)) + q.y,
               (m.z + div_rn(dint.z * dt, MJ_HBAR, MJ_RHBAR)) + q.z);
    double tn = norm3(t);
    if (tn > 1.0e-8) {
        const double r = mj_rcp_rn(tn);
        return mk3(div_rn(t.x * mn, tn, r), div_rn(t.y * mn, tn, r), div_rn(t.z * mn, tn, r));
    }
    return m;
}

__device__ __forceinline__ v3 euler_step(v3 m, v3 dint, v3 dtf, double dt)
{
    return euler_step_q(m, dint, thermal_increment(dtf, dt), dt);
}

// euler without a thermal term (the reference adds Dmag_T_zero*dt/hbar = +0)
__device__ __forceinline__ v3 euler_step0(v3 m, v3 dint, double dt)
{
    double mn = norm3(m);
    v3 t = mk3(m.x + div_rn(dint.x * dt, MJ_HBAR, MJ_RHBAR), m.y + div_rn(dint.y * dt, MJ_HBAR, MJ_RHBAR),
               m.z + div_rn(dint.z * dt, MJ_HBAR, MJ_RHBAR));
    double tn = norm3(t);
    if (tn > 1.0e-8) {
        const double r = mj_rcp_rn(tn);
        return mk3(div_rn(t.x * mn, tn, r), div_rn(t.y * mn, tn, r), div_rn(t.z * mn, tn, r));
    }
    return m;
}

// Implicit-midpoint (Cayley) solve (I + [a]x) x = m - a x m  (SIB, Mentink et al. 2010)
__device__ __forceinline__ v3 cayley(v3 m, v3 a)
{
    v3 axm = cross3(a, m);
    v3 y = mk3(m.x - axm.x, m.y - axm.y, m.z - axm.z);
    v3 axy = cross3(a, y);
    double ay = dot3(a, y);
    double aa = dot3(a, a);
    double inv = mj_rcp_rn(1.0 + aa);
    return mk3(((y.x - axy.x) + ay * a.x) * inv, ((y.y - axy.y) + ay * a.y) * inv,
               ((y.z - axy.z) + ay * a.z) * inv);
}

// a = (dt/2) Omega, Omega = -(B + alpha p x B)/((1+alpha^2) hbar)
__device__ __forceinline__ v3 sib_vec(v3 p, v3 B, double alpha, double dt)
{
    v3 c = cross3(p, B);
    double f = -0.5 * dt / ((1.0 + alpha * alpha) * MJ_HBAR);
    return mk3(f * (B.x + alpha * c.x), f * (B.y + alpha * c.y), f * (B.z + alpha * c.z));
}

// Depondt effective rotation field Bt = (B + alpha m^ x B)/(1+alpha^2)
__device__ __forceinline__ v3 dep_field(v3 m, v3 B, double alpha, double den, double rden)
{
    v3 mn = unit3(m);
    v3 c = cross3(mn, B);
    return mk3(div_rn(B.x + alpha * c.x, den, rden), div_rn(B.y + alpha * c.y, den, rden),
               div_rn(B.z + alpha * c.z, den, rden));
}

// rotate m about b by |b| dt / hbar (Rodrigues), dm/dt = (1/hbar) b x m
__device__ __forceinline__ v3 dep_rotate(v3 m, v3 b, double dt)
{
    double bn = norm3(b);
    if (!(bn > 0.0)) return m;
    const double rb = mj_rcp_rn(bn);
    v3 k = mk3(div_rn(b.x, bn, rb), div_rn(b.y, bn, rb), div_rn(b.z, bn, rb));
    double th = div_rn(bn * dt, MJ_HBAR, MJ_RHBAR);
    double st, ct;
    sincos(th, &st, &ct);
    v3 kxm = cross3(k, m);
    double km = dot3(k, m);
    double f = km * (1.0 - ct);
    return mk3((m.x * ct + kxm.x * st) + k.x * f, (m.y * ct + kxm.y * st) + k.y * f,
               (m.z * ct + kxm.z * st) + k.z * f);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller normals.
// Key = seed, counter = (site, step lo, step hi, stream).
// ---------------------------------------------------------------------------
__host__ __device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c[0] = n0;
        c[1] = n1;
        c[2] = n2;
        c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// three N(0,1) for (site, step) from one Philox block, counter (site, step, 0):
// Box-Muller on the 32-bit words, words (0, 1) -> (n0, n1), words (2, 3) -> n2
// (radius from u = (w + 1) 2^-32 in (0, 1], angle 2 pi w 2^-32 in [0, 2 pi);
// mj_rngmath.h's word forms)
__device__ __forceinline__ v3 normals3(uint64_t seed, uint64_t step, uint32_t site)
{
    uint32_t c[4] = {site, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double r1 = mj_sqrt_pos(-2.0 * mj_ln_u32(c[0]));
    const double r2 = mj_sqrt_pos(-2.0 * mj_ln_u32(c[2]));
    double s1, c1;
    mj_sincos2pi_u32(c[1], &s1, &c1);
    const double c2 = mj_cos2pi_u32(c[3]);
    return mk3(r1 * c1, r1 * s1, r2 * c2);
}
