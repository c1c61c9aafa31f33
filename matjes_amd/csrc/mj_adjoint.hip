// Differentiable SIB trajectory: forward with checkpoints and the hand-written
// reverse (adjoint) step on the device.
//
// Forward step (per site i; T = 0; Mentink et al. 2010 SIB as in the plain path):
//   B(x)_i = -(sum_d C_d x_{i+d}) - Z^T b_i,   C = sum_k theta_k C^(k)
//   a1 = f (B(m) + alpha m x B(m)),  mt = cay(m, a1),  p = (m + mt)/2
//   a2 = f (B(p) + alpha p x B(p)),  m' = cay(m, a2),   f = -dt / (2 (1+alpha^2) hbar)
// with cay(m, a) = (I + [a]x)^-1 (I - [a]x) m.  Reverse of one step, given
// g' = dL/dm':
//   d cay(m,a)/dm ^T g = cay(g, -a);  d cay/da ^T g = M(-a) g x (cay(m,a) + m),
//   M(-a) g = (g + a x g + (a.g) a)/(1+|a|^2);
//   a = f (B + alpha x x B):  gB = f (ga + alpha ga x x),  gx += f alpha (B x ga);
//   B(x) field: gx += -(C * gB) (the operator is symmetric: C_{-d}^T = C_d);
//   dL/dtheta_k += sum_i gB_i . (-(C^(k) x)_i);  dL/db_i += -Z gB_i.
// Kernels of a reverse step: R1 (stage-2 site math), R2 (transpose of the
// stage-2 field + stage-1 site math), R3 (transpose of the stage-1 field); R3
// of a step and R1 of the step before it run as one kernel.
// Checkpoints every `ckpt` steps; a segment is recomputed forward (storing m_n
// and p_n) before its reverse sweep.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <vector>

#include "mj_device.h"
#include "mj_internal.h"
#include "mj_stage_body.h"   // mj_block_sum

struct AdjArgs {
    int nx, ny, nz, px, py, pz, bx, by, bz, rx, ry, rz, sx, sy, sz, ntx, nty, nblocks;
    int64_t np;
    int nofs;
    const int *ofs;     // LDS offsets of the total stencil
    const int *ofs0;    // the same in the fused step's S0 layout (tile + 2 reach)
    const double *cf;   // 9 per offset (row-major C_d)
    int nparam;
    const int *pstart;  // unit stencils: entries [pstart[k], pstart[k+1]) of pofs/pcf
    const int *pofs;
    const double *pcf;
    double z[9];
    const double *bmap; // b_i (SoA)
    double alpha, f;
    // state
    const double *x;    // stencil state of the kernel (SoA)
    const double *x2;   // second stencil state (fused R3 + R1: the midpoint of the earlier step)
    const double *m;    // own-site state m_n
    const double *m1;   // own-site state m_{n+1}
    const double *gin;  // own-site input gradient
    const double *gacc_in, *gpd_in;
    double *out0, *out1, *out2;
    double *ghmap;
    double *ppart;      // [nblocks][nparam + 3]
};

__device__ __forceinline__ v3 v_add(v3 a, v3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ v3 v_scale(double s, v3 a) { return mk3(s * a.x, s * a.y, s * a.z); }

// (I - [a]x + a a^T)^T ... M(-a) g = (g + a x g + (a.g) a) / (1 + |a|^2)
__device__ __forceinline__ v3 minv_t(v3 a, v3 g)
{
    v3 axg = cross3(a, g);
    double ag = dot3(a, g);
    double inv = 1.0 / (1.0 + dot3(a, a));
    return mk3((g.x + axg.x + ag * a.x) * inv, (g.y + axg.y + ag * a.y) * inv, (g.z + axg.z + ag * a.z) * inv);
}

__device__ __forceinline__ void adj_tile(const AdjArgs &a, const double *src, double *L, int x0, int y0, int z0)
{
    const int T = a.sx * a.sy * a.sz;
    const int nthr = a.bx * a.by * a.bz;
    for (int t = threadIdx.x; t < T; t += nthr) {
        int lx = t % a.sx, r = t / a.sx, ly = r % a.sy, lz = r / a.sy;
        int gx = x0 + lx - a.rx, gy = y0 + ly - a.ry, gz = z0 + lz - a.rz;
        double vx = 0.0, vy = 0.0, vz = 0.0;
        bool ok = true;
        int g3[3] = {gx, gy, gz};
        const int n3[3] = {a.nx, a.ny, a.nz}, p3[3] = {a.px, a.py, a.pz};
        for (int d = 0; d < 3; ++d) {
            if (g3[d] < 0 || g3[d] >= n3[d]) {
                if (!p3[d]) ok = false;
                g3[d] = ((g3[d] % n3[d]) + n3[d]) % n3[d];
            }
        }
        if (ok) {
            int64_t g = (int64_t)g3[0] + (int64_t)a.nx * ((int64_t)g3[1] + (int64_t)a.ny * g3[2]);
            vx = src[g];
            vy = src[g + a.np];
            vz = src[g + 2 * a.np];
        }
        L[t] = vx;
        L[T + t] = vy;
        L[2 * T + t] = vz;
    }
}

__device__ __forceinline__ v3 apply_blocks(const double *L, int T, int c, int n, const int *ofs, const double *cf)
{
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int k = 0; k < n; ++k) {
        const int j = c + ofs[k];
        const double *q = cf + 9 * k;
        const double mx = L[j], my = L[T + j], mz = L[2 * T + j];
        ax = fma(q[0], mx, ax);
        ax = fma(q[1], my, ax);
        ax = fma(q[2], mz, ax);
        ay = fma(q[3], mx, ay);
        ay = fma(q[4], my, ay);
        ay = fma(q[5], mz, ay);
        az = fma(q[6], mx, az);
        az = fma(q[7], my, az);
        az = fma(q[8], mz, az);
    }
    return mk3(ax, ay, az);
}

struct Site {
    int x0, y0, z0, c;
    bool active;
    int64_t gi;
};

__device__ __forceinline__ Site adj_site(const AdjArgs &a)
{
    Site s;
    int bid = blockIdx.x;
    int tx = bid % a.ntx, tr = bid / a.ntx, ty = tr % a.nty, tz = tr / a.nty;
    s.x0 = tx * a.bx;
    s.y0 = ty * a.by;
    s.z0 = tz * a.bz;
    int lx = threadIdx.x % a.bx, ly = (threadIdx.x / a.bx) % a.by, lz = threadIdx.x / (a.bx * a.by);
    int gx = s.x0 + lx, gy = s.y0 + ly, gz = s.z0 + lz;
    s.active = gx < a.nx && gy < a.ny && gz < a.nz;
    s.gi = s.active ? (int64_t)gx + (int64_t)a.nx * ((int64_t)gy + (int64_t)a.ny * gz) : 0;
    s.c = (lz + a.rz) * a.sx * a.sy + (ly + a.ry) * a.sx + (lx + a.rx);
    return s;
}

__device__ __forceinline__ v3 ld3(const double *p, int64_t i, int64_t np) { return mk3(p[i], p[i + np], p[i + 2 * np]); }
__device__ __forceinline__ void st3(double *p, int64_t i, int64_t np, v3 v)
{
    p[i] = v.x;
    p[i + np] = v.y;
    p[i + 2 * np] = v.z;
}

__device__ __forceinline__ v3 zeeman_t(const AdjArgs &a, int64_t gi)   // Z^T b_i
{
    v3 b = ld3(a.bmap, gi, a.np);
    return mk3((b.x * a.z[0] + b.y * a.z[3]) + b.z * a.z[6], (b.x * a.z[1] + b.y * a.z[4]) + b.z * a.z[7],
               (b.x * a.z[2] + b.y * a.z[5]) + b.z * a.z[8]);
}

__device__ __forceinline__ v3 zeeman_vjp(const AdjArgs &a, v3 g)   // -Z g
{
    return mk3(-((a.z[0] * g.x + a.z[1] * g.y) + a.z[2] * g.z), -((a.z[3] * g.x + a.z[4] * g.y) + a.z[5] * g.z),
               -((a.z[6] * g.x + a.z[7] * g.y) + a.z[8] * g.z));
}

// per-block deterministic sums of nparam + 3 values, accumulated into
// ppart[block][*]: every value's wavefront butterflies, one barrier, then
// thread q adds value q's wave sums in wave order (mj_block_sum's order for
// each value, without a barrier pair per value)
__device__ __forceinline__ void adj_reduce(const AdjArgs &a, double *red, const double *vals, int nv)
{
    const int nthr = a.bx * a.by * a.bz;
    __syncthreads();
    if ((nthr & 63) != 0) {
        for (int q = 0; q < nv; ++q) {
            const double t = mj_block_sum(vals[q], red, nthr);
            if (threadIdx.x == 0) a.ppart[(size_t)blockIdx.x * nv + q] += t;
        }
        return;
    }
    const int nw = nthr >> 6;
    for (int q = 0; q < nv; ++q) {
        double v = vals[q];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0) red[q * nw + (threadIdx.x >> 6)] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < nv) {
        const int q = threadIdx.x;
        double r = red[q * nw];
        for (int w = 1; w < nw; ++w) r = r + red[q * nw + w];
        a.ppart[(size_t)blockIdx.x * nv + q] += r;
    }
}

#define MJ_ADJ_MAXP 13

// forward stage 1: p = (m + cay(m, a1(m)))/2
__global__ void k_adj_sib1(const AdjArgs a)
{
    extern __shared__ double L[];
    Site s = adj_site(a);
    adj_tile(a, a.x, L, s.x0, s.y0, s.z0);
    __syncthreads();
    if (!s.active) return;
    const int T = a.sx * a.sy * a.sz;
    v3 m = mk3(L[s.c], L[T + s.c], L[2 * T + s.c]);
    v3 acc = apply_blocks(L, T, s.c, a.nofs, a.ofs, a.cf);
    v3 zb = zeeman_t(a, s.gi);
    v3 B = mk3(-(acc.x + zb.x), -(acc.y + zb.y), -(acc.z + zb.z));
    v3 c = cross3(m, B);
    v3 av = v_scale(a.f, mk3(B.x + a.alpha * c.x, B.y + a.alpha * c.y, B.z + a.alpha * c.z));
    v3 mt = cayley(m, av);
    st3(a.out0, s.gi, a.np, mk3(0.5 * (m.x + mt.x), 0.5 * (m.y + mt.y), 0.5 * (m.z + mt.z)));
}

// forward stage 2: m' = cay(m, a2(p))
__global__ void k_adj_sib2(const AdjArgs a)
{
    extern __shared__ double L[];
    Site s = adj_site(a);
    adj_tile(a, a.x, L, s.x0, s.y0, s.z0);
    __syncthreads();
    if (!s.active) return;
    const int T = a.sx * a.sy * a.sz;
    v3 p = mk3(L[s.c], L[T + s.c], L[2 * T + s.c]);
    v3 acc = apply_blocks(L, T, s.c, a.nofs, a.ofs, a.cf);
    v3 zb = zeeman_t(a, s.gi);
    v3 B = mk3(-(acc.x + zb.x), -(acc.y + zb.y), -(acc.z + zb.z));
    v3 c = cross3(p, B);
    v3 av = v_scale(a.f, mk3(B.x + a.alpha * c.x, B.y + a.alpha * c.y, B.z + a.alpha * c.z));
    st3(a.out0, s.gi, a.np, cayley(ld3(a.m, s.gi, a.np), av));
}

// lattice site of a region position (wrapped on periodic axes; false off an open edge)
__device__ __forceinline__ bool adj_site_of(const AdjArgs &a, int gx, int gy, int gz, int64_t &g)
{
    int g3[3] = {gx, gy, gz};
    const int n3[3] = {a.nx, a.ny, a.nz}, p3[3] = {a.px, a.py, a.pz};
    for (int d = 0; d < 3; ++d) {
        if (g3[d] < 0 || g3[d] >= n3[d]) {
            if (!p3[d]) return false;
            g3[d] = ((g3[d] % n3[d]) + n3[d]) % n3[d];
        }
    }
    g = (int64_t)g3[0] + (int64_t)a.nx * ((int64_t)g3[1] + (int64_t)a.ny * g3[2]);
    return true;
}

// the midpoint p of SIB's first stage at one site from the S0 tile
__device__ __forceinline__ v3 adj_midpoint(const AdjArgs &a, const double *L0, int T0, int c0, int64_t gi, v3 &m)
{
    m = mk3(L0[c0], L0[T0 + c0], L0[2 * T0 + c0]);
    v3 acc = apply_blocks(L0, T0, c0, a.nofs, a.ofs0, a.cf);
    v3 zb = zeeman_t(a, gi);
    v3 B = mk3(-(acc.x + zb.x), -(acc.y + zb.y), -(acc.z + zb.z));
    v3 c = cross3(m, B);
    v3 av = v_scale(a.f, mk3(B.x + a.alpha * c.x, B.y + a.alpha * c.y, B.z + a.alpha * c.z));
    v3 mt = cayley(m, av);
    return mk3(0.5 * (m.x + mt.x), 0.5 * (m.y + mt.y), 0.5 * (m.z + mt.z));
}

// a whole forward step in one kernel: the tile plus twice the reach (S0) in
// LDS, the midpoint on the tile plus the reach (S1: the tile's own sites and
// a ring repeating the neighbouring tiles' first stage with the same
// operations), the second stage on the tile.  out0 = the midpoint p (kept for
// the reverse sweep), out1 = m'; bit for bit k_adj_sib1 then k_adj_sib2.
__global__ void k_adj_sibf(const AdjArgs a)
{
    extern __shared__ double L[];
    const int s0x = a.bx + 4 * a.rx, s0y = a.by + 4 * a.ry, s0z = a.bz + 4 * a.rz, T0 = s0x * s0y * s0z;
    const int T1 = a.sx * a.sy * a.sz;
    const int nthr = a.bx * a.by * a.bz;
    double *L0 = L;
    double *L1 = L + 3 * T0;
    Site s = adj_site(a);
    for (int t = threadIdx.x; t < T0; t += nthr) {
        const int lx = t % s0x, r = t / s0x, ly = r % s0y, lz = r / s0y;
        int64_t g = 0;
        double vx = 0.0, vy = 0.0, vz = 0.0;
        if (adj_site_of(a, s.x0 + lx - 2 * a.rx, s.y0 + ly - 2 * a.ry, s.z0 + lz - 2 * a.rz, g)) {
            vx = a.x[g];
            vy = a.x[g + a.np];
            vz = a.x[g + 2 * a.np];
        }
        L0[t] = vx;
        L0[T0 + t] = vy;
        L0[2 * T0 + t] = vz;
    }
    __syncthreads();
    const int lx = threadIdx.x % a.bx, ly = (threadIdx.x / a.bx) % a.by, lz = threadIdx.x / (a.bx * a.by);
    v3 m0 = mk3(0.0, 0.0, 0.0);
    {
        int64_t gw = 0;
        v3 p = mk3(0.0, 0.0, 0.0);
        if (adj_site_of(a, s.x0 + lx, s.y0 + ly, s.z0 + lz, gw)) {
            const int c0 = (lz + 2 * a.rz) * s0x * s0y + (ly + 2 * a.ry) * s0x + (lx + 2 * a.rx);
            p = adj_midpoint(a, L0, T0, c0, gw, m0);
        }
        L1[s.c] = p.x;
        L1[T1 + s.c] = p.y;
        L1[2 * T1 + s.c] = p.z;
    }
    // the ring of S1: z slabs, then y slabs of the tile's planes, then x slabs of its rows
    const int nzs = 2 * a.rz * a.sx * a.sy, nys = 2 * a.ry * a.bz * a.sx;
    for (int k = threadIdx.x; k < T1 - nthr; k += nthr) {
        int sx, sy, sz;
        if (k < nzs) {
            const int plane = k / (a.sx * a.sy), rem = k % (a.sx * a.sy);
            sz = plane < a.rz ? plane : plane + a.bz;
            sy = rem / a.sx;
            sx = rem % a.sx;
        } else if (k < nzs + nys) {
            const int q = k - nzs, row = q / a.sx, yy = row % (2 * a.ry);
            sz = a.rz + row / (2 * a.ry);
            sy = yy < a.ry ? yy : yy + a.by;
            sx = q % a.sx;
        } else {
            const int q = k - nzs - nys, col = q % (2 * a.rx), row = q / (2 * a.rx);
            sy = a.ry + row % a.by;
            sz = a.rz + row / a.by;
            sx = col < a.rx ? col : col + a.bx;
        }
        const int c1 = sz * a.sx * a.sy + sy * a.sx + sx;
        int64_t gw = 0;
        v3 p = mk3(0.0, 0.0, 0.0);
        if (adj_site_of(a, s.x0 - a.rx + sx, s.y0 - a.ry + sy, s.z0 - a.rz + sz, gw)) {
            const int c0 = (sz + a.rz) * s0x * s0y + (sy + a.ry) * s0x + (sx + a.rx);
            v3 mm;
            p = adj_midpoint(a, L0, T0, c0, gw, mm);
        }
        L1[c1] = p.x;
        L1[T1 + c1] = p.y;
        L1[2 * T1 + c1] = p.z;
    }
    __syncthreads();
    if (!s.active) return;
    v3 p = mk3(L1[s.c], L1[T1 + s.c], L1[2 * T1 + s.c]);
    v3 acc = apply_blocks(L1, T1, s.c, a.nofs, a.ofs, a.cf);
    v3 zb = zeeman_t(a, s.gi);
    v3 B = mk3(-(acc.x + zb.x), -(acc.y + zb.y), -(acc.z + zb.z));
    v3 c = cross3(p, B);
    v3 av = v_scale(a.f, mk3(B.x + a.alpha * c.x, B.y + a.alpha * c.y, B.z + a.alpha * c.z));
    st3(a.out0, s.gi, a.np, p);
    st3(a.out1, s.gi, a.np, cayley(m0, av));
}

// R1: reverse of the stage-2 site math.  x = p (stencil), m, m1 = m', gin = g'.
// out0 = gB2, out1 = gacc (cay(g', -a2)), out2 = gp_direct
// R1 site algebra: the midpoint tile p is in L, g = dL/dm' at the own site
__device__ __forceinline__ void r1_site(const AdjArgs &a, const double *L, int T, const Site &s, v3 g, double *vals)
{
    {
        v3 p = mk3(L[s.c], L[T + s.c], L[2 * T + s.c]);
        v3 acc = apply_blocks(L, T, s.c, a.nofs, a.ofs, a.cf);
        v3 zb = zeeman_t(a, s.gi);
        v3 B = mk3(-(acc.x + zb.x), -(acc.y + zb.y), -(acc.z + zb.z));
        v3 c = cross3(p, B);
        v3 av = v_scale(a.f, mk3(B.x + a.alpha * c.x, B.y + a.alpha * c.y, B.z + a.alpha * c.z));
        v3 m = ld3(a.m, s.gi, a.np), m1 = ld3(a.m1, s.gi, a.np);
        v3 gm = cayley(g, mk3(-av.x, -av.y, -av.z));
        v3 w = minv_t(av, g);
        v3 ga = cross3(w, v_add(m1, m));
        v3 gB = v_scale(a.f, v_add(ga, v_scale(a.alpha, cross3(ga, p))));
        v3 gp = v_scale(a.f * a.alpha, cross3(B, ga));
        st3(a.out0, s.gi, a.np, gB);
        st3(a.out1, s.gi, a.np, gm);
        st3(a.out2, s.gi, a.np, gp);
        for (int k = 0; k < a.nparam; ++k) {
            v3 u = apply_blocks(L, T, s.c, a.pstart[k + 1] - a.pstart[k], a.pofs + a.pstart[k], a.pcf + 9 * a.pstart[k]);
            vals[k] = -dot3(gB, u);
        }
        v3 gb = zeeman_vjp(a, gB);
        vals[a.nparam] = gb.x;
        vals[a.nparam + 1] = gb.y;
        vals[a.nparam + 2] = gb.z;
        st3(a.ghmap, s.gi, a.np, v_add(ld3(a.ghmap, s.gi, a.np), gb));
    }
}

__global__ void k_adj_r1(const AdjArgs a)
{
    extern __shared__ double L[];
    const int T = a.sx * a.sy * a.sz;
    double *red = L + 3 * T;
    Site s = adj_site(a);
    adj_tile(a, a.x, L, s.x0, s.y0, s.z0);
    __syncthreads();
    double vals[MJ_ADJ_MAXP + 3];
    for (int q = 0; q < a.nparam + 3; ++q) vals[q] = 0.0;
    if (s.active) r1_site(a, L, T, s, ld3(a.gin, s.gi, a.np), vals);
    adj_reduce(a, red, vals, a.nparam + 3);
}

// R3 of step q fused with R1 of step q-1: lambda_q = gacc - C * gB1 at the own
// site (tile of gB1 in x), then the stage-2 reverse of step q-1 with
// g = lambda_q (tile of its midpoint in x2); lambda_q never goes to memory
__global__ void k_adj_r3r1(const AdjArgs a)
{
    extern __shared__ double L[];
    const int T = a.sx * a.sy * a.sz;
    double *LP = L + 3 * T;
    double *red = L + 6 * T;
    Site s = adj_site(a);
    adj_tile(a, a.x, L, s.x0, s.y0, s.z0);
    adj_tile(a, a.x2, LP, s.x0, s.y0, s.z0);
    __syncthreads();
    double vals[MJ_ADJ_MAXP + 3];
    for (int q = 0; q < a.nparam + 3; ++q) vals[q] = 0.0;
    if (s.active) {
        const v3 t = apply_blocks(L, T, s.c, a.nofs, a.ofs, a.cf);
        const v3 g = ld3(a.gacc_in, s.gi, a.np);
        r1_site(a, LP, T, s, mk3(g.x - t.x, g.y - t.y, g.z - t.z), vals);
    }
    adj_reduce(a, red, vals, a.nparam + 3);
}

// R2: transpose of the stage-2 field (stencil over gB2 in x) + reverse of the
// stage-1 site math (needs the stencil of m: second tile).
// gin = gB2 is x; m stencil read from a.m; gacc_in, gpd_in own.
// out0 = gB1, out1 = gacc2
__global__ void k_adj_r2(const AdjArgs a)
{
    extern __shared__ double L[];
    const int T = a.sx * a.sy * a.sz;
    double *LM = L + 3 * T;
    double *red = L + 6 * T;
    Site s = adj_site(a);
    adj_tile(a, a.x, L, s.x0, s.y0, s.z0);
    adj_tile(a, a.m, LM, s.x0, s.y0, s.z0);
    __syncthreads();
    double vals[MJ_ADJ_MAXP + 3];
    for (int q = 0; q < a.nparam + 3; ++q) vals[q] = 0.0;
    if (s.active) {
        v3 tgb = apply_blocks(L, T, s.c, a.nofs, a.ofs, a.cf);
        v3 gp = mk3(a.gpd_in[s.gi] - tgb.x, a.gpd_in[s.gi + a.np] - tgb.y, a.gpd_in[s.gi + 2 * a.np] - tgb.z);
        v3 gmt = v_scale(0.5, gp);
        v3 gm = v_add(ld3(a.gacc_in, s.gi, a.np), gmt);
        v3 m = mk3(LM[s.c], LM[T + s.c], LM[2 * T + s.c]);
        v3 acc = apply_blocks(LM, T, s.c, a.nofs, a.ofs, a.cf);
        v3 zb = zeeman_t(a, s.gi);
        v3 B = mk3(-(acc.x + zb.x), -(acc.y + zb.y), -(acc.z + zb.z));
        v3 c = cross3(m, B);
        v3 av = v_scale(a.f, mk3(B.x + a.alpha * c.x, B.y + a.alpha * c.y, B.z + a.alpha * c.z));
        v3 mt = cayley(m, av);
        gm = v_add(gm, cayley(gmt, mk3(-av.x, -av.y, -av.z)));
        v3 w = minv_t(av, gmt);
        v3 ga = cross3(w, v_add(mt, m));
        v3 gB = v_scale(a.f, v_add(ga, v_scale(a.alpha, cross3(ga, m))));
        gm = v_add(gm, v_scale(a.f * a.alpha, cross3(B, ga)));
        st3(a.out0, s.gi, a.np, gB);
        st3(a.out1, s.gi, a.np, gm);
        for (int k = 0; k < a.nparam; ++k) {
            v3 u = apply_blocks(LM, T, s.c, a.pstart[k + 1] - a.pstart[k], a.pofs + a.pstart[k], a.pcf + 9 * a.pstart[k]);
            vals[k] = -dot3(gB, u);
        }
        v3 gb = zeeman_vjp(a, gB);
        vals[a.nparam] = gb.x;
        vals[a.nparam + 1] = gb.y;
        vals[a.nparam + 2] = gb.z;
        st3(a.ghmap, s.gi, a.np, v_add(ld3(a.ghmap, s.gi, a.np), gb));
    }
    adj_reduce(a, red, vals, a.nparam + 3);
}

// R3: lambda_n = gacc2 - C * gB1   (x = gB1)
__global__ void k_adj_r3(const AdjArgs a)
{
    extern __shared__ double L[];
    const int T = a.sx * a.sy * a.sz;
    Site s = adj_site(a);
    adj_tile(a, a.x, L, s.x0, s.y0, s.z0);
    __syncthreads();
    if (!s.active) return;
    v3 t = apply_blocks(L, T, s.c, a.nofs, a.ofs, a.cf);
    v3 g = ld3(a.gacc_in, s.gi, a.np);
    st3(a.out0, s.gi, a.np, mk3(g.x - t.x, g.y - t.y, g.z - t.z));
}

__global__ void k_adj_bmap(const double *bext, const double *h_aos, double *bmap, int64_t n, int64_t np)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; ++c) bmap[i + c * np] = bext[c] + (h_aos ? h_aos[3 * i + c] : 0.0);
}

__global__ void k_adj_aos_soa(const double *in, double *out, int64_t n, int64_t np, int to_soa)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < 3; ++c) {
        if (to_soa) out[i + c * np] = in[3 * i + c];
        else out[3 * i + c] = in[i + c * np];
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct AdjState {
    int nparam = 0;
    std::vector<std::vector<int>> pofs3;          // per parameter: offsets (3 each) of unit H blocks
    std::vector<std::vector<double>> pblk;        // per parameter: 9 per offset (H_MM(i, i+d) = A)
    // device
    int *d_ofs = nullptr, *d_pstart = nullptr, *d_pofs = nullptr;
    int *d_ofs0 = nullptr;           // the total stencil in the fused step's S0 layout (null: two kernels per step)
    size_t shm_fused = 0;
    double *d_cf = nullptr, *d_pcf = nullptr;
    double *d_bext = nullptr, *d_bmap = nullptr;
    double *ckpt = nullptr;          // ncp * 3np
    double *segm = nullptr, *segp = nullptr;   // (ckpt + 1) * 3np and ckpt * 3np
    double *g = nullptr, *gB = nullptr, *gacc = nullptr, *gpd = nullptr, *ghmap = nullptr, *ppart = nullptr;
    int nofs = 0, rx = 0, ry = 0, rz = 0, bx = 0, by = 0, bz = 0, nblocks = 0;
    int nsteps = 0, ckpt_every = 0, ncp = 0;
    double alpha = 0.0, dt = 0.0;
    std::vector<double> theta;
    int last_valid = 0;              // segm/segp hold the last segment of the forward pass
    // graphs of one full segment (ckpt_every steps): forward recompute, reverse sweep
    hipGraphExec_t g_fwd = nullptr, g_rev = nullptr;
};

static void adj_drop_graphs(AdjState *s)
{
    if (s->g_fwd) (void)hipGraphExecDestroy(s->g_fwd);
    if (s->g_rev) (void)hipGraphExecDestroy(s->g_rev);
    s->g_fwd = s->g_rev = nullptr;
}

#define ADJ(c) ((AdjState *)(c)->adj)

static void adj_free(AdjState *s)
{
    adj_drop_graphs(s);
    int *ip[] = {s->d_ofs, s->d_pstart, s->d_pofs, s->d_ofs0};
    for (int *p : ip)
        if (p) hipFree(p);
    double *dp[] = {s->d_cf, s->d_pcf, s->d_bext, s->d_bmap, s->ckpt, s->segm, s->segp, s->g, s->gB, s->gacc,
                    s->gpd, s->ghmap, s->ppart};
    for (double *p : dp)
        if (p) hipFree(p);
}

void mj_adj_destroy(mj_ctx *c)
{
    if (!c->adj) return;
    adj_free(ADJ(c));
    delete ADJ(c);
    c->adj = nullptr;
}

static inline int canon_d(int d, int n, int periodic)
{
    if (!periodic) return d;
    d %= n;
    if (d < 0) d += n;
    if (2 * d > n) d -= n;
    return d;
}

// symmetrised stencil C_d = A_d + A_{-d}^T of a list of H blocks
static void symmetrise(const mj_ctx *c, const std::vector<int> &o3, const std::vector<double> &blk,
                       std::vector<std::array<int, 3>> &offs, std::vector<std::array<double, 9>> &C)
{
    auto slot = [&](int dx, int dy, int dz) -> std::array<double, 9> & {
        std::array<int, 3> o = {dx, dy, dz};
        for (size_t k = 0; k < offs.size(); ++k)
            if (offs[k] == o) return C[k];
        offs.push_back(o);
        C.push_back({0, 0, 0, 0, 0, 0, 0, 0, 0});
        return C.back();
    };
    for (size_t k = 0; k < blk.size() / 9; ++k) {
        int d[3], md[3];
        for (int q = 0; q < 3; ++q) {
            d[q] = canon_d(o3[3 * k + q], c->dim[q], c->periodic[q]);
            md[q] = canon_d(-o3[3 * k + q], c->dim[q], c->periodic[q]);
        }
        auto &f = slot(d[0], d[1], d[2]);
        for (int q = 0; q < 9; ++q) f[q] += blk[9 * k + q];
        auto &g = slot(md[0], md[1], md[2]);
        for (int r = 0; r < 3; ++r)
            for (int q = 0; q < 3; ++q) g[3 * r + q] += blk[9 * k + 3 * q + r];
    }
}

extern "C" int mj_adjoint_setup(mj_ctx *c, int nparam, const int *nofs, const int *ofs3, const double *blocks9)
{
    if (!c || nparam < 0 || nparam > MJ_ADJ_MAXP || (nparam && (!nofs || !ofs3 || !blocks9))) {
        mj_set_error("mj_adjoint_setup: bad argument (at most 13 parameters)");
        return MJ_EINVAL;
    }
    if (!c->has_z) {
        mj_set_error("mj_adjoint_setup: register the Zeeman (B-M) term first");
        return MJ_ESTATE;
    }
    mj_adj_destroy(c);
    AdjState *s = new AdjState();
    c->adj = s;
    s->nparam = nparam;
    int pos = 0;
    for (int k = 0; k < nparam; ++k) {
        s->pofs3.emplace_back(ofs3 + 3 * pos, ofs3 + 3 * (pos + nofs[k]));
        s->pblk.emplace_back(blocks9 + 9 * pos, blocks9 + 9 * (pos + nofs[k]));
        pos += nofs[k];
    }
    return MJ_OK;
}

static void adj_args(const mj_ctx *c, const AdjState *s, AdjArgs &a)
{
    memset(&a, 0, sizeof a);
    a.nx = c->dim[0];
    a.ny = c->dim[1];
    a.nz = c->dim[2];
    a.px = c->periodic[0];
    a.py = c->periodic[1];
    a.pz = c->periodic[2];
    a.bx = s->bx;
    a.by = s->by;
    a.bz = s->bz;
    a.rx = s->rx;
    a.ry = s->ry;
    a.rz = s->rz;
    a.sx = a.bx + 2 * a.rx;
    a.sy = a.by + 2 * a.ry;
    a.sz = a.bz + 2 * a.rz;
    a.ntx = (a.nx + a.bx - 1) / a.bx;
    a.nty = (a.ny + a.by - 1) / a.by;
    a.nblocks = a.ntx * a.nty * ((a.nz + a.bz - 1) / a.bz);
    a.np = c->np;
    a.nofs = s->nofs;
    a.ofs = s->d_ofs;
    a.ofs0 = s->d_ofs0;
    a.cf = s->d_cf;
    a.nparam = s->nparam;
    a.pstart = s->d_pstart;
    a.pofs = s->d_pofs;
    a.pcf = s->d_pcf;
    for (int q = 0; q < 9; ++q) a.z[q] = c->zblk[q];
    a.bmap = s->d_bmap;
    a.alpha = s->alpha;
    a.f = -0.5 * s->dt / ((1.0 + s->alpha * s->alpha) * MJ_HBAR);
}

static int adj_launch(void (*k)(const AdjArgs), const AdjArgs &a, int ntiles, hipStream_t st)
{
    const int T = a.sx * a.sy * a.sz;
    const int nthr = a.bx * a.by * a.bz;
    size_t shm = sizeof(double) * ((size_t)3 * T * ntiles + nthr);
    if (shm > 65536) {
        mj_set_error("adjoint: the stencil reach needs more LDS than a workgroup has");
        return MJ_ENOTSUP;
    }
    hipLaunchKernelGGL(k, dim3(a.nblocks), dim3(nthr), shm, st, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("adjoint kernel launch: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    return MJ_OK;
}

#define ATRY(x)                     \
    do {                            \
        int r_ = (x);               \
        if (r_ != MJ_OK) return r_; \
    } while (0)

static int adj_alloc(double **p, size_t n)
{
    if (*p) hipFree(*p);
    *p = nullptr;
    if (hipMalloc(p, sizeof(double) * std::max<size_t>(n, 1)) != hipSuccess) {
        mj_set_error("adjoint: device allocation failed");
        return MJ_ENOMEM;
    }
    return MJ_OK;
}

// one forward SIB step from x_in to x_out (SoA), storing the midpoint in pbuf:
// one fused kernel when its tiles fit, else the two stage kernels
static int adj_step(mj_ctx *c, AdjState *s, const double *xin, double *pbuf, double *xout)
{
    AdjArgs a;
    adj_args(c, s, a);
    a.x = xin;
    a.out0 = pbuf;
    if (s->d_ofs0) {
        a.out1 = xout;
        hipLaunchKernelGGL(k_adj_sibf, dim3(a.nblocks), dim3(a.bx * a.by * a.bz), s->shm_fused, c->stream, a);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            mj_set_error(std::string("adjoint fused step launch: ") + hipGetErrorString(e));
            return MJ_EHIP;
        }
        return MJ_OK;
    }
    ATRY(adj_launch(k_adj_sib1, a, 1, c->stream));
    a.x = pbuf;
    a.m = xin;
    a.out0 = xout;
    return adj_launch(k_adj_sib2, a, 1, c->stream);
}

// `len` steps of one segment: segm[q] -> (segp[q], segm[q+1]); a full segment
// replays the captured graph (captured on first use after each forward set-up)
static int adj_segment_fwd(mj_ctx *c, AdjState *s, int len)
{
    const size_t n3 = 3 * (size_t)c->np;
    auto enqueue = [&](int n) -> int {
        for (int q = 0; q < n; ++q) ATRY(adj_step(c, s, s->segm + n3 * q, s->segp + n3 * q, s->segm + n3 * (q + 1)));
        return MJ_OK;
    };
    if (len != s->ckpt_every) return enqueue(len);
    if (!s->g_fwd) {
        hipGraph_t g = nullptr;
        MJ_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue(len);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != MJ_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (rc != MJ_OK) return rc;
            mj_set_error(std::string("adjoint forward capture: ") + hipGetErrorString(e));
            return MJ_EHIP;
        }
        e = hipGraphInstantiate(&s->g_fwd, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            s->g_fwd = nullptr;
            mj_set_error(std::string("adjoint forward graph: ") + hipGetErrorString(e));
            return MJ_EHIP;
        }
    }
    MJ_HIP_CHECK(hipGraphLaunch(s->g_fwd, c->stream));
    return MJ_OK;
}

extern "C" int mj_adjoint_forward(mj_ctx *c, const double *theta, const double *bext, const double *hmap_dev,
                                  const double *m0_dev, int nsteps, double timestep, double damping, int ckpt_every,
                                  double *mfinal_dev)
{
    if (!c || !c->adj || (!theta && ADJ(c)->nparam) || !bext || !m0_dev || !mfinal_dev || nsteps < 1 ||
        ckpt_every < 1 || !(timestep > 0.0) || !(damping >= 0.0)) {
        mj_set_error("mj_adjoint_forward: bad argument or mj_adjoint_setup not called");
        return MJ_EINVAL;
    }
    MJ_HIP_CHECK(hipSetDevice(c->device));
    AdjState *s = ADJ(c);
    s->theta.assign(theta, theta + s->nparam);
    s->nsteps = nsteps;
    s->ckpt_every = ckpt_every;
    s->alpha = damping;
    s->dt = timestep;
    // total and unit stencils (symmetrised), LDS offsets for the tile
    std::vector<std::array<int, 3>> offs;
    std::vector<std::array<double, 9>> C;
    std::vector<int> all3;
    std::vector<double> allb;
    for (int k = 0; k < s->nparam; ++k) {
        all3.insert(all3.end(), s->pofs3[k].begin(), s->pofs3[k].end());
        for (size_t q = 0; q < s->pblk[k].size(); ++q) allb.push_back(theta[k] * s->pblk[k][q]);
    }
    symmetrise(c, all3, allb, offs, C);
    std::vector<std::vector<std::array<int, 3>>> poffs(s->nparam);
    std::vector<std::vector<std::array<double, 9>>> pC(s->nparam);
    for (int k = 0; k < s->nparam; ++k) symmetrise(c, s->pofs3[k], s->pblk[k], poffs[k], pC[k]);
    s->rx = s->ry = s->rz = 0;
    for (auto &o : offs) {
        s->rx = std::max(s->rx, std::abs(o[0]));
        s->ry = std::max(s->ry, std::abs(o[1]));
        s->rz = std::max(s->rz, std::abs(o[2]));
    }
    // tile: up to 256 sites, smaller (down to one wavefront) while that gives
    // fewer than 512 workgroups, so that small lattices still spread over the CUs
    {
        const int px = 1 << (int)ceil(log2((double)std::max(c->dim[0], 1)));
        const int py = 1 << (int)ceil(log2((double)std::max(c->dim[1], 1)));
        int nthr = 256;
        while (nthr > 64 && (int64_t)c->n / nthr < 512) nthr >>= 1;
        s->bx = std::min(px, nthr / 8);   // 32x8, 16x8 or 8x8
        s->by = std::max(1, std::min(nthr / s->bx, py));
        s->bz = 1;
    }
    const int sx = s->bx + 2 * s->rx, sy = s->by + 2 * s->ry;
    auto lofs = [&](const std::array<int, 3> &o) { return o[2] * sx * sy + o[1] * sx + o[0]; };
    std::vector<int> hofs;
    std::vector<double> hcf;
    for (size_t k = 0; k < offs.size(); ++k) {
        hofs.push_back(lofs(offs[k]));
        hcf.insert(hcf.end(), C[k].begin(), C[k].end());
    }
    std::vector<int> hps(1, 0), hpofs;
    std::vector<double> hpcf;
    for (int k = 0; k < s->nparam; ++k) {
        for (size_t q = 0; q < poffs[k].size(); ++q) {
            hpofs.push_back(lofs(poffs[k][q]));
            hpcf.insert(hpcf.end(), pC[k][q].begin(), pC[k][q].end());
        }
        hps.push_back((int)hpofs.size());
    }
    s->nofs = (int)hofs.size();
    int *ip[4] = {s->d_ofs, s->d_pstart, s->d_pofs, s->d_ofs0};
    for (int *p : ip)
        if (p) hipFree(p);
    s->d_ofs0 = nullptr;
    // the fused forward step (MJ_ADJ_FUSE=0: two kernels per step; the same bits)
    {
        const int s0x = s->bx + 4 * s->rx, s0y = s->by + 4 * s->ry, s0z = s->bz + 4 * s->rz;
        const size_t shm = sizeof(double) * 3 * ((size_t)s0x * s0y * s0z + (size_t)sx * sy * (s->bz + 2 * s->rz));
        const char *fe = getenv("MJ_ADJ_FUSE");
        if (shm <= 65536 && !(fe && fe[0] == '0')) {
            std::vector<int> hofs0;
            for (auto &o : offs) hofs0.push_back(o[2] * s0x * s0y + o[1] * s0x + o[0]);
            MJ_HIP_CHECK(hipMalloc(&s->d_ofs0, sizeof(int) * std::max<size_t>(1, hofs0.size())));
            if (!hofs0.empty())
                MJ_HIP_CHECK(hipMemcpy(s->d_ofs0, hofs0.data(), sizeof(int) * hofs0.size(), hipMemcpyHostToDevice));
            s->shm_fused = shm;
        }
    }
    MJ_HIP_CHECK(hipMalloc(&s->d_ofs, sizeof(int) * std::max<size_t>(1, hofs.size())));
    MJ_HIP_CHECK(hipMalloc(&s->d_pstart, sizeof(int) * hps.size()));
    MJ_HIP_CHECK(hipMalloc(&s->d_pofs, sizeof(int) * std::max<size_t>(1, hpofs.size())));
    ATRY(adj_alloc(&s->d_cf, hcf.size()));
    ATRY(adj_alloc(&s->d_pcf, hpcf.size()));
    if (!hofs.empty()) {
        MJ_HIP_CHECK(hipMemcpy(s->d_ofs, hofs.data(), sizeof(int) * hofs.size(), hipMemcpyHostToDevice));
        MJ_HIP_CHECK(hipMemcpy(s->d_cf, hcf.data(), sizeof(double) * hcf.size(), hipMemcpyHostToDevice));
    }
    MJ_HIP_CHECK(hipMemcpy(s->d_pstart, hps.data(), sizeof(int) * hps.size(), hipMemcpyHostToDevice));
    if (!hpofs.empty()) {
        MJ_HIP_CHECK(hipMemcpy(s->d_pofs, hpofs.data(), sizeof(int) * hpofs.size(), hipMemcpyHostToDevice));
        MJ_HIP_CHECK(hipMemcpy(s->d_pcf, hpcf.data(), sizeof(double) * hpcf.size(), hipMemcpyHostToDevice));
    }
    // buffers
    const size_t n3 = 3 * (size_t)c->np;
    s->ncp = (nsteps + ckpt_every - 1) / ckpt_every;
    ATRY(adj_alloc(&s->d_bext, 3));
    ATRY(adj_alloc(&s->d_bmap, n3));
    ATRY(adj_alloc(&s->ckpt, n3 * s->ncp));
    ATRY(adj_alloc(&s->segm, n3 * (ckpt_every + 1)));
    ATRY(adj_alloc(&s->segp, n3 * ckpt_every));
    MJ_HIP_CHECK(hipMemcpy(s->d_bext, bext, sizeof(double) * 3, hipMemcpyHostToDevice));
    const unsigned nb = (unsigned)((c->n + 255) / 256);
    hipLaunchKernelGGL(k_adj_bmap, dim3(nb), dim3(256), 0, c->stream, s->d_bext, hmap_dev, s->d_bmap, c->n, c->np);
    // forward with checkpoints: segment by segment, states kept as in the recompute
    adj_drop_graphs(s);
    s->last_valid = 0;
    hipLaunchKernelGGL(k_adj_aos_soa, dim3(nb), dim3(256), 0, c->stream, m0_dev, s->segm, c->n, c->np, 1);
    int len = 0;
    for (int seg = 0; seg < s->ncp; ++seg) {
        len = std::min(ckpt_every, nsteps - seg * ckpt_every);
        if (seg > 0)
            MJ_HIP_CHECK(hipMemcpyAsync(s->segm, s->segm + n3 * ckpt_every, sizeof(double) * n3,
                                        hipMemcpyDeviceToDevice, c->stream));
        MJ_HIP_CHECK(hipMemcpyAsync(s->ckpt + n3 * seg, s->segm, sizeof(double) * n3, hipMemcpyDeviceToDevice,
                                    c->stream));
        ATRY(adj_segment_fwd(c, s, len));
    }
    s->last_valid = 1;
    hipLaunchKernelGGL(k_adj_aos_soa, dim3(nb), dim3(256), 0, c->stream, s->segm + n3 * len, mfinal_dev, c->n, c->np,
                       0);
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

// reverse sweep over one segment (q = len-1 .. 0): R1, R2, R3 per step
static int adj_segment_rev(mj_ctx *c, AdjState *s, int len)
{
    const size_t n3 = 3 * (size_t)c->np;
    AdjArgs a0;
    adj_args(c, s, a0);
    a0.ppart = s->ppart;
    a0.ghmap = s->ghmap;
    double *gB2 = s->gB, *gB1 = s->gB + n3;
    auto enqueue = [&](int n) -> int {
        {
            AdjArgs a = a0;          // R1 of the segment's last step: stage-2 site math
            a.x = s->segp + n3 * (n - 1);
            a.m = s->segm + n3 * (n - 1);
            a.m1 = s->segm + n3 * n;
            a.gin = s->g;
            a.out0 = gB2;
            a.out1 = s->gacc;
            a.out2 = s->gpd;
            ATRY(adj_launch(k_adj_r1, a, 1, c->stream));
        }
        for (int q = n - 1; q >= 0; --q) {
            AdjArgs b = a0;          // R2: transpose of the stage-2 field + stage-1 site math
            b.x = gB2;
            b.m = s->segm + n3 * q;
            b.gacc_in = s->gacc;
            b.gpd_in = s->gpd;
            b.out0 = gB1;
            b.out1 = s->gacc;        // own-site read-then-write
            ATRY(adj_launch(k_adj_r2, b, 2, c->stream));
            if (q > 0) {             // R3 of step q + R1 of step q-1
                AdjArgs f = a0;
                f.x = gB1;
                f.x2 = s->segp + n3 * (q - 1);
                f.gacc_in = s->gacc;
                f.m = s->segm + n3 * (q - 1);
                f.m1 = s->segm + n3 * q;
                f.out0 = gB2;
                f.out1 = s->gacc;    // own-site read-then-write
                f.out2 = s->gpd;
                ATRY(adj_launch(k_adj_r3r1, f, 2, c->stream));
            } else {                 // R3 of the segment's first step: lambda to memory
                AdjArgs r = a0;
                r.ppart = nullptr;
                r.ghmap = nullptr;
                r.x = gB1;
                r.gacc_in = s->gacc;
                r.out0 = s->g;
                ATRY(adj_launch(k_adj_r3, r, 1, c->stream));
            }
        }
        return MJ_OK;
    };
    if (len != s->ckpt_every) return enqueue(len);
    if (!s->g_rev) {
        hipGraph_t g = nullptr;
        MJ_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue(len);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc != MJ_OK || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            if (rc != MJ_OK) return rc;
            mj_set_error(std::string("adjoint reverse capture: ") + hipGetErrorString(e));
            return MJ_EHIP;
        }
        e = hipGraphInstantiate(&s->g_rev, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            s->g_rev = nullptr;
            mj_set_error(std::string("adjoint reverse graph: ") + hipGetErrorString(e));
            return MJ_EHIP;
        }
    }
    MJ_HIP_CHECK(hipGraphLaunch(s->g_rev, c->stream));
    return MJ_OK;
}

extern "C" int mj_adjoint_backward(mj_ctx *c, const double *gfinal_dev, double *gtheta, double *gbext,
                                   double *ghmap_dev, double *gm0_dev)
{
    if (!c || !c->adj || !gfinal_dev || (!gtheta && ADJ(c)->nparam) || !gbext) {
        mj_set_error("mj_adjoint_backward: bad argument or no forward pass");
        return MJ_EINVAL;
    }
    AdjState *s = ADJ(c);
    if (s->nsteps < 1 || !s->ckpt) {
        mj_set_error("mj_adjoint_backward: call mj_adjoint_forward first");
        return MJ_ESTATE;
    }
    MJ_HIP_CHECK(hipSetDevice(c->device));
    const size_t n3 = 3 * (size_t)c->np;
    AdjArgs a0;
    adj_args(c, s, a0);
    const int nv = s->nparam + 3;
    if (!s->g) {   // kept until the next mj_adjoint_setup (sizes fixed per context and parameter set)
        ATRY(adj_alloc(&s->g, n3));
        ATRY(adj_alloc(&s->gB, 2 * n3));
        ATRY(adj_alloc(&s->gacc, n3));
        ATRY(adj_alloc(&s->gpd, n3));
        ATRY(adj_alloc(&s->ghmap, n3));
        ATRY(adj_alloc(&s->ppart, (size_t)a0.nblocks * nv));
    }
    MJ_HIP_CHECK(hipMemsetAsync(s->ghmap, 0, sizeof(double) * n3, c->stream));
    MJ_HIP_CHECK(hipMemsetAsync(s->ppart, 0, sizeof(double) * a0.nblocks * nv, c->stream));
    const unsigned nb = (unsigned)((c->n + 255) / 256);
    hipLaunchKernelGGL(k_adj_aos_soa, dim3(nb), dim3(256), 0, c->stream, gfinal_dev, s->g, c->n, c->np, 1);
    for (int seg = s->ncp - 1; seg >= 0; --seg) {
        const int j0 = seg * s->ckpt_every;
        const int len = std::min(s->ckpt_every, s->nsteps - j0);
        if (!(seg == s->ncp - 1 && s->last_valid)) {   // the forward pass left the last segment in place
            MJ_HIP_CHECK(hipMemcpyAsync(s->segm, s->ckpt + n3 * seg, sizeof(double) * n3, hipMemcpyDeviceToDevice,
                                        c->stream));
            ATRY(adj_segment_fwd(c, s, len));
        }
        ATRY(adj_segment_rev(c, s, len));
    }
    s->last_valid = s->ncp == 1;
    std::vector<double> part((size_t)a0.nblocks * nv);
    MJ_HIP_CHECK(hipMemcpyAsync(part.data(), s->ppart, sizeof(double) * part.size(), hipMemcpyDeviceToHost, c->stream));
    if (ghmap_dev)
        hipLaunchKernelGGL(k_adj_aos_soa, dim3(nb), dim3(256), 0, c->stream, s->ghmap, ghmap_dev, c->n, c->np, 0);
    if (gm0_dev) hipLaunchKernelGGL(k_adj_aos_soa, dim3(nb), dim3(256), 0, c->stream, s->g, gm0_dev, c->n, c->np, 0);
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int q = 0; q < nv; ++q) {
        double t = 0.0;
        for (int b = 0; b < a0.nblocks; ++b) t += part[(size_t)b * nv + q];
        if (q < s->nparam) gtheta[q] = t;
        else gbext[q - s->nparam] = t;
    }
    return MJ_OK;
}
