// Metropolis Monte Carlo on the device: the reference's single-site update
// (Metropolis-MC/MCstep.f90:17-65) applied to all sites of one colour class
// at once.  Sites of one class share no term of the Hamiltonian, so their
// updates commute and a sweep is one pass per class.
//
//   single-site energy  hamiltonian%energy_single (hamiltonian_collection.f90:402-431):
//                       pair terms twice their row (mult_M_single = 2), on-site
//                       terms and Zeeman once;  here e(s) = s.(h + Z^T b) + s^T C_0 s / 2
//                       with h = sum_{d != 0} C_d m_{i+d}, C_d = A_d + A_{-d}^T
//   proposal            sphereft (paratemp-MC/sampling.f90:9-62): uniform in the
//                       cap of half-angle `cone` around the old spin, or the
//                       Ising flip (track_value.f90:83-91)
//   acceptance          accept (MCstep.f90:70-88): dE < 0, else u < exp(max(-dE/kT, -200)),
//                       evaluated as ln u < max(-dE/kT, -200)
//   cone adaptation     cone_update (track_value.f90:149-160) once per proposal,
//                       here for the N proposals of a whole sweep at the running
//                       acceptance rate of the sweeps before
//
// Colouring: class(x, y, z) = (x + a y + b z) mod p with the smallest p (then a,
// then b) for which no offset of the stencil joins two sites of one class and
// the classes stay consistent across periodic boundaries (p | nx, p | a ny,
// p | b nz where periodic).  Random numbers: Philox4x32-10 keyed by the seed,
// counter (site, sweep, 'MC'), one block per site and sweep: u1, u2 for the
// proposal, u3 for the acceptance, 32-bit resolution.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <vector>

#include "mj_device.h"
#include "mj_internal.h"
#include "mj_stage_body.h"

#define MJ_FAIL(code, msg)        \
    do {                          \
        mj_set_error(msg);        \
        return (code);            \
    } while (0)
#define MJ_TRY(x)                   \
    do {                            \
        int r_ = (x);               \
        if (r_ != MJ_OK) return r_; \
    } while (0)

namespace {

constexpr uint32_t kMcStream = 0x4D430000u;   // counter word 3 of the Monte Carlo draws
constexpr int kMcThreads = 256;

struct McState {
    double cone, cos_cone, acc, nb, rate;
};

struct McArgs {
    double *m;              // state, SoA
    int64_t np;
    int nx, ny, nz, px, py, pz;
    int p, ca, cb, color;
    int64_t nthreads;
    int n_iso, n_dm, n_gen;
    const int *o3;          // offsets (3 each), classes iso | dm | gen, no d = 0
    const double *cf;       // iso: s; dm: s, w; gen: 9
    double c0[9];           // on-site C_0 = A_0 + A_0^T
    int bmode;
    double zb[3], z[9];
    const double *bmap;
    double kT;
    uint64_t seed, sweep;
    int sampler;
    McState *st;
    double *acc_part;
    // replicas (parallel tempering): blockIdx.y is the replica; its state, counters,
    // partials and temperature are at these strides (kT_rep null: one temperature)
    int64_t rep_cm, rep_part;
    const double *kT_rep;
    // colour-major copy of the state used by the passes: slot of site (x, y, z) =
    // class * R * W + (y + ny z) * W + x / p, R = ny nz rows, W = ceil(nx / p)
    double *cm;
    const double *cbmap;    // per-site field map in the same order (or null)
    int64_t cs;             // component stride of cm (p R W)
    int W;
    const int *csh;         // per stencil entry: class shift (dx + a dy + b dz) mod p
    const double *cf9;      // every stencil entry as a full 3x3 block (passes)
    // rows at least the stencil reach from the y and z boundaries take the
    // neighbour slots from a table: slot + tab[x0 * ne + e] (+-W across x)
    int fast, ne;
    const int *tab;
    int ylo, yhi, zlo, zhi;
    // single-site energies
    int64_t nsites;
    const int64_t *sites;
    const double *spins;    // AoS 3 per site, or null for the current spin
    double *e_out;
};

__device__ __forceinline__ bool mc_nb(const McArgs &a, int x, int y, int z, const int *o, int64_t &j)
{
    int q[3] = {x + o[0], y + o[1], z + o[2]};
    const int n[3] = {a.nx, a.ny, a.nz};
    const int per[3] = {a.px, a.py, a.pz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (q[d] < 0 || q[d] >= n[d]) {
            if (!per[d]) return false;
            q[d] = ((q[d] % n[d]) + n[d]) % n[d];
        }
    }
    j = (int64_t)q[0] + (int64_t)a.nx * ((int64_t)q[1] + (int64_t)a.ny * q[2]);
    return true;
}

// colour-major slot of the neighbour at offset o of the site (x0 + jj p, y, z) of class k
__device__ __forceinline__ bool mc_nb_cm(const McArgs &a, int k, int x0, int jj, int y, int z, const int *o, int kd,
                                         int64_t &slot)
{
    int u = x0 + o[0], q = 0;
    while (u < 0) {
        u += a.p;
        --q;
    }
    while (u >= a.p) {
        u -= a.p;
        ++q;
    }
    int j2 = jj + q;
    const int x2 = u + j2 * a.p;
    if (x2 < 0 || x2 >= a.nx) {
        if (!a.px) return false;
        j2 += (x2 < 0 ? 1 : -1) * (a.nx / a.p);
    }
    int y2 = y + o[1], z2 = z + o[2];
    if (y2 < 0 || y2 >= a.ny) {
        if (!a.py) return false;
        y2 = ((y2 % a.ny) + a.ny) % a.ny;
    }
    if (z2 < 0 || z2 >= a.nz) {
        if (!a.pz) return false;
        z2 = ((z2 % a.nz) + a.nz) % a.nz;
    }
    int k2 = k + kd;
    if (k2 >= a.p) k2 -= a.p;
    slot = ((int64_t)k2 * a.ny * a.nz + (int64_t)y2 + (int64_t)a.ny * z2) * a.W + j2;
    return true;
}

// neighbour index in either layout: site order (energy_single) or colour-major (passes)
template <bool CM>
struct McSite {
    int x, y, z, k, x0, jj;
    __device__ __forceinline__ bool nb(const McArgs &a, int e, int64_t &j) const
    {
        if (CM) return mc_nb_cm(a, k, x0, jj, y, z, a.o3 + 3 * e, a.csh[e], j);
        return mc_nb(a, x, y, z, a.o3 + 3 * e, j);
    }
};

// h = sum_{d != 0} C_d m_{i+d}, same term order and operations as the field kernels
template <bool CM>
__device__ __forceinline__ v3 mc_pair_field(const McArgs &a, const McSite<CM> &site)
{
    const double *m = CM ? a.cm : a.m;
    const int64_t np = CM ? a.cs : a.np;
    double ax = 0.0, ay = 0.0, az = 0.0;
    int64_t j;
    for (int k = 0; k < a.n_iso; ++k) {
        if (!site.nb(a, k, j)) continue;
        const double s = a.cf[k];
        ax = fma(s, m[j], ax);
        ay = fma(s, m[j + np], ay);
        az = fma(s, m[j + 2 * np], az);
    }
    const double *cd = a.cf + a.n_iso;
    for (int k = 0; k < a.n_dm; ++k) {
        if (!site.nb(a, a.n_iso + k, j)) continue;
        const double s = cd[4 * k], wx = cd[4 * k + 1], wy = cd[4 * k + 2], wz = cd[4 * k + 3];
        const double mx = m[j], my = m[j + np], mz = m[j + 2 * np];
        ax = fma(s, mx, ax);
        ax = fma(wy, mz, ax);
        ax = fma(-wz, my, ax);
        ay = fma(s, my, ay);
        ay = fma(wz, mx, ay);
        ay = fma(-wx, mz, ay);
        az = fma(s, mz, az);
        az = fma(wx, my, az);
        az = fma(-wy, mx, az);
    }
    const double *cg = cd + 4 * a.n_dm;
    for (int k = 0; k < a.n_gen; ++k) {
        if (!site.nb(a, a.n_iso + a.n_dm + k, j)) continue;
        const double *q = cg + 9 * k;
        const double mx = m[j], my = m[j + np], mz = m[j + 2 * np];
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

// pair field in the passes: the neighbour slots of 8 stencil entries at a time
// are formed first so that their 24 loads are in flight together
constexpr int kMcChunk = 8;
__device__ __forceinline__ v3 mc_pair_field_cm(const McArgs &a, const double *cm, int k, int x0, int jj, int y, int z,
                                                int64_t own)
{
    const int n = a.n_iso + a.n_dm + a.n_gen;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int e0 = 0; e0 < n; e0 += kMcChunk) {
        int64_t sl[kMcChunk];
        bool ok[kMcChunk];
#pragma unroll
        for (int q = 0; q < kMcChunk; ++q) {
            const int e = e0 + q;
            ok[q] = e < n && mc_nb_cm(a, k, x0, jj, y, z, a.o3 + 3 * e, a.csh[e], sl[q]);
            if (!ok[q]) sl[q] = own;
        }
        double vx[kMcChunk], vy[kMcChunk], vz[kMcChunk];
#pragma unroll
        for (int q = 0; q < kMcChunk; ++q) {
            vx[q] = cm[sl[q]];
            vy[q] = cm[sl[q] + a.cs];
            vz[q] = cm[sl[q] + 2 * a.cs];
        }
#pragma unroll
        for (int q = 0; q < kMcChunk; ++q) {
            if (!ok[q]) continue;
            const double *c9 = a.cf9 + 9 * (e0 + q);
            ax = fma(c9[0], vx[q], ax);
            ax = fma(c9[1], vy[q], ax);
            ax = fma(c9[2], vz[q], ax);
            ay = fma(c9[3], vx[q], ay);
            ay = fma(c9[4], vy[q], ay);
            ay = fma(c9[5], vz[q], ay);
            az = fma(c9[6], vx[q], az);
            az = fma(c9[7], vy[q], az);
            az = fma(c9[8], vz[q], az);
        }
    }
    return mk3(ax, ay, az);
}

__device__ __forceinline__ v3 mc_zeeman(const McArgs &a, const double *bmap, int64_t stride, int64_t i)
{
    if (a.bmode == 1) return mk3(a.zb[0], a.zb[1], a.zb[2]);
    if (a.bmode == 2) {
        const double bx = bmap[i], by = bmap[i + stride], bz = bmap[i + 2 * stride];
        return mk3((bx * a.z[0] + by * a.z[3]) + bz * a.z[6], (bx * a.z[1] + by * a.z[4]) + bz * a.z[7],
                   (bx * a.z[2] + by * a.z[5]) + bz * a.z[8]);
    }
    return mk3(0.0, 0.0, 0.0);
}

// energy of site i with spin s: s.(h + z) + s^T C_0 s / 2
__device__ __forceinline__ double mc_site_energy(const McArgs &a, v3 s, v3 h)
{
    const double *c = a.c0;
    const double qx = (c[0] * s.x + c[1] * s.y) + c[2] * s.z;
    const double qy = (c[3] * s.x + c[4] * s.y) + c[5] * s.z;
    const double qz = (c[6] * s.x + c[7] * s.y) + c[8] * s.z;
    return dot3(s, h) + 0.5 * dot3(s, mk3(qx, qy, qz));
}

// sphereft: uniform in the cap of half-angle `cone` around s (cc = cos cone)
__device__ __forceinline__ v3 mc_sphere(v3 s, double cc, double u1, double u2)
{
    const double ct = u1 + cc * (1.0 - u1);
    const double st = sqrt(1.0 - ct * ct);
    double sp, cp;
    mj_sincos2pi(u2, &sp, &cp);
    const double cti = s.z;
    const double rr = s.x * s.x + s.y * s.y;
    const double sti = sqrt(rr);
    double cpi = 1.0, spi = 0.0;
    if (!(rr < 0.00001)) {
        cpi = s.x / sqrt(rr);
        spi = s.y / sqrt(rr);
    }
    const double t = st * sp * cti + ct * sti;
    const double ny = st * cp * cpi + t * spi;
    const double nx = -(st * cp * spi) + t * cpi;
    const double nz = -(st * sp * sti) + ct * cti;
    const double ss = sqrt((nx * nx + ny * ny) + nz * nz);
    return mk3(nx / ss, ny / ss, nz / ss);
}

// acceptance counters after a sweep (s accepted of n_done), and the cone for the
// next sweep's n_next proposals (cone_update n_next times at the running rate)
__device__ __forceinline__ void mc_fold_state(McState *st, double s, double n_done, double n_next, int adapt)
{
    const double acc = st->acc + s, nb = st->nb + n_done;
    st->acc = acc;
    st->nb = nb;
    if (nb > 0.0) st->rate = acc / nb;
    if (adapt && n_next > 0.0) {
        const double pi = 3.14159265358979323846264338327950288;   // constant.f90:32
        const double rate = st->rate;
        double cone = st->cone;
        if (rate > 0.5 && cone < pi) {
            const double k = fmin(ceil((pi - cone) / 1.0e-4), n_next);
            cone = cone + k * 1.0e-4;
        } else if (rate < 0.5 && cone > 0.01) {
            const double k = fmin(ceil((cone - 0.01) / 1.0e-4), n_next);
            cone = cone - k * 1.0e-4;
        }
        st->cone = cone;
        st->cos_cone = cos(cone);
    }
}

constexpr int kMcTab = 2048;   // neighbour-slot table entries held in LDS

// pair field of an interior row: neighbour slots from the table, x wrapped by +-W
__device__ __forceinline__ v3 mc_pair_field_tab(const McArgs &a, const double *cm, const int *tab, int x0, int x,
                                                 int64_t own)
{
    const int n = a.ne;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int e0 = 0; e0 < n; e0 += kMcChunk) {
        int64_t sl[kMcChunk];
        bool ok[kMcChunk];
#pragma unroll
        for (int q = 0; q < kMcChunk; ++q) {
            const int e = e0 + q;
            ok[q] = e < n;
            sl[q] = own;
            if (ok[q]) {
                const int xq = x + a.o3[3 * e];
                int64_t d = tab[x0 * n + e];
                if (xq < 0 || xq >= a.nx) {
                    ok[q] = a.px != 0;
                    d += xq < 0 ? a.W : -a.W;
                }
                if (ok[q]) sl[q] = own + d;
            }
        }
        double vx[kMcChunk], vy[kMcChunk], vz[kMcChunk];
#pragma unroll
        for (int q = 0; q < kMcChunk; ++q) {
            vx[q] = cm[sl[q]];
            vy[q] = cm[sl[q] + a.cs];
            vz[q] = cm[sl[q] + 2 * a.cs];
        }
#pragma unroll
        for (int q = 0; q < kMcChunk; ++q) {
            if (!ok[q]) continue;
            const double *c9 = a.cf9 + 9 * (e0 + q);
            ax = fma(c9[0], vx[q], ax);
            ax = fma(c9[1], vy[q], ax);
            ax = fma(c9[2], vz[q], ax);
            ay = fma(c9[3], vx[q], ay);
            ay = fma(c9[4], vy[q], ay);
            ay = fma(c9[5], vz[q], ay);
            az = fma(c9[6], vx[q], az);
            az = fma(c9[7], vy[q], az);
            az = fma(c9[8], vz[q], az);
        }
    }
    return mk3(ax, ay, az);
}

// one colour class
__global__ void __launch_bounds__(kMcThreads) k_mc_pass(const McArgs a)
{
    __shared__ double red[kMcThreads];
    __shared__ int tab[kMcTab];
    if (a.fast) {
        for (int q = threadIdx.x; q < a.p * a.ne; q += blockDim.x) tab[q] = a.tab[q];
        __syncthreads();
    }
    const int rep = blockIdx.y;
    double *cm = a.cm + rep * a.rep_cm;
    const McState *st = a.st + rep;
    const double kT = a.kT_rep ? a.kT_rep[rep] : a.kT;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double accepted = 0.0;
    if (t < a.nthreads) {
        const int64_t row = t / a.W;
        const int jj = (int)(t - row * a.W);
        const int y = (int)(row % a.ny), z = (int)(row / a.ny);
        int x0 = (a.color - a.ca * y - a.cb * z) % a.p;
        if (x0 < 0) x0 += a.p;
        const int x = x0 + jj * a.p;
        if (x < a.nx && z < a.nz) {
            const int64_t i = (int64_t)x + (int64_t)a.nx * row;
            const int64_t slot = (int64_t)a.color * a.ny * a.nz * a.W + t;
            const v3 s = mk3(cm[slot], cm[slot + a.cs], cm[slot + 2 * a.cs]);
            uint32_t c[4] = {(uint32_t)i, (uint32_t)a.sweep, (uint32_t)(a.sweep >> 32), kMcStream + (uint32_t)rep};
            philox4x32_10(c, (uint32_t)a.seed, (uint32_t)(a.seed >> 32));
            const double two_m32 = 2.3283064365386962890625e-10;
            v3 sn;
            if (a.sampler == MJ_MC_ISING)
                sn = mk3(-s.x, -s.y, -s.z);
            else
                sn = mc_sphere(s, st->cos_cone, (double)c[0] * two_m32, (double)c[1] * two_m32);
            const bool interior = a.fast && y >= a.ylo && y < a.ny - a.yhi && z >= a.zlo && z < a.nz - a.zhi;
            const v3 pf = interior ? mc_pair_field_tab(a, cm, tab, x0, x, slot)
                                   : mc_pair_field_cm(a, cm, a.color, x0, jj, y, z, slot);
            const v3 zf = mc_zeeman(a, a.cbmap, a.cs, slot);
            const v3 h = mk3(pf.x + zf.x, pf.y + zf.y, pf.z + zf.z);
            const double e_old = mc_site_energy(a, s, h);
            const double e_new = mc_site_energy(a, sn, h);
            const double de = e_new - e_old;
            bool acc = de < 0.0;
            if (!acc && kT > 0.0) {
                const double tmp = fmax(-de / kT, -200.0);
                acc = c[2] == 0u || mj_log01((double)c[2] * two_m32) < tmp;
            }
            if (acc) {
                cm[slot] = sn.x;
                cm[slot + a.cs] = sn.y;
                cm[slot + 2 * a.cs] = sn.z;
                accepted = 1.0;
            }
        }
    }
    const double s = mj_block_sum(accepted, red, blockDim.x);
    if (threadIdx.x == 0) a.acc_part[rep * a.rep_part + blockIdx.x] = s;
}

// acceptance counts of a sweep (the partials of its passes) and the cone of the next sweep
// (one block per replica: state st[blockIdx.x], partials part + blockIdx.x * rep_part)
__global__ void __launch_bounds__(kMcThreads) k_mc_fold(McState *st, const double *part, int nparts, int64_t rep_part,
                                                        double n_done, double n_next, int adapt)
{
    __shared__ double red[kMcThreads];
    const double *pp = part + blockIdx.x * rep_part;
    double v = 0.0;
    for (int b = threadIdx.x; b < nparts; b += blockDim.x) v += pp[b];
    const double tot = mj_block_sum(v, red, blockDim.x);
    if (threadIdx.x == 0) mc_fold_state(st + blockIdx.x, tot, n_done, n_next, adapt);
}

// site order <-> colour-major order of a 3-component SoA field
__global__ void __launch_bounds__(kMcThreads) k_mc_layout(const McArgs a, const double *src, int64_t sstride,
                                                          double *dst, int64_t dstride, int to_cm)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = (int64_t)a.nx * a.ny * a.nz;
    if (i >= n) return;
    const int x = (int)(i % a.nx);
    const int64_t r = i / a.nx;
    const int y = (int)(r % a.ny), z = (int)(r / a.ny);
    const int k = (int)(((int64_t)x + (int64_t)a.ca * y + (int64_t)a.cb * z) % a.p);
    const int64_t slot = ((int64_t)k * a.ny * a.nz + r) * a.W + x / a.p;
    const int64_t is = to_cm ? i : slot, id = to_cm ? slot : i;
    for (int c = 0; c < 3; ++c) dst[id + c * dstride] = src[is + c * sstride];
}

__global__ void __launch_bounds__(kMcThreads) k_mc_energy_single(const McArgs a)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nsites) return;
    const int64_t i = a.sites[t];
    const int x = (int)(i % a.nx);
    const int64_t r = i / a.nx;
    const int y = (int)(r % a.ny), z = (int)(r / a.ny);
    v3 s = mk3(a.m[i], a.m[i + a.np], a.m[i + 2 * a.np]);
    if (a.spins) s = mk3(a.spins[3 * t], a.spins[3 * t + 1], a.spins[3 * t + 2]);
    const McSite<false> site = {x, y, z, 0, 0, 0};
    const v3 pf = mc_pair_field<false>(a, site);
    const v3 zf = mc_zeeman(a, a.bmap, a.np, i);
    a.e_out[t] = mc_site_energy(a, s, mk3(pf.x + zf.x, pf.y + zf.y, pf.z + zf.z));
}

// energy and summed spins of every replica's colour-major state: per-block
// partials (E, Mx, My, Mz) at part + (replica * nblk + block) * 4
//   E = sum_i [ m_i.h_i / 2 + m_i^T C_0 m_i / 2 + m_i.(Z^T b_i) ],  h_i the pair field
__global__ void __launch_bounds__(kMcThreads) k_mc_rep_energy(const McArgs a, double *part)
{
    __shared__ double red[kMcThreads];
    const int rep = blockIdx.y;
    const double *cm = a.cm + rep * a.rep_cm;
    const int64_t RW = (int64_t)a.ny * a.nz * a.W;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (t < a.p * RW) {
        const int k = (int)(t / RW);
        const int64_t rest = t - k * RW;
        const int64_t row = rest / a.W;
        const int jj = (int)(rest - row * a.W);
        const int y = (int)(row % a.ny), z = (int)(row / a.ny);
        int x0 = (k - a.ca * y - a.cb * z) % a.p;
        if (x0 < 0) x0 += a.p;
        const int x = x0 + jj * a.p;
        if (x < a.nx) {
            const v3 m = mk3(cm[t], cm[t + a.cs], cm[t + 2 * a.cs]);
            const bool interior = a.fast && y >= a.ylo && y < a.ny - a.yhi && z >= a.zlo && z < a.nz - a.zhi;
            const v3 pf = interior ? mc_pair_field_tab(a, cm, a.tab + (size_t)k * a.p * a.ne, x0, x, t)
                                   : mc_pair_field_cm(a, cm, k, x0, jj, y, z, t);
            const v3 zf = mc_zeeman(a, a.cbmap, a.cs, t);
            v[0] = mc_site_energy(a, m, mk3(0.5 * pf.x + zf.x, 0.5 * pf.y + zf.y, 0.5 * pf.z + zf.z));
            v[1] = m.x;
            v[2] = m.y;
            v[3] = m.z;
        }
    }
    const int64_t base = ((int64_t)rep * gridDim.x + blockIdx.x) * 4;
    for (int q = 0; q < 4; ++q) {
        const double sq = mj_block_sum(v[q], red, blockDim.x);
        if (threadIdx.x == 0) part[base + q] = sq;
    }
}

__global__ void __launch_bounds__(kMcThreads) k_mc_rep_reduce(const double *part, int nblk, double *out)
{
    __shared__ double red[kMcThreads];
    const double *pp = part + (int64_t)blockIdx.x * nblk * 4;
    for (int q = 0; q < 4; ++q) {
        double v = 0.0;
        for (int b = threadIdx.x; b < nblk; b += blockDim.x) v += pp[4 * b + q];
        const double sq = mj_block_sum(v, red, blockDim.x);
        if (threadIdx.x == 0) out[4 * blockIdx.x + q] = sq;
    }
}

// reorder_temperature (paratemp-MC/paratemp.f90:47-103) with accept (:104-121):
// neighbouring rungs (q, q+1), q from (n_ex + 1) mod 2 in steps of 2, swap their
// replicas when delta = (1/kT_{q+1} - 1/kT_q)(E_{q+1} - E_q) > 0 or
// u < exp(max(delta, -200)); u from Philox (key seed, counter (q, n_ex, 'PT')).
__global__ void k_mc_exchange(const double *E4, const double *ladder, int *image, double *kT_rep, double *swaps,
                              int nrep, uint64_t seed, uint64_t n_ex)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    const double two_m32 = 2.3283064365386962890625e-10;
    for (int q = (int)((n_ex + 1) % 2); q + 1 < nrep; q += 2) {
        const int r1 = image[q], r2 = image[q + 1];
        const double kt1 = ladder[q], kt2 = ladder[q + 1];
        const double delta = (1.0 / kt2 - 1.0 / kt1) * (E4[4 * r2] - E4[4 * r1]);
        bool acc = delta > 0.0;
        if (!acc) {
            uint32_t c[4] = {(uint32_t)q, (uint32_t)n_ex, (uint32_t)(n_ex >> 32), 0x50540000u};
            philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
            acc = exp(fmax(delta, -200.0)) > (double)c[0] * two_m32;
        }
        swaps[2 * q + 1] += 1.0;
        if (acc) {
            image[q] = r2;
            image[q + 1] = r1;
            kT_rep[r1] = kt2;
            kT_rep[r2] = kt1;
            swaps[2 * q] += 1.0;
        }
    }
}

// samples by rung: E and M of the replica at each rung
__global__ void k_mc_rep_samples(const double *E4, const int *image, int nrep, double *e_out, double *m_out)
{
    for (int q = threadIdx.x; q < nrep; q += blockDim.x) {
        const int r = image[q];
        e_out[q] = E4[4 * r];
        for (int c = 0; c < 3; ++c) m_out[3 * q + c] = E4[4 * r + 1 + c];
    }
}

// sum of the spins: per-block partials (3 per block), then one block
__global__ void __launch_bounds__(kMcThreads) k_mc_msum(const double *m, int64_t n, int64_t np, double *part)
{
    __shared__ double red[kMcThreads];
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        for (int c = 0; c < 3; ++c) v[c] += m[i + c * np];
    for (int c = 0; c < 3; ++c) {
        const double s = mj_block_sum(v[c], red, blockDim.x);
        if (threadIdx.x == 0) part[3 * blockIdx.x + c] = s;
    }
}

__global__ void __launch_bounds__(kMcThreads) k_mc_msum_final(const double *part, int nparts, double *out)
{
    __shared__ double red[kMcThreads];
    for (int c = 0; c < 3; ++c) {
        double v = 0.0;
        for (int i = threadIdx.x; i < nparts; i += blockDim.x) v += part[3 * i + c];
        const double s = mj_block_sum(v, red, blockDim.x);
        if (threadIdx.x == 0) out[c] = s;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct McCtx {
    // stencil without the on-site entry, and the colouring it admits
    std::vector<int> o3;
    std::vector<double> cf;
    int n_iso = 0, n_dm = 0, n_gen = 0;
    double c0[9] = {0};
    int p = 0, ca = 0, cb = 0;
    std::vector<int64_t> count;        // sites per colour
    std::vector<double> key_cf;        // the classified stencil the above was built from
    std::vector<int> key_o3;
    // device
    int *d_o3 = nullptr;
    double *d_cf = nullptr;
    size_t o3_cap = 0, cf_cap = 0;
    McState *d_st = nullptr;
    double *d_part = nullptr;          // acceptance partials, then spin-sum partials
    int64_t part_cap = 0;
    double *d_samples = nullptr;       // E, then Mx My Mz per sample
    int64_t samples_cap = 0;
    std::vector<int> csh;              // class shift of each stencil entry
    int *d_csh = nullptr;
    size_t csh_cap = 0;
    std::vector<double> cf9;           // each entry as a full block
    double *d_cf9 = nullptr;
    size_t cf9_cap = 0;
    double *d_cm = nullptr;            // colour-major state (and field map)
    int64_t cm_cap = 0;
    // neighbour-slot offsets of interior rows, per class: tab[(k * p + x0) * ne + e]
    std::vector<int> tab;
    int *d_tab = nullptr;
    size_t tab_cap = 0;
    int fast = 0, ylo = 0, yhi = 0, zlo = 0, zhi = 0;
    // parallel tempering: replicas of the lattice, one temperature each
    int nrep = 0;
    std::vector<double> ladder;        // kT of each rung
    std::vector<double> pt_key_cf;     // the Hamiltonian the replicas were set up for
    std::vector<int> pt_key_o3;
    int pt_bmode = 0;
    double *d_rcm = nullptr;           // nrep colour-major states
    McState *d_rst = nullptr;          // counters and cone of each replica
    double *d_rkT = nullptr;           // kT of each replica
    double *d_ladder = nullptr;
    int *d_image = nullptr;            // replica at each rung
    double *d_swaps = nullptr;         // accepted, attempted per rung pair
    double *d_E4 = nullptr;            // E, Mx, My, Mz of each replica
    double *d_rpart = nullptr;         // partials
    int64_t rpart_cap = 0;
    double *d_rsamples = nullptr;
    int64_t rsamples_cap = 0;
    uint64_t n_ex = 0;                 // exchanges so far
};

static int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// colour class of x = (x + a y + b z) mod p such that no stencil offset stays in
// one class; p must divide the periodic wrap of each direction
static bool find_colouring(const std::vector<int> &o3, const int dim[3], const int per[3], int &p_out, int &a_out,
                           int &b_out)
{
    const size_t n = o3.size() / 3;
    bool uses_y = false, uses_z = false;
    for (size_t k = 0; k < n; ++k) {
        if (o3[3 * k] == 0 && o3[3 * k + 1] == 0 && o3[3 * k + 2] == 0) return false;   // self-coupling
        uses_y = uses_y || o3[3 * k + 1] != 0;
        uses_z = uses_z || o3[3 * k + 2] != 0;
    }
    for (int p = 1; p <= 64; ++p) {
        if (per[0] && dim[0] % p) continue;
        for (int a = 0; a < p; ++a) {
            if (per[1] && ((int64_t)a * dim[1]) % p) continue;
            if (!uses_y && a) break;
            for (int b = 0; b < p; ++b) {
                if (per[2] && ((int64_t)b * dim[2]) % p) continue;
                if (!uses_z && b) break;
                bool ok = true;
                for (size_t k = 0; k < n && ok; ++k) {
                    const int64_t v = (int64_t)o3[3 * k] + (int64_t)a * o3[3 * k + 1] + (int64_t)b * o3[3 * k + 2];
                    ok = (v % p) != 0;
                }
                if (ok) {
                    p_out = p;
                    a_out = a;
                    b_out = b;
                    return true;
                }
            }
        }
    }
    return false;
}

static McCtx *mc_of(mj_ctx *c)
{
    if (!c->mc) c->mc = new McCtx();
    return (McCtx *)c->mc;
}

static int mc_prepare(mj_ctx *c, const char *who)
{
    if (c->dip) MJ_FAIL(MJ_ENOTSUP, std::string(who) + ": the FFT term couples every pair of sites (no colouring)");
    if (c->mm_generic) MJ_FAIL(MJ_ENOTSUP, std::string(who) + ": the M-M operator is not translation invariant");
    if (c->dim_mode != 3) MJ_FAIL(MJ_ENOTSUP, std::string(who) + ": one magnetic site per cell only");
    McCtx *mc = mc_of(c);
    if (!mc->d_st) {   // [0] the state between calls, [1] the other half of the in-pass double buffer
        MJ_HIP_CHECK(hipMalloc(&mc->d_st, 2 * sizeof(McState)));
        const McState s0[2] = {{0.0, 1.0, 0.0, 0.0, 0.0}, {0.0, 1.0, 0.0, 0.0, 0.0}};
        MJ_HIP_CHECK(hipMemcpy(mc->d_st, s0, sizeof s0, hipMemcpyHostToDevice));
    }
    if (mc->p > 0 && mc->key_cf == c->h_cf && mc->key_o3 == c->h_ofs3) return MJ_OK;
    // split the on-site entry off the classified stencil
    mc->o3.clear();
    mc->cf.clear();
    std::vector<int> o_iso, o_dm, o_gen;
    std::vector<double> c_iso, c_dm, c_gen;
    for (int q = 0; q < 9; ++q) mc->c0[q] = 0.0;
    const double *cf = c->h_cf.data();
    const int nent = c->n_iso + c->n_dm + c->n_gen;
    if ((int)c->h_ofs3.size() != 3 * nent) MJ_FAIL(MJ_ESTATE, std::string(who) + ": stencil not built");
    for (int k = 0; k < nent; ++k) {
        const int *o = &c->h_ofs3[3 * k];
        const bool onsite = o[0] == 0 && o[1] == 0 && o[2] == 0;
        if (k < c->n_iso) {
            const double s = cf[k];
            if (onsite) {
                mc->c0[0] += s;
                mc->c0[4] += s;
                mc->c0[8] += s;
            } else {
                o_iso.insert(o_iso.end(), o, o + 3);
                c_iso.push_back(s);
            }
        } else if (k < c->n_iso + c->n_dm) {
            const double *d = cf + c->n_iso + 4 * (k - c->n_iso);
            if (onsite) {
                const double s = d[0], wx = d[1], wy = d[2], wz = d[3];
                const double blk[9] = {s, -wz, wy, wz, s, -wx, -wy, wx, s};
                for (int q = 0; q < 9; ++q) mc->c0[q] += blk[q];
            } else {
                o_dm.insert(o_dm.end(), o, o + 3);
                c_dm.insert(c_dm.end(), d, d + 4);
            }
        } else {
            const double *g = cf + c->n_iso + 4 * c->n_dm + 9 * (k - c->n_iso - c->n_dm);
            if (onsite) {
                for (int q = 0; q < 9; ++q) mc->c0[q] += g[q];
            } else {
                o_gen.insert(o_gen.end(), o, o + 3);
                c_gen.insert(c_gen.end(), g, g + 9);
            }
        }
    }
    mc->n_iso = (int)c_iso.size();
    mc->n_dm = (int)c_dm.size() / 4;
    mc->n_gen = (int)c_gen.size() / 9;
    mc->o3 = o_iso;
    mc->o3.insert(mc->o3.end(), o_dm.begin(), o_dm.end());
    mc->o3.insert(mc->o3.end(), o_gen.begin(), o_gen.end());
    mc->cf = c_iso;
    mc->cf.insert(mc->cf.end(), c_dm.begin(), c_dm.end());
    mc->cf.insert(mc->cf.end(), c_gen.begin(), c_gen.end());
    int p = 0, a = 0, b = 0;
    if (!find_colouring(mc->o3, c->dim, c->periodic, p, a, b))
        MJ_FAIL(MJ_ENOTSUP, std::string(who) + ": no colouring of this lattice and stencil (p <= 64)");
    mc->p = p;
    mc->ca = a;
    mc->cb = b;
    mc->count.assign(p, 0);
    for (int z = 0; z < c->dim[2]; ++z)
        for (int y = 0; y < c->dim[1]; ++y)
            for (int k = 0; k < p; ++k) {
                int x0 = (k - a * y - b * z) % p;
                if (x0 < 0) x0 += p;
                if (x0 < c->dim[0]) mc->count[k] += (c->dim[0] - x0 + p - 1) / p;
            }
    // device copies of the stencil
    if (mc->o3.size() > mc->o3_cap) {
        if (mc->d_o3) hipFree(mc->d_o3);
        mc->d_o3 = nullptr;
        mc->o3_cap = mc->o3.size();
        MJ_HIP_CHECK(hipMalloc(&mc->d_o3, sizeof(int) * mc->o3_cap));
    }
    if (mc->cf.size() > mc->cf_cap) {
        if (mc->d_cf) hipFree(mc->d_cf);
        mc->d_cf = nullptr;
        mc->cf_cap = mc->cf.size();
        MJ_HIP_CHECK(hipMalloc(&mc->d_cf, sizeof(double) * mc->cf_cap));
    }
    if (!mc->o3.empty())
        MJ_HIP_CHECK(hipMemcpy(mc->d_o3, mc->o3.data(), sizeof(int) * mc->o3.size(), hipMemcpyHostToDevice));
    if (!mc->cf.empty())
        MJ_HIP_CHECK(hipMemcpy(mc->d_cf, mc->cf.data(), sizeof(double) * mc->cf.size(), hipMemcpyHostToDevice));
    const size_t ne = mc->o3.size() / 3;
    mc->csh.assign(ne, 0);
    for (size_t e = 0; e < ne; ++e) {
        const int64_t v = (int64_t)mc->o3[3 * e] + (int64_t)a * mc->o3[3 * e + 1] + (int64_t)b * mc->o3[3 * e + 2];
        mc->csh[e] = (int)(((v % p) + p) % p);
    }
    if (ne > mc->csh_cap) {
        if (mc->d_csh) hipFree(mc->d_csh);
    if (mc->d_cf9) hipFree(mc->d_cf9);
        mc->d_csh = nullptr;
        mc->csh_cap = ne;
        MJ_HIP_CHECK(hipMalloc(&mc->d_csh, sizeof(int) * ne));
    }
    if (ne) MJ_HIP_CHECK(hipMemcpy(mc->d_csh, mc->csh.data(), sizeof(int) * ne, hipMemcpyHostToDevice));
    mc->cf9.clear();
    for (int e = 0; e < mc->n_iso; ++e) {
        const double v = mc->cf[e];
        const double blk[9] = {v, 0.0, 0.0, 0.0, v, 0.0, 0.0, 0.0, v};
        mc->cf9.insert(mc->cf9.end(), blk, blk + 9);
    }
    for (int e = 0; e < mc->n_dm; ++e) {
        const double *d = &mc->cf[mc->n_iso + 4 * e];
        const double v = d[0], wx = d[1], wy = d[2], wz = d[3];
        const double blk[9] = {v, -wz, wy, wz, v, -wx, -wy, wx, v};
        mc->cf9.insert(mc->cf9.end(), blk, blk + 9);
    }
    mc->cf9.insert(mc->cf9.end(), mc->cf.begin() + mc->n_iso + 4 * mc->n_dm, mc->cf.end());
    if (mc->cf9.size() > mc->cf9_cap) {
        if (mc->d_cf9) hipFree(mc->d_cf9);
        mc->d_cf9 = nullptr;
        mc->cf9_cap = mc->cf9.size();
        MJ_HIP_CHECK(hipMalloc(&mc->d_cf9, sizeof(double) * mc->cf9_cap));
    }
    if (!mc->cf9.empty())
        MJ_HIP_CHECK(hipMemcpy(mc->d_cf9, mc->cf9.data(), sizeof(double) * mc->cf9.size(), hipMemcpyHostToDevice));
    // slot offsets for rows whose neighbours need no wrap in y and z
    mc->ylo = mc->yhi = mc->zlo = mc->zhi = 0;
    for (size_t e = 0; e < ne; ++e) {
        mc->ylo = std::max(mc->ylo, -mc->o3[3 * e + 1]);
        mc->yhi = std::max(mc->yhi, mc->o3[3 * e + 1]);
        mc->zlo = std::max(mc->zlo, -mc->o3[3 * e + 2]);
        mc->zhi = std::max(mc->zhi, mc->o3[3 * e + 2]);
    }
    {
        const int64_t W = ceil_div(c->dim[0], p), RW = (int64_t)c->dim[1] * c->dim[2] * W;
        mc->fast = ne > 0 && (int64_t)p * (int64_t)ne <= kMcTab && (int64_t)p * RW < INT32_MAX &&
                   (!c->periodic[0] || c->dim[0] % p == 0);
        mc->tab.assign(mc->fast ? (size_t)p * p * ne : 0, 0);
        if (mc->fast) {
            for (int k = 0; k < p; ++k)
                for (int x0 = 0; x0 < p; ++x0)
                    for (size_t e = 0; e < ne; ++e) {
                        const int dx = mc->o3[3 * e], dy = mc->o3[3 * e + 1], dz = mc->o3[3 * e + 2];
                        int u = x0 + dx, q = 0;
                        while (u < 0) {
                            u += p;
                            --q;
                        }
                        while (u >= p) {
                            u -= p;
                            ++q;
                        }
                        const int k2 = (k + mc->csh[e]) % p;
                        const int64_t d = (int64_t)(k2 - k) * RW + ((int64_t)dy + (int64_t)c->dim[1] * dz) * W + q;
                        mc->tab[((size_t)k * p + x0) * ne + e] = (int)d;
                    }
            if (mc->tab.size() > mc->tab_cap) {
                if (mc->d_tab) hipFree(mc->d_tab);
                mc->d_tab = nullptr;
                mc->tab_cap = mc->tab.size();
                MJ_HIP_CHECK(hipMalloc(&mc->d_tab, sizeof(int) * mc->tab_cap));
            }
            MJ_HIP_CHECK(hipMemcpy(mc->d_tab, mc->tab.data(), sizeof(int) * mc->tab.size(), hipMemcpyHostToDevice));
        }
    }
    mc->key_cf = c->h_cf;
    mc->key_o3 = c->h_ofs3;
    return MJ_OK;
}

static void mc_args(const mj_ctx *c, const McCtx *mc, McArgs &a)
{
    memset(&a, 0, sizeof a);
    a.m = c->A;
    a.np = c->np;
    a.nx = c->dim[0];
    a.ny = c->dim[1];
    a.nz = c->dim[2];
    a.px = c->periodic[0];
    a.py = c->periodic[1];
    a.pz = c->periodic[2];
    a.p = mc->p;
    a.ca = mc->ca;
    a.cb = mc->cb;
    a.n_iso = mc->n_iso;
    a.n_dm = mc->n_dm;
    a.n_gen = mc->n_gen;
    a.o3 = mc->d_o3;
    a.cf = mc->d_cf;
    for (int q = 0; q < 9; ++q) a.c0[q] = mc->c0[q];
    a.bmode = c->bmode;
    for (int q = 0; q < 3; ++q) a.zb[q] = c->zb[q];
    for (int q = 0; q < 9; ++q) a.z[q] = c->zblk[q];
    a.bmap = c->d_bmap;
    a.st = mc->d_st;
    a.W = (int)ceil_div(c->dim[0], mc->p);
    a.cs = (int64_t)mc->p * c->dim[1] * c->dim[2] * a.W;
    a.csh = mc->d_csh;
    a.cf9 = mc->d_cf9;
    a.fast = mc->fast;
    a.ne = (int)(mc->o3.size() / 3);
    a.ylo = mc->ylo;
    a.yhi = mc->yhi;
    a.zlo = mc->zlo;
    a.zhi = mc->zhi;
}

static int ensure_part(McCtx *mc, int64_t n)
{
    if (mc->part_cap >= n) return MJ_OK;
    if (mc->d_part) hipFree(mc->d_part);
    mc->d_part = nullptr;
    MJ_HIP_CHECK(hipMalloc(&mc->d_part, sizeof(double) * (size_t)n));
    mc->part_cap = n;
    return MJ_OK;
}

static int launch_check(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) MJ_FAIL(MJ_EHIP, std::string(what) + ": " + hipGetErrorString(e));
    return MJ_OK;
}

}  // namespace



void mj_mc_destroy(mj_ctx *c)
{
    McCtx *mc = (McCtx *)c->mc;
    if (!mc) return;
    if (mc->d_o3) hipFree(mc->d_o3);
    if (mc->d_cf) hipFree(mc->d_cf);
    if (mc->d_st) hipFree(mc->d_st);
    if (mc->d_part) hipFree(mc->d_part);
    if (mc->d_samples) hipFree(mc->d_samples);
    if (mc->d_csh) hipFree(mc->d_csh);
    if (mc->d_cm) hipFree(mc->d_cm);
    if (mc->d_tab) hipFree(mc->d_tab);
    void *pt[] = {mc->d_rcm, mc->d_rst, mc->d_rkT, mc->d_ladder, mc->d_image, mc->d_swaps, mc->d_E4, mc->d_rpart,
                  mc->d_rsamples};
    for (void *q : pt)
        if (q) hipFree(q);
    delete mc;
    c->mc = nullptr;
}

extern "C" int mj_energy_single(mj_ctx *c, int64_t n, const int64_t *sites1, const double *spins, double *E)
{
    if (!c || n < 0 || (n > 0 && (!sites1 || !E))) MJ_FAIL(MJ_EINVAL, "mj_energy_single: null argument");
    if (!c->state_set) MJ_FAIL(MJ_ESTATE, "mj_energy_single: no state set");
    if (n == 0) return MJ_OK;
    for (int64_t k = 0; k < n; ++k)
        if (sites1[k] < 1 || sites1[k] > c->n) MJ_FAIL(MJ_EINVAL, "mj_energy_single: site index outside 1..Ncell");
    MJ_HIP_CHECK(hipSetDevice(c->device));
    MJ_TRY(mj_finalize(c));
    MJ_TRY(mc_prepare(c, "mj_energy_single"));
    McCtx *mc = mc_of(c);
    std::vector<int64_t> s0(n);
    for (int64_t k = 0; k < n; ++k) s0[k] = sites1[k] - 1;
    int64_t *d_sites = nullptr;
    double *d_spins = nullptr, *d_e = nullptr;
    int rc = MJ_OK;
    if (hipMalloc(&d_sites, sizeof(int64_t) * n) != hipSuccess || hipMalloc(&d_e, sizeof(double) * n) != hipSuccess ||
        (spins && hipMalloc(&d_spins, sizeof(double) * 3 * n) != hipSuccess)) {
        mj_set_error("mj_energy_single: device allocation failed");
        rc = MJ_ENOMEM;
    }
    if (rc == MJ_OK && (hipMemcpy(d_sites, s0.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice) != hipSuccess ||
                        (spins && hipMemcpy(d_spins, spins, sizeof(double) * 3 * n, hipMemcpyHostToDevice) != hipSuccess))) {
        mj_set_error("mj_energy_single: copy to the device failed");
        rc = MJ_EHIP;
    }
    if (rc == MJ_OK) {
        McArgs a;
        mc_args(c, mc, a);
        a.nsites = n;
        a.sites = d_sites;
        a.spins = d_spins;
        a.e_out = d_e;
        hipLaunchKernelGGL(k_mc_energy_single, dim3((unsigned)ceil_div(n, kMcThreads)), dim3(kMcThreads), 0, c->stream,
                           a);
        rc = launch_check("energy_single launch");
        if (rc == MJ_OK && (hipMemcpyAsync(E, d_e, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                            hipStreamSynchronize(c->stream) != hipSuccess)) {
            mj_set_error("mj_energy_single: copy back failed");
            rc = MJ_EHIP;
        }
    }
    if (d_sites) hipFree(d_sites);
    if (d_spins) hipFree(d_spins);
    if (d_e) hipFree(d_e);
    return rc;
}

extern "C" int mj_mc_set_cone(mj_ctx *c, double cone)
{
    if (!c || !(cone >= 0.0)) MJ_FAIL(MJ_EINVAL, "mj_mc_set_cone: invalid argument");
    MJ_HIP_CHECK(hipSetDevice(c->device));
    McCtx *mc = mc_of(c);
    if (!mc->d_st) MJ_HIP_CHECK(hipMalloc(&mc->d_st, 2 * sizeof(McState)));
    const McState s0 = {cone, cos(cone), 0.0, 0.0, 0.0};
    MJ_HIP_CHECK(hipMemcpy(mc->d_st, &s0, sizeof s0, hipMemcpyHostToDevice));
    return MJ_OK;
}

extern "C" int mj_mc_stats(mj_ctx *c, double *out4, int *colouring3)
{
    if (!c || !out4) MJ_FAIL(MJ_EINVAL, "mj_mc_stats: null argument");
    MJ_HIP_CHECK(hipSetDevice(c->device));
    McCtx *mc = mc_of(c);
    McState s = {0.0, 1.0, 0.0, 0.0, 0.0};
    if (mc->d_st) {
        MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
        MJ_HIP_CHECK(hipMemcpy(&s, mc->d_st, sizeof s, hipMemcpyDeviceToHost));
    }
    out4[0] = s.cone;
    out4[1] = s.acc;
    out4[2] = s.nb;
    out4[3] = s.rate;
    if (colouring3) {
        if (c->state_set) {
            MJ_TRY(mj_finalize(c));
            MJ_TRY(mc_prepare(c, "mj_mc_stats"));
        }
        colouring3[0] = mc->p;
        colouring3[1] = mc->ca;
        colouring3[2] = mc->cb;
    }
    return MJ_OK;
}

extern "C" int mj_mc_sweeps(mj_ctx *c, int nsweeps, int sampler, double kT, uint64_t seed, uint64_t sweep0,
                            int measure_every, double *E_samples, double *M_samples)
{
    if (!c || nsweeps < 0) MJ_FAIL(MJ_EINVAL, "mj_mc_sweeps: invalid argument");
    if (sampler != MJ_MC_SPHERE && sampler != MJ_MC_ISING) MJ_FAIL(MJ_EINVAL, "mj_mc_sweeps: unknown sampler");
    if (!(kT >= 0.0)) MJ_FAIL(MJ_EINVAL, "mj_mc_sweeps: kT must be >= 0");
    if (!c->state_set) MJ_FAIL(MJ_ESTATE, "mj_mc_sweeps: no state set");
    if (measure_every < 0) MJ_FAIL(MJ_EINVAL, "mj_mc_sweeps: measure_every must be >= 0");
    const int64_t nsamp = measure_every > 0 ? nsweeps / measure_every : 0;
    if (nsamp > 0 && !E_samples && !M_samples) MJ_FAIL(MJ_EINVAL, "mj_mc_sweeps: no sample array");
    MJ_HIP_CHECK(hipSetDevice(c->device));
    MJ_TRY(mj_finalize(c));
    MJ_TRY(mc_prepare(c, "mj_mc_sweeps"));
    McCtx *mc = mc_of(c);
    if (nsweeps == 0) return MJ_OK;
    const int adapt = sampler == MJ_MC_SPHERE ? 1 : 0;
    const int p = mc->p;
    McArgs a;
    mc_args(c, mc, a);
    const int64_t nthreads = (int64_t)c->dim[1] * c->dim[2] * a.W;
    const int64_t nblocks = ceil_div(nthreads, kMcThreads);
    if (nblocks > INT32_MAX) MJ_FAIL(MJ_ENOTSUP, "mj_mc_sweeps: lattice too large");
    const int64_t msum_blocks = std::min<int64_t>(ceil_div(c->n, kMcThreads), 4096);
    if ((int64_t)p * nblocks > INT32_MAX) MJ_FAIL(MJ_ENOTSUP, "mj_mc_sweeps: lattice too large");
    MJ_TRY(ensure_part(mc, (int64_t)p * nblocks + 3 * msum_blocks));
    double *msum_part = mc->d_part + (size_t)p * nblocks;
    if (nsamp > 0 && mc->samples_cap < 4 * nsamp) {
        if (mc->d_samples) hipFree(mc->d_samples);
        mc->d_samples = nullptr;
        MJ_HIP_CHECK(hipMalloc(&mc->d_samples, sizeof(double) * 4 * (size_t)nsamp));
        mc->samples_cap = 4 * nsamp;
    }
    const int64_t need = (c->bmode == 2 ? 6 : 3) * a.cs;
    if (mc->cm_cap < need) {
        if (mc->d_cm) hipFree(mc->d_cm);
        mc->d_cm = nullptr;
        MJ_HIP_CHECK(hipMalloc(&mc->d_cm, sizeof(double) * (size_t)need));
        mc->cm_cap = need;
    }
    a.cm = mc->d_cm;
    a.cbmap = c->bmode == 2 ? mc->d_cm + 3 * a.cs : nullptr;
    a.nthreads = nthreads;
    a.kT = kT;
    a.seed = seed;
    a.sampler = sampler;
    const unsigned site_blocks = (unsigned)ceil_div(c->n, kMcThreads);
    auto to_sites = [&]() -> int {
        hipLaunchKernelGGL(k_mc_layout, dim3(site_blocks), dim3(kMcThreads), 0, c->stream, a, (const double *)a.cm,
                           a.cs, c->A, c->np, 0);
        return launch_check("layout launch");
    };
    hipLaunchKernelGGL(k_mc_layout, dim3(site_blocks), dim3(kMcThreads), 0, c->stream, a, (const double *)c->A, c->np,
                       a.cm, a.cs, 1);
    MJ_TRY(launch_check("layout launch"));
    if (a.cbmap) {
        hipLaunchKernelGGL(k_mc_layout, dim3(site_blocks), dim3(kMcThreads), 0, c->stream, a,
                           (const double *)c->d_bmap, c->np, (double *)a.cbmap, a.cs, 1);
        MJ_TRY(launch_check("layout launch"));
    }
    McState *st = mc->d_st;
    const double nsite = (double)c->n;
    // cone of the first sweep
    hipLaunchKernelGGL(k_mc_fold, dim3(1), dim3(kMcThreads), 0, c->stream, st, (const double *)mc->d_part, 0,
                       (int64_t)0, 0.0, nsite, adapt);
    MJ_TRY(launch_check("fold launch"));
    a.st = st;
    int64_t js = 0;
    for (int s = 0; s < nsweeps; ++s) {
        a.sweep = sweep0 + (uint64_t)s;
        for (int k = 0; k < p; ++k) {
            a.color = k;
            a.tab = mc->fast ? mc->d_tab + (size_t)k * p * a.ne : nullptr;
            a.acc_part = mc->d_part + (size_t)k * nblocks;
            hipLaunchKernelGGL(k_mc_pass, dim3((unsigned)nblocks), dim3(kMcThreads), 0, c->stream, a);
            MJ_TRY(launch_check("Monte Carlo pass launch"));
        }
        const double n_next = s == nsweeps - 1 ? 0.0 : nsite;
        hipLaunchKernelGGL(k_mc_fold, dim3(1), dim3(kMcThreads), 0, c->stream, st, (const double *)mc->d_part,
                           (int)(p * nblocks), (int64_t)0, nsite, n_next, adapt);
        MJ_TRY(launch_check("fold launch"));
        if (measure_every > 0 && (s + 1) % measure_every == 0 && js < nsamp) {
            MJ_TRY(to_sites());
            MJ_TRY(mj_energy_into(c, mc->d_samples + js));
            hipLaunchKernelGGL(k_mc_msum, dim3((unsigned)msum_blocks), dim3(kMcThreads), 0, c->stream,
                               (const double *)c->A, c->n, c->np, msum_part);
            MJ_TRY(launch_check("spin sum launch"));
            hipLaunchKernelGGL(k_mc_msum_final, dim3(1), dim3(kMcThreads), 0, c->stream, (const double *)msum_part,
                               (int)msum_blocks, mc->d_samples + nsamp + 3 * js);
            MJ_TRY(launch_check("spin sum launch"));
            ++js;
        }
    }
    MJ_TRY(to_sites());
    if (nsamp > 0) {
        if (E_samples)
            MJ_HIP_CHECK(hipMemcpyAsync(E_samples, mc->d_samples, sizeof(double) * nsamp, hipMemcpyDeviceToHost,
                                        c->stream));
        if (M_samples)
            MJ_HIP_CHECK(hipMemcpyAsync(M_samples, mc->d_samples + nsamp, sizeof(double) * 3 * nsamp,
                                        hipMemcpyDeviceToHost, c->stream));
    }
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

// ---------------------------------------------------------------------------
// parallel tempering (paratemp-MC/parallel_tempering.f90:24-217)
// ---------------------------------------------------------------------------
static void pt_free(McCtx *mc)
{
    void *pt[] = {mc->d_rcm, mc->d_rst, mc->d_rkT, mc->d_ladder, mc->d_image, mc->d_swaps, mc->d_E4};
    for (void *q : pt)
        if (q) hipFree(q);
    mc->d_rcm = nullptr;
    mc->d_rst = nullptr;
    mc->d_rkT = nullptr;
    mc->d_ladder = nullptr;
    mc->d_image = nullptr;
    mc->d_swaps = nullptr;
    mc->d_E4 = nullptr;
    mc->nrep = 0;
}

extern "C" int mj_mc_pt_setup(mj_ctx *c, int nrep, const double *kT_ladder, double cone)
{
    if (!c || nrep < 1 || !kT_ladder || !(cone >= 0.0)) MJ_FAIL(MJ_EINVAL, "mj_mc_pt_setup: invalid argument");
    for (int q = 0; q < nrep; ++q)
        if (!(kT_ladder[q] > 0.0)) MJ_FAIL(MJ_EINVAL, "mj_mc_pt_setup: temperatures must be > 0");
    if (!c->state_set) MJ_FAIL(MJ_ESTATE, "mj_mc_pt_setup: no state set");
    MJ_HIP_CHECK(hipSetDevice(c->device));
    MJ_TRY(mj_finalize(c));
    MJ_TRY(mc_prepare(c, "mj_mc_pt_setup"));
    McCtx *mc = mc_of(c);
    pt_free(mc);
    McArgs a;
    mc_args(c, mc, a);
    const size_t rep_bytes = sizeof(double) * 3 * (size_t)a.cs;
    MJ_HIP_CHECK(hipMalloc(&mc->d_rcm, rep_bytes * nrep));
    MJ_HIP_CHECK(hipMalloc(&mc->d_rst, sizeof(McState) * nrep));
    MJ_HIP_CHECK(hipMalloc(&mc->d_rkT, sizeof(double) * nrep));
    MJ_HIP_CHECK(hipMalloc(&mc->d_ladder, sizeof(double) * nrep));
    MJ_HIP_CHECK(hipMalloc(&mc->d_image, sizeof(int) * nrep));
    MJ_HIP_CHECK(hipMalloc(&mc->d_swaps, sizeof(double) * 2 * nrep));
    MJ_HIP_CHECK(hipMalloc(&mc->d_E4, sizeof(double) * 4 * nrep));
    mc->nrep = nrep;
    mc->ladder.assign(kT_ladder, kT_ladder + nrep);
    mc->pt_key_cf = c->h_cf;
    mc->pt_key_o3 = c->h_ofs3;
    mc->pt_bmode = c->bmode;
    mc->n_ex = 0;
    std::vector<McState> st(nrep, McState{cone, cos(cone), 0.0, 0.0, 0.0});
    std::vector<int> image(nrep);
    for (int q = 0; q < nrep; ++q) image[q] = q;
    std::vector<double> zeros(2 * nrep, 0.0);
    MJ_HIP_CHECK(hipMemcpy(mc->d_rst, st.data(), sizeof(McState) * nrep, hipMemcpyHostToDevice));
    MJ_HIP_CHECK(hipMemcpy(mc->d_rkT, kT_ladder, sizeof(double) * nrep, hipMemcpyHostToDevice));
    MJ_HIP_CHECK(hipMemcpy(mc->d_ladder, kT_ladder, sizeof(double) * nrep, hipMemcpyHostToDevice));
    MJ_HIP_CHECK(hipMemcpy(mc->d_image, image.data(), sizeof(int) * nrep, hipMemcpyHostToDevice));
    MJ_HIP_CHECK(hipMemcpy(mc->d_swaps, zeros.data(), sizeof(double) * 2 * nrep, hipMemcpyHostToDevice));
    // every replica starts from the context's state
    const unsigned site_blocks = (unsigned)ceil_div(c->n, kMcThreads);
    hipLaunchKernelGGL(k_mc_layout, dim3(site_blocks), dim3(kMcThreads), 0, c->stream, a, (const double *)c->A, c->np,
                       mc->d_rcm, a.cs, 1);
    MJ_TRY(launch_check("layout launch"));
    for (int r = 1; r < nrep; ++r)
        MJ_HIP_CHECK(hipMemcpyAsync(mc->d_rcm + (size_t)r * 3 * a.cs, mc->d_rcm, rep_bytes, hipMemcpyDeviceToDevice,
                                    c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

extern "C" int mj_mc_pt_sweeps(mj_ctx *c, int nsweeps, int sampler, uint64_t seed, uint64_t sweep0,
                               int exchange_every, int measure_every, double *E_samples, double *M_samples)
{
    if (!c || nsweeps < 0 || exchange_every < 0 || measure_every < 0)
        MJ_FAIL(MJ_EINVAL, "mj_mc_pt_sweeps: invalid argument");
    if (sampler != MJ_MC_SPHERE && sampler != MJ_MC_ISING) MJ_FAIL(MJ_EINVAL, "mj_mc_pt_sweeps: unknown sampler");
    McCtx *mc = (McCtx *)c->mc;
    if (!mc || mc->nrep < 1) MJ_FAIL(MJ_ESTATE, "mj_mc_pt_sweeps: no replicas (mj_mc_pt_setup)");
    MJ_HIP_CHECK(hipSetDevice(c->device));
    MJ_TRY(mj_finalize(c));
    if (mc->pt_key_cf != c->h_cf || mc->pt_key_o3 != c->h_ofs3 || mc->pt_bmode != c->bmode)
        MJ_FAIL(MJ_ESTATE, "mj_mc_pt_sweeps: the Hamiltonian changed since mj_mc_pt_setup");
    MJ_TRY(mc_prepare(c, "mj_mc_pt_sweeps"));
    const int64_t nsamp = measure_every > 0 ? nsweeps / measure_every : 0;
    if (nsamp > 0 && !E_samples && !M_samples) MJ_FAIL(MJ_EINVAL, "mj_mc_pt_sweeps: no sample array");
    if (nsweeps == 0) return MJ_OK;
    const int nrep = mc->nrep, p = mc->p;
    const int adapt = sampler == MJ_MC_SPHERE ? 1 : 0;
    McArgs a;
    mc_args(c, mc, a);
    const int64_t nthreads = (int64_t)c->dim[1] * c->dim[2] * a.W;
    const int64_t nblocks = ceil_div(nthreads, kMcThreads);
    const int64_t eblocks = ceil_div((int64_t)p * c->dim[1] * c->dim[2] * a.W, kMcThreads);
    if (nblocks > INT32_MAX || eblocks > INT32_MAX || nrep > 65535)
        MJ_FAIL(MJ_ENOTSUP, "mj_mc_pt_sweeps: lattice or replica count too large");
    const int64_t rep_part = (int64_t)p * nblocks;
    const int64_t need = std::max<int64_t>(rep_part, 4 * eblocks) * nrep;
    if (mc->rpart_cap < need) {
        if (mc->d_rpart) hipFree(mc->d_rpart);
        mc->d_rpart = nullptr;
        MJ_HIP_CHECK(hipMalloc(&mc->d_rpart, sizeof(double) * (size_t)need));
        mc->rpart_cap = need;
    }
    if (nsamp > 0 && mc->rsamples_cap < 4 * nsamp * nrep) {
        if (mc->d_rsamples) hipFree(mc->d_rsamples);
        mc->d_rsamples = nullptr;
        MJ_HIP_CHECK(hipMalloc(&mc->d_rsamples, sizeof(double) * 4 * (size_t)nsamp * nrep));
        mc->rsamples_cap = 4 * nsamp * nrep;
    }
    if (c->bmode == 2) {   // field map in colour-major order (shared by the replicas)
        if (mc->cm_cap < 6 * a.cs) {
            if (mc->d_cm) hipFree(mc->d_cm);
            mc->d_cm = nullptr;
            MJ_HIP_CHECK(hipMalloc(&mc->d_cm, sizeof(double) * 6 * (size_t)a.cs));
            mc->cm_cap = 6 * a.cs;
        }
        a.cbmap = mc->d_cm + 3 * a.cs;
        hipLaunchKernelGGL(k_mc_layout, dim3((unsigned)ceil_div(c->n, kMcThreads)), dim3(kMcThreads), 0, c->stream, a,
                           (const double *)c->d_bmap, c->np, (double *)a.cbmap, a.cs, 1);
        MJ_TRY(launch_check("layout launch"));
    }
    a.cm = mc->d_rcm;
    a.rep_cm = 3 * a.cs;
    a.rep_part = rep_part;
    a.kT_rep = mc->d_rkT;
    a.st = mc->d_rst;
    a.nthreads = nthreads;
    a.seed = seed;
    a.sampler = sampler;
    const double nsite = (double)c->n;
    hipLaunchKernelGGL(k_mc_fold, dim3(nrep), dim3(kMcThreads), 0, c->stream, mc->d_rst, (const double *)mc->d_rpart,
                       0, rep_part, 0.0, nsite, adapt);
    MJ_TRY(launch_check("fold launch"));
    auto energies = [&]() -> int {
        McArgs e = a;
        e.tab = mc->d_tab;
        hipLaunchKernelGGL(k_mc_rep_energy, dim3((unsigned)eblocks, nrep), dim3(kMcThreads), 0, c->stream, e,
                           mc->d_rpart);
        MJ_TRY(launch_check("replica energy launch"));
        hipLaunchKernelGGL(k_mc_rep_reduce, dim3(nrep), dim3(kMcThreads), 0, c->stream, (const double *)mc->d_rpart,
                           (int)eblocks, mc->d_E4);
        return launch_check("replica energy launch");
    };
    int64_t js = 0;
    for (int s = 0; s < nsweeps; ++s) {
        a.sweep = sweep0 + (uint64_t)s;
        for (int k = 0; k < p; ++k) {
            a.color = k;
            a.tab = mc->fast ? mc->d_tab + (size_t)k * p * a.ne : nullptr;
            a.acc_part = mc->d_rpart + (size_t)k * nblocks;
            hipLaunchKernelGGL(k_mc_pass, dim3((unsigned)nblocks, nrep), dim3(kMcThreads), 0, c->stream, a);
            MJ_TRY(launch_check("Monte Carlo pass launch"));
        }
        const double n_next = s == nsweeps - 1 ? 0.0 : nsite;
        hipLaunchKernelGGL(k_mc_fold, dim3(nrep), dim3(kMcThreads), 0, c->stream, mc->d_rst,
                           (const double *)mc->d_rpart, (int)rep_part, rep_part, nsite, n_next, adapt);
        MJ_TRY(launch_check("fold launch"));
        const bool ex = exchange_every > 0 && (s + 1) % exchange_every == 0;
        const bool me = measure_every > 0 && (s + 1) % measure_every == 0 && js < nsamp;
        if (ex || me) MJ_TRY(energies());
        if (me) {   // sampled before the exchange, by rung
            hipLaunchKernelGGL(k_mc_rep_samples, dim3(1), dim3(kMcThreads), 0, c->stream, (const double *)mc->d_E4,
                               (const int *)mc->d_image, nrep, mc->d_rsamples + js * nrep,
                               mc->d_rsamples + nsamp * nrep + 3 * js * nrep);
            MJ_TRY(launch_check("sample launch"));
            ++js;
        }
        if (ex && nrep > 1) {
            hipLaunchKernelGGL(k_mc_exchange, dim3(1), dim3(64), 0, c->stream, (const double *)mc->d_E4,
                               (const double *)mc->d_ladder, mc->d_image, mc->d_rkT, mc->d_swaps, nrep, seed, mc->n_ex);
            MJ_TRY(launch_check("exchange launch"));
            ++mc->n_ex;
        }
    }
    if (nsamp > 0) {
        if (E_samples)
            MJ_HIP_CHECK(hipMemcpyAsync(E_samples, mc->d_rsamples, sizeof(double) * nsamp * nrep,
                                        hipMemcpyDeviceToHost, c->stream));
        if (M_samples)
            MJ_HIP_CHECK(hipMemcpyAsync(M_samples, mc->d_rsamples + nsamp * nrep, sizeof(double) * 3 * nsamp * nrep,
                                        hipMemcpyDeviceToHost, c->stream));
    }
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

extern "C" int mj_mc_pt_state(mj_ctx *c, int *image_of_rung, double *swaps_accepted, double *swaps_attempted,
                              double *replica_stats)
{
    if (!c) MJ_FAIL(MJ_EINVAL, "mj_mc_pt_state: null argument");
    McCtx *mc = (McCtx *)c->mc;
    if (!mc || mc->nrep < 1) MJ_FAIL(MJ_ESTATE, "mj_mc_pt_state: no replicas");
    MJ_HIP_CHECK(hipSetDevice(c->device));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    const int nrep = mc->nrep;
    if (image_of_rung) MJ_HIP_CHECK(hipMemcpy(image_of_rung, mc->d_image, sizeof(int) * nrep, hipMemcpyDeviceToHost));
    std::vector<double> sw(2 * nrep);
    MJ_HIP_CHECK(hipMemcpy(sw.data(), mc->d_swaps, sizeof(double) * 2 * nrep, hipMemcpyDeviceToHost));
    for (int q = 0; q + 1 < nrep; ++q) {
        if (swaps_accepted) swaps_accepted[q] = sw[2 * q];
        if (swaps_attempted) swaps_attempted[q] = sw[2 * q + 1];
    }
    if (replica_stats) {
        std::vector<McState> st(nrep);
        MJ_HIP_CHECK(hipMemcpy(st.data(), mc->d_rst, sizeof(McState) * nrep, hipMemcpyDeviceToHost));
        for (int r = 0; r < nrep; ++r) {
            replica_stats[4 * r] = st[r].cone;
            replica_stats[4 * r + 1] = st[r].acc;
            replica_stats[4 * r + 2] = st[r].nb;
            replica_stats[4 * r + 3] = st[r].rate;
        }
    }
    return MJ_OK;
}

int mj_launch_soa2aos(const double *in, double *out, int64_t n, int64_t np, double eps, hipStream_t s);

extern "C" int mj_mc_pt_get_state(mj_ctx *c, int rung, double *m, int64_t n3)
{
    if (!c || !m || n3 != 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_mc_pt_get_state: size must be 3*Ncell");
    McCtx *mc = (McCtx *)c->mc;
    if (!mc || mc->nrep < 1) MJ_FAIL(MJ_ESTATE, "mj_mc_pt_get_state: no replicas");
    if (rung < 0 || rung >= mc->nrep) MJ_FAIL(MJ_EINVAL, "mj_mc_pt_get_state: no such rung");
    MJ_HIP_CHECK(hipSetDevice(c->device));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    int r = 0;
    MJ_HIP_CHECK(hipMemcpy(&r, mc->d_image + rung, sizeof(int), hipMemcpyDeviceToHost));
    McArgs a;
    mc_args(c, mc, a);
    hipLaunchKernelGGL(k_mc_layout, dim3((unsigned)ceil_div(c->n, kMcThreads)), dim3(kMcThreads), 0, c->stream, a,
                       (const double *)(mc->d_rcm + (size_t)r * 3 * a.cs), a.cs, c->Bb, c->np, 0);
    MJ_TRY(launch_check("layout launch"));
    MJ_TRY(mj_launch_soa2aos(c->Bb, c->Nb, c->n, c->np, 1.0e-50, c->stream));
    MJ_HIP_CHECK(hipMemcpyAsync(m, c->Nb, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}
