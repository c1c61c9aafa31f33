// Batched LLG: independent systems of uneven sizes and parallel-tempering
// replicas advanced together, one launch per integrator stage (integration 2:
// one fused launch per step) for the whole batch, the steps between replica
// exchanges captured in one hipGraph; integration 2 can also run each such
// segment as one persistent launch with grid barriers.
//
// Every system keeps the single-system path's arithmetic: its translation-
// invariant stencil (classified iso | dm | gen by mj_finalize), Zeeman term,
// damping and Philox stream (seed, step, site) are copied from a finalized
// mj_ctx, and the stage kernel runs the same body (mj_stage_body.h) with a
// geometry policy that reads the system of each workgroup from a table.  A
// system's trajectory therefore equals the single-system one bit for bit,
// whatever the tile shape.
//
// Layout in HBM: the systems' sites are concatenated (each starting on a
// multiple of 64), SoA [mx | my | mz] with one stride for the batch.  The
// work list is the concatenation of every system's tiles (256 threads each;
// per-system tile shapes); workgroup -> tile goes through the XCD remap over
// the whole batch.
//
// Parallel tempering: replicas of one Hamiltonian hold the temperatures of a
// ladder; every `every` steps the energies of the states entering the step
// are reduced per replica (in the stage-1 kernel) and neighbouring
// temperatures are swapped with probability min(1, exp((b_t - b_t+1)(E_a - E_b)))
// (even pairs on even exchanges, odd pairs on odd ones) by a one-thread kernel;
// only the temperature assignment moves, the states stay in place.  The step
// counter lives in device memory, so the captured graph is replayed as is.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "mj_device.h"
#include "mj_internal.h"
#include "mj_stage_body.h"

struct BatchSys {
    int nx, ny, nz, px, py, pz;
    int bx, by, bz, rx, ry, rz, sx, sy, sz, ntx, nty, ntz, nblocks, block0;
    int n_iso, n_dm, n_gen, ofs_off, cf_off, bmode;
    int ofs0_off;         // the stencil's offsets in the fused step's S0 layout (tile + 2R)
    int64_t base;
    const double *hdip;   // dipolar systems: the padded grid of H_fft (null: none), its row / plane / component strides
    int dnx, dny;
    int64_t hstride;
    double zb[3];
    double z9[9];         // the Zeeman block, for a per-site B map (bmode 2)
    double alpha, den, rden;
    uint64_t seed;
};

struct BatchGeom : RtGeom {
    static __device__ __forceinline__ int block(const KArgs &a) { return a.vbid; }
    static __device__ __forceinline__ int part(const KArgs &a) { return a.vbase + a.vbid; }
};

#define MJ_BATCH_NTHR 256

// the arguments of tile gb (a tile of system S) from the batch-wide ones:
// the system's geometry, stencil, Zeeman term, damping and Philox stream, its
// slice of the state buffers, its thermal amplitude and the absolute step;
// DIP: the batch holds dipolar systems (their H_fft grids)
template <bool DIP = false>
__device__ __forceinline__ KArgs sys_args(const KArgs &base, const BatchSys &S, int gb, double sigma, uint64_t step)
{
    KArgs a = base;
    a.nx = S.nx;
    a.ny = S.ny;
    a.nz = S.nz;
    a.px = S.px;
    a.py = S.py;
    a.pz = S.pz;
    a.bx = S.bx;
    a.by = S.by;
    a.bz = S.bz;
    a.rx = S.rx;
    a.ry = S.ry;
    a.rz = S.rz;
    a.sx = S.sx;
    a.sy = S.sy;
    a.sz = S.sz;
    a.ntx = S.ntx;
    a.nty = S.nty;
    a.ntz = S.ntz;
    a.nblocks = S.nblocks;
    a.vbid = gb - S.block0;
    a.vbase = S.block0;
    a.src = base.src + S.base;
    a.own = base.own ? base.own + S.base : nullptr;
    a.dst = base.dst ? base.dst + S.base : nullptr;
    a.aux = base.aux ? base.aux + S.base : nullptr;
    a.ofs = base.ofs + S.ofs_off;
    a.cf = base.cf + S.cf_off;
    a.n_iso = S.n_iso;
    a.n_dm = S.n_dm;
    a.n_gen = S.n_gen;
    a.bmode = S.bmode;
    a.zb[0] = S.zb[0];
    a.zb[1] = S.zb[1];
    a.zb[2] = S.zb[2];
    a.bmap = base.bmap ? base.bmap + S.base : nullptr;
#pragma unroll
    for (int k = 0; k < 9; ++k) a.z[k] = S.z9[k];
    if constexpr (DIP) {
        a.hdip = S.hdip;
        a.dnx = S.dnx;
        a.dny = S.dny;
        a.hstride = S.hstride;
    }
    a.alpha = S.alpha;
    a.den = S.den;
    a.rden = S.rden;
    a.sigma = sigma;
    a.seed = S.seed;
    a.step = step;
    return a;
}

#ifndef MJ_STAGE_WAVES
#define MJ_STAGE_WAVES 6
#endif
template <int MODE, bool NOISE, bool DIP>
__global__ void __launch_bounds__(MJ_BATCH_NTHR, MJ_STAGE_WAVES)
    k_batch_stage(const KArgs base, const BatchSys *__restrict__ sys, const int *__restrict__ blk2sys,
                  const double *__restrict__ sigma, const uint64_t *__restrict__ stepp)
{
    extern __shared__ double lds[];
    const int gb = mj_xcd_remap(blockIdx.x, base.nblocks);
    const int s = blk2sys[gb];
    const KArgs a = sys_args<DIP>(base, sys[s], gb, sigma[s], *stepp + base.step);
    stage_body<MODE, NOISE, BatchGeom>(a, lds);
}

// ---------------------------------------------------------------------------
// Both stages of integration 2 for one tile of a batched system, the
// runtime-geometry form of the single-system fused step.  Regions (tile BX x
// BY x BZ, reach R):
//   S0 = the tile plus 2R on every side: m0 in LDS;
//   S1 = the tile plus R (the per-stage kernels' tile): stage 1 runs on every
//        S1 site from S0, its m1 goes to LDS, the tile's own sites keep D1;
// then stage 2 runs on the tile from S1 and writes m2 out of place (other
// tiles still read m0 around it).  S1 sites outside the tile repeat stage 1 of
// the neighbouring tiles with the same operands and operations, and every
// operation is the per-stage kernels', so m2 equals theirs bit for bit while
// the state crosses HBM once per step instead of three times.  The energies
// the exchange needs come from a field pass before the step.

// Zeeman term at a site (the per-stage body's: uniform, or the site's B map)
__device__ __forceinline__ v3 batch_zeeman(const KArgs &a, bool live, int64_t gi)
{
    double zx = 0.0, zy = 0.0, zz = 0.0;
    if (a.bmode == 1) {
        zx = a.zb[0];
        zy = a.zb[1];
        zz = a.zb[2];
    } else if (a.bmode == 2 && live) {
        const int64_t np = a.np;
        const double bxv = a.bmap[gi], byv = a.bmap[gi + np], bzv = a.bmap[gi + 2 * np];
        zx = (bxv * a.z[0] + byv * a.z[3]) + bzv * a.z[6];
        zy = (bxv * a.z[1] + byv * a.z[4]) + bzv * a.z[7];
        zz = (bxv * a.z[2] + byv * a.z[5]) + bzv * a.z[8];
    }
    return mk3(zx, zy, zz);
}

// lattice site of a region position, wrapped on periodic axes; false off an open edge
__device__ __forceinline__ bool batch_site(const KArgs &a, int gx, int gy, int gz, int64_t &gi)
{
    if (!mj_wrap_idx(gx, a.nx, a.px) || !mj_wrap_idx(gy, a.ny, a.py) || !mj_wrap_idx(gz, a.nz, a.pz)) return false;
    gi = (int64_t)gx + (int64_t)a.nx * ((int64_t)gy + (int64_t)a.ny * gz);
    return true;
}

// stage 1 at one site from S0, the per-stage kernels' arithmetic:
//   integration 2 (HEUN1): X = D1 = LLG(B(m0), m0), m1 = euler(m0, D1, dt);
//   SIB (SIB1): m1 = the midpoint (m0 + cay(m0, a1(B + bt)))/2;
//   Depondt (DEP1): X = b1 = its field of B + bt, m1 = m0 rotated about b1
template <int INTEG>
__device__ __forceinline__ void batch_stage1(const KArgs &a0, const double *L0, int T0, int c0, int64_t gi, v3 bt,
                                             v3 &m0, v3 &X, v3 &m1)
{
    double ax = 0.0, ay = 0.0, az = 0.0;
    m0 = mk3(L0[c0], L0[T0 + c0], L0[2 * T0 + c0]);
    RtGeom::stencil(a0, L0, T0, c0, ax, ay, az);
    const v3 z = batch_zeeman(a0, true, gi);
    const double hx = 0.0, hy = 0.0, hz = 0.0;
    v3 B = mk3(-(ax + z.x) - 2.0 * hx, -(ay + z.y) - 2.0 * hy, -(az + z.z) - 2.0 * hz);
    if constexpr (INTEG == MJ_HEUN) {
        (void)bt;
        X = llg_rhs(m0, B, a0.alpha, a0.den, a0.rden);
        m1 = euler_step0(m0, X, a0.dt);
    } else {
        B = mk3(B.x + bt.x, B.y + bt.y, B.z + bt.z);
        if constexpr (INTEG == MJ_SIB) {
            const v3 av = sib_vec(m0, B, a0.alpha, a0.dt);
            const v3 mt = cayley(m0, av);
            m1 = mk3(0.5 * (m0.x + mt.x), 0.5 * (m0.y + mt.y), 0.5 * (m0.z + mt.z));
            X = mk3(0.0, 0.0, 0.0);
        } else {
            X = dep_field(m0, B, a0.alpha, a0.den, a0.rden);
            m1 = dep_rotate(m0, X, a0.dt);
        }
    }
}

// the thermal field of a site and step (the per-stage kernels' draw)
__device__ __forceinline__ v3 batch_thermal(const KArgs &a, int64_t gi)
{
    const v3 g = normals3(a.seed, a.step, (uint32_t)gi);
    return mk3(a.sigma * g.x, a.sigma * g.y, a.sigma * g.z);
}

// S1 position of ring site k (S1 minus the tile): z slabs, then y slabs of
// the tile's planes, then x slabs of its rows
__device__ __forceinline__ void batch_ring_pos(const KArgs &a, int k, int &sx, int &sy, int &sz)
{
    const int S1X = a.sx, S1Y = a.sy;
    const int nzs = 2 * a.rz * S1X * S1Y, nys = 2 * a.ry * a.bz * S1X;
    if (k < nzs) {
        const int plane = k / (S1X * S1Y), rem = k % (S1X * S1Y);
        sz = plane < a.rz ? plane : plane + a.bz;
        sy = rem / S1X;
        sx = rem % S1X;
    } else if (k < nzs + nys) {
        const int q = k - nzs, row = q / S1X, yy = row % (2 * a.ry);
        sz = a.rz + row / (2 * a.ry);
        sy = yy < a.ry ? yy : yy + a.by;
        sx = q % S1X;
    } else {
        const int q = k - nzs - nys, col = q % (2 * a.rx), row = q / (2 * a.rx);
        sy = a.ry + row % a.by;
        sz = a.rz + row / a.by;
        sx = col < a.rx ? col : col + a.bx;
    }
}

// one tile of integration 2, SIB or Depondt: a.src = m0, a.dst = m2 (the
// system's slices), a.ofs the S1 offsets, ofs0 the S0 offsets; 256 threads,
// one tile site each.  SIB and Depondt add the thermal field in both stages,
// so the ring draws it too.
template <int INTEG>
__device__ __forceinline__ void batch_fused_tile(const KArgs &a, const int *ofs0, int bid, double *lds)
{
    const int BX = a.bx, BY = a.by, BZ = a.bz, RX = a.rx, RY = a.ry, RZ = a.rz;
    const int S0X = BX + 4 * RX, S0Y = BY + 4 * RY, S0Z = BZ + 4 * RZ, T0 = S0X * S0Y * S0Z;
    const int S1X = a.sx, S1Y = a.sy, T1 = S1X * S1Y * a.sz;
    const int nthr = BX * BY * BZ;
    const int64_t np = a.np;
    double *L0 = lds;
    double *L1 = lds + 3 * T0;
    KArgs a0 = a;
    a0.ofs = ofs0;

    const int tx = bid % a.ntx, tr = bid / a.ntx, ty = tr % a.nty, tz = tr / a.nty;
    const int x0 = tx * BX, y0 = ty * BY, z0 = tz * BZ;
    // S0 from HBM (interior tiles without the wrap)
    const bool interior = x0 - 2 * RX >= 0 && x0 + BX + 2 * RX <= a.nx && y0 - 2 * RY >= 0 &&
                          y0 + BY + 2 * RY <= a.ny && z0 - 2 * RZ >= 0 && z0 + BZ + 2 * RZ <= a.nz;
    for (int t = threadIdx.x; t < T0; t += nthr) {
        const int lx = t % S0X, r = t / S0X, ly = r % S0Y, lz = r / S0Y;
        const int gx = x0 + lx - 2 * RX, gy = y0 + ly - 2 * RY, gz = z0 + lz - 2 * RZ;
        int64_t g = 0;
        double vx = 0.0, vy = 0.0, vz = 0.0;
        bool ok = true;
        if (interior) g = (int64_t)gx + (int64_t)a.nx * ((int64_t)gy + (int64_t)a.ny * gz);
        else ok = batch_site(a, gx, gy, gz, g);
        if (ok) {
            vx = a.src[g];
            vy = a.src[g + np];
            vz = a.src[g + 2 * np];
        }
        L0[t] = vx;
        L0[T0 + t] = vy;
        L0[2 * T0 + t] = vz;
    }
    // the own site, and its thermal field while the tile lands
    const int lx = threadIdx.x % BX, ly = (threadIdx.x / BX) % BY, lz = threadIdx.x / (BX * BY);
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    const bool active = gx < a.nx && gy < a.ny && gz < a.nz;
    const int64_t gi = active ? (int64_t)gx + (int64_t)a.nx * ((int64_t)gy + (int64_t)a.ny * gz) : 0;
    v3 bt = mk3(0.0, 0.0, 0.0);
    if (active) bt = batch_thermal(a, gi);
    __syncthreads();

    // stage 1 at the own position (wrapped where the tile passes a periodic edge)
    v3 m0 = mk3(0.0, 0.0, 0.0), D1 = mk3(0.0, 0.0, 0.0);
    {
        const int c1 = (lz + RZ) * S1X * S1Y + (ly + RY) * S1X + (lx + RX);
        const int c0 = (lz + 2 * RZ) * S0X * S0Y + (ly + 2 * RY) * S0X + (lx + 2 * RX);
        int64_t gw = 0;
        v3 m1 = mk3(0.0, 0.0, 0.0);
        if (batch_site(a, gx, gy, gz, gw)) {
            const v3 bw = (INTEG == MJ_HEUN || active) ? bt : batch_thermal(a, gw);
            batch_stage1<INTEG>(a0, L0, T0, c0, gw, bw, m0, D1, m1);
        }
        L1[c1] = m1.x;
        L1[T1 + c1] = m1.y;
        L1[2 * T1 + c1] = m1.z;
    }
    // stage 1 on the ring of S1
    const int nring = T1 - nthr;
    for (int k = threadIdx.x; k < nring; k += nthr) {
        int sx, sy, sz;
        batch_ring_pos(a, k, sx, sy, sz);
        const int c1 = sz * S1X * S1Y + sy * S1X + sx;
        const int c0 = (sz + RZ) * S0X * S0Y + (sy + RY) * S0X + (sx + RX);
        int64_t gw = 0;
        v3 m1 = mk3(0.0, 0.0, 0.0);
        if (batch_site(a, x0 - RX + sx, y0 - RY + sy, z0 - RZ + sz, gw)) {
            v3 mm, dd;
            v3 bw = mk3(0.0, 0.0, 0.0);
            if constexpr (INTEG != MJ_HEUN) bw = batch_thermal(a, gw);
            batch_stage1<INTEG>(a0, L0, T0, c0, gw, bw, mm, dd, m1);
        }
        L1[c1] = m1.x;
        L1[T1 + c1] = m1.y;
        L1[2 * T1 + c1] = m1.z;
    }
    __syncthreads();

    // stage 2 at the own site from S1: HEUN2 with the Langevin term on m0;
    // SIB2 (m0 about the midpoint's field); DEP2 (m0 about (b1 + b2)/2)
    if (active) {
        const int c1 = (lz + RZ) * S1X * S1Y + (ly + RY) * S1X + (lx + RX);
        const v3 m = mk3(L1[c1], L1[T1 + c1], L1[2 * T1 + c1]);
        double ax = 0.0, ay = 0.0, az = 0.0;
        RtGeom::stencil(a, L1, T1, c1, ax, ay, az);
        const v3 z = batch_zeeman(a, true, gi);
        const double hx = 0.0, hy = 0.0, hz = 0.0;
        v3 B = mk3(-(ax + z.x) - 2.0 * hx, -(ay + z.y) - 2.0 * hy, -(az + z.z) - 2.0 * hz);
        v3 o;
        if constexpr (INTEG == MJ_HEUN) {
            const v3 D2 = llg_rhs(m, B, a.alpha, a.den, a.rden);
            const v3 dint = mk3(0.5 * D1.x + 0.5 * D2.x, 0.5 * D1.y + 0.5 * D2.y, 0.5 * D1.z + 0.5 * D2.z);
            o = euler_step(m, dint, llg_rhs(m0, bt, a.alpha, a.den, a.rden), a.dt);
        } else if constexpr (INTEG == MJ_SIB) {
            B = mk3(B.x + bt.x, B.y + bt.y, B.z + bt.z);
            const v3 av = sib_vec(m, B, a.alpha, a.dt);
            o = cayley(m0, av);
        } else {
            B = mk3(B.x + bt.x, B.y + bt.y, B.z + bt.z);
            const v3 b2 = dep_field(m, B, a.alpha, a.den, a.rden);
            const v3 ba = mk3(0.5 * (D1.x + b2.x), 0.5 * (D1.y + b2.y), 0.5 * (D1.z + b2.z));
            o = dep_rotate(m0, ba, a.dt);
        }
        a.dst[gi] = o.x;
        a.dst[gi + np] = o.y;
        a.dst[gi + 2 * np] = o.z;
    }
}

// one step of every system: base.src -> base.dst
#ifndef MJ_FUSED_WAVES
#define MJ_FUSED_WAVES 6
#endif
template <int INTEG>
__global__ void __launch_bounds__(MJ_BATCH_NTHR, MJ_FUSED_WAVES)
    k_batch_fused(const KArgs base, const BatchSys *__restrict__ sys, const int *__restrict__ blk2sys,
                  const double *__restrict__ sigma, const uint64_t *__restrict__ stepp, const int *__restrict__ ofs0)
{
    extern __shared__ double lds[];
    const int gb = mj_xcd_remap(blockIdx.x, base.nblocks);
    const int s = blk2sys[gb];
    const BatchSys &S = sys[s];
    const KArgs a = sys_args(base, S, gb, sigma[s], *stepp + base.step);
    batch_fused_tile<INTEG>(a, ofs0 + S.ofs0_off, gb - S.block0, lds);
}

// per-system energy: the sum of the system's block partials (total, not per
// site), strided over 256 threads then reduced in wave order
__device__ __forceinline__ double sys_energy(const double *part, const BatchSys &S, double *red)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < S.nblocks; i += MJ_BATCH_NTHR) v += part[S.block0 + i];
    return mj_block_sum(v, red, MJ_BATCH_NTHR);
}

__global__ void __launch_bounds__(MJ_BATCH_NTHR) k_batch_esum(const double *part, const BatchSys *sys, double *E)
{
    __shared__ double red[MJ_BATCH_NTHR];
    const double t = sys_energy(part, sys[blockIdx.x], red);
    if (threadIdx.x == 0) E[blockIdx.x] = t;
}

// a load at agent scope: values another workgroup of the same launch wrote
// (before a grid barrier) are read from memory, never from a stale cache line
__device__ __forceinline__ double ld_agent(const double *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// the parallel-tempering group: replicas first .. first+count-1 on a ladder
struct PtArgs {
    double *E;                      // per-system energies of the entering states
    const double *beta, *sig_ladder;
    int *rep_of_temp, *temp_of_rep;
    double *sigma;                  // per-system thermal amplitudes (moved by the exchange)
    unsigned long long *acc, *att;
    int first, count, every;
    uint64_t seed;
};

// the exchange of absolute step `step` (one thread): rungs t, t+1 with t of
// the exchange number's parity swap with probability min(1, exp((b_t - b_t+1)
// (E_a - E_b))), the uniform from Philox (t, exchange, 2) keyed by the PT seed
__device__ __forceinline__ void pt_exchange(const PtArgs &p, uint64_t step)
{
    if (step == 0 || step % (uint64_t)p.every != 0) return;
    const uint64_t e = step / (uint64_t)p.every;
    for (int t = (int)(e & 1); t + 1 < p.count; t += 2) {
        const int ra = p.rep_of_temp[t], rb = p.rep_of_temp[t + 1];
        const double d = (p.beta[t] - p.beta[t + 1]) * (ld_agent(p.E + p.first + ra) - ld_agent(p.E + p.first + rb));
        uint32_t c[4] = {(uint32_t)t, (uint32_t)e, (uint32_t)(e >> 32), 2u};
        philox4x32_10(c, (uint32_t)p.seed, (uint32_t)(p.seed >> 32));
        const double u = (double)((((uint64_t)c[1] << 32) | c[0]) >> 11) * 1.1102230246251565404236316680908203125e-16;
        p.att[t] += 1ull;
        if (d >= 0.0 || u < exp(d)) {
            p.rep_of_temp[t] = rb;
            p.rep_of_temp[t + 1] = ra;
            p.temp_of_rep[ra] = t + 1;
            p.temp_of_rep[rb] = t;
            p.sigma[p.first + ra] = p.sig_ladder[t + 1];
            p.sigma[p.first + rb] = p.sig_ladder[t];
            p.acc[t] += 1ull;
        }
    }
}

__global__ void k_pt_exchange(const PtArgs p, const uint64_t *stepp, uint64_t joff)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    pt_exchange(p, *stepp + joff);   // joff: the step's place in a captured segment
}

// ---------------------------------------------------------------------------
// Persistent form of a segment of integration 2: one launch runs `len` steps
// of every system.  Its workgroups (no more than fit on the device at once)
// walk the batch's tiles stage by stage, in the order the per-stage launches
// give them (the XCD remap of the virtual block index), and meet at a grid
// barrier after each stage and, on exchange steps, after the energy sums and
// after the exchange.  Every tile runs the per-stage kernels' body with the
// same arguments, so the trajectories and the exchange decisions are theirs
// bit for bit.
//
// Grid barrier: arrivals on a counter, a generation number flipped by the
// last arrival.  Each side fences at agent scope (the L2 of one XCD is not
// coherent with the others' within a launch).  A waiting workgroup gives up
// after `spin` sleeps (about 0.25 us each) and raises the abort flag, which
// every other waiter and every later barrier honours, so a launch whose
// workgroups are not all resident ends instead of hanging; the host reports
// the flag as an error.
struct GridBar {
    unsigned count, gen, abort, spare;
};

struct PersistArgs {
    GridBar *bar;
    const BatchSys *sys;
    const int *blk2sys;
    double *A, *Bb, *Db, *epart;
    const int *ofs0;   // fused tiles: the stencil offsets in the S0 layout
    uint64_t *stepp;
    PtArgs pt;   // count 0: no exchange
    int len, nsys;
    unsigned spin;
};

__device__ __forceinline__ bool grid_sync(GridBar *g, unsigned nwg, unsigned &gen, unsigned spin, int *flag)
{
    __builtin_amdgcn_s_waitcnt(0);   // this wave's stores have reached L2
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned g0 = gen;
        int ab = 0;
        if (__hip_atomic_fetch_add(&g->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == nwg - 1u) {
            __hip_atomic_store(&g->count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&g->gen, g0 + 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            for (unsigned k = 0;; ++k) {
                if (__hip_atomic_load(&g->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != g0) break;
                if (__hip_atomic_load(&g->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u || k >= spin) {
                    __hip_atomic_store(&g->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ab = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(8);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        gen = g0 + 1u;
        *flag = ab;
    }
    __syncthreads();
    return *flag == 0;
}

__global__ void __launch_bounds__(MJ_BATCH_NTHR) k_batch_heun_persist(const KArgs base, const PersistArgs p)
{
    extern __shared__ double lds[];
    __shared__ int flag;
    const unsigned nwg = gridDim.x;
    unsigned gen = 0;
    if (threadIdx.x == 0) gen = __hip_atomic_load(&p.bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t step0 = *p.stepp;
    for (int j = 0; j < p.len; ++j) {
        const bool with_e = p.pt.count && j % p.pt.every == 0;
        // stage 1 reads A, writes m1 to Bb and D1 to Db
        for (int v = blockIdx.x; v < base.nblocks; v += (int)nwg) {
            const int gb = mj_xcd_remap(v, base.nblocks);
            const BatchSys &S = p.sys[p.blk2sys[gb]];
            KArgs a = sys_args(base, S, gb, 0.0, step0 + j);
            a.src = p.A + S.base;
            a.dst = p.Bb + S.base;
            a.aux = p.Db + S.base;
            a.energy = with_e ? 1 : 0;
            a.epart = p.epart;
            stage_body<MODE_HEUN1, false, BatchGeom>(a, lds);
            __syncthreads();
        }
        if (!grid_sync(p.bar, nwg, gen, p.spin, &flag)) return;
        if (with_e) {
            for (int s = blockIdx.x; s < p.nsys; s += (int)nwg) {
                const double t = sys_energy(p.epart, p.sys[s], lds);
                if (threadIdx.x == 0) p.pt.E[s] = t;
            }
            if (!grid_sync(p.bar, nwg, gen, p.spin, &flag)) return;
            if (blockIdx.x == 0 && threadIdx.x == 0) pt_exchange(p.pt, step0 + j);
            if (!grid_sync(p.bar, nwg, gen, p.spin, &flag)) return;
        }
        // stage 2 reads Bb, Db and its own site of A, and writes A in place
        for (int v = blockIdx.x; v < base.nblocks; v += (int)nwg) {
            const int gb = mj_xcd_remap(v, base.nblocks);
            const int s = p.blk2sys[gb];
            const BatchSys &S = p.sys[s];
            KArgs a = sys_args(base, S, gb, ld_agent(p.pt.sigma + s), step0 + j);
            a.src = p.Bb + S.base;
            a.own = p.A + S.base;
            a.aux = p.Db + S.base;
            a.dst = p.A + S.base;
            stage_body<MODE_HEUN2, true, BatchGeom>(a, lds);
            __syncthreads();
        }
        if (!grid_sync(p.bar, nwg, gen, p.spin, &flag)) return;
    }
    // every workgroup has read the step counter: advance it for the next launch
    if (blockIdx.x == 0 && threadIdx.x == 0) *p.stepp = step0 + (uint64_t)p.len;
}

// the same with fused tiles: one grid barrier per step (the state alternates
// between A and Bb, `len` is even); on exchange steps a field pass forms the
// entering energies first
__global__ void __launch_bounds__(MJ_BATCH_NTHR) k_batch_heun_persist_fused(const KArgs base, const PersistArgs p)
{
    extern __shared__ double lds[];
    __shared__ int flag;
    const unsigned nwg = gridDim.x;
    unsigned gen = 0;
    if (threadIdx.x == 0) gen = __hip_atomic_load(&p.bar->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint64_t step0 = *p.stepp;
    for (int j = 0; j < p.len; ++j) {
        double *src = (j & 1) ? p.Bb : p.A;
        double *dst = (j & 1) ? p.A : p.Bb;
        if (p.pt.count && j % p.pt.every == 0) {
            for (int v = blockIdx.x; v < base.nblocks; v += (int)nwg) {
                const int gb = mj_xcd_remap(v, base.nblocks);
                const BatchSys &S = p.sys[p.blk2sys[gb]];
                KArgs a = sys_args(base, S, gb, 0.0, step0 + j);
                a.src = src + S.base;
                a.energy = 1;
                a.epart = p.epart;
                stage_body<MODE_FIELD, false, BatchGeom>(a, lds);
                __syncthreads();
            }
            if (!grid_sync(p.bar, nwg, gen, p.spin, &flag)) return;
            for (int s = blockIdx.x; s < p.nsys; s += (int)nwg) {
                const double t = sys_energy(p.epart, p.sys[s], lds);
                if (threadIdx.x == 0) p.pt.E[s] = t;
            }
            if (!grid_sync(p.bar, nwg, gen, p.spin, &flag)) return;
            if (blockIdx.x == 0 && threadIdx.x == 0) pt_exchange(p.pt, step0 + j);
            if (!grid_sync(p.bar, nwg, gen, p.spin, &flag)) return;
        }
        for (int v = blockIdx.x; v < base.nblocks; v += (int)nwg) {
            const int gb = mj_xcd_remap(v, base.nblocks);
            const int s = p.blk2sys[gb];
            const BatchSys &S = p.sys[s];
            KArgs a = sys_args(base, S, gb, ld_agent(p.pt.sigma + s), step0 + j);
            a.src = src + S.base;
            a.dst = dst + S.base;
            batch_fused_tile<MJ_HEUN>(a, p.ofs0 + S.ofs0_off, gb - S.block0, lds);
            __syncthreads();
        }
        if (!grid_sync(p.bar, nwg, gen, p.spin, &flag)) return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *p.stepp = step0 + (uint64_t)p.len;
}

__global__ void k_bump(uint64_t *stepp, uint64_t n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *stepp += n;
}

__global__ void k_batch_soa2aos(const double *in, double *out, int64_t n, int64_t np)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[3 * i] = in[i];
    out[3 * i + 1] = in[i + np];
    out[3 * i + 2] = in[i + 2 * np];
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct HostSys {
    int dim[3], periodic[3];
    int64_t n, base;
    int rx, ry, rz, n_iso, n_dm, n_gen, bmode;
    std::vector<int> ofs3;
    std::vector<double> cf;
    double zb[3];
    double z9[9] = {0};
    std::vector<double> bmap;   // per-site B (SoA, n each) for bmode 2
    double alpha, kT;
    uint64_t seed;
    int bx = 0, by = 0;   // 0: automatic
    std::vector<double> m0;   // initial state (SoA, n each) until uploaded
    mj_ctx *dipc = nullptr;   // dipolar term: a context holding the batch's own copy of the transform (owned)
};

struct mj_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<HostSys> sys;
    int dirty = 1;
    int64_t np = 0;
    int nblocks = 0;
    size_t shm = 0;
    double *A = nullptr, *Bb = nullptr, *Db = nullptr, *epart = nullptr, *E = nullptr, *sigma = nullptr;
    double *bmap = nullptr;   // the systems' per-site B maps, laid out as the state (null: none)
    BatchSys *d_sys = nullptr;
    int *d_blk2sys = nullptr, *d_ofs = nullptr;
    double *d_cf = nullptr;
    uint64_t *d_step = nullptr;
    uint64_t h_step = 0;
    // parallel tempering (one group)
    int pt_first = -1, pt_count = 0, pt_every = 0;
    uint64_t pt_seed = 0;
    std::vector<double> pt_kT;
    double *d_beta = nullptr, *d_sigl = nullptr;
    int *d_rep_of_temp = nullptr, *d_temp_of_rep = nullptr;
    unsigned long long *d_acc = nullptr, *d_att = nullptr;
    // how integration 2's segments launch (MJ_BATCH_*), whether the fused tiles
    // fit in LDS, their LDS; persistent segments: grid barrier, workgroups per
    // kernel, sleeps before a barrier gives up
    int launch = MJ_BATCH_FUSED;
    int has_dip = 0;   // a system has a dipolar term: per-stage launches, its H_fft before each
    int fused_ok = 0;
    size_t shm_fused = 0;
    int *d_ofs0 = nullptr;
    GridBar *d_bar = nullptr;
    int pgrid = 0, pgrid_fused = 0;
    int ran_persist = 0;   // a persistent segment ran since the barrier was last checked
    unsigned spin = 1u << 22;
    // graph of one segment
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int seg = 0;
    double g_timestep = -1.0;
    int integ = MJ_HEUN, g_integ = 0;   // the integration the batch runs; the one sigma/graphs were made for
    // PT assignment kept across a rebuild (empty: start from the identity)
    std::vector<int> k_tor, k_rot;
    std::vector<uint64_t> k_acc, k_att;
};

#define BFAIL(code, msg)    \
    do {                    \
        mj_set_error(msg);  \
        return (code);      \
    } while (0)
#define BTRY(x)                     \
    do {                            \
        int r_ = (x);               \
        if (r_ != MJ_OK) return r_; \
    } while (0)

static const double kHbarB = 6.58211928e-1;

static void bfree(void *p)
{
    if (p) (void)hipFree(p);
}

static void drop_graph(mj_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
}

static void free_dev(mj_batch *b)
{
    drop_graph(b);
    void *ps[] = {b->A, b->Bb, b->Db, b->epart, b->E, b->sigma, b->d_sys, b->d_blk2sys, b->d_ofs, b->d_cf,
                  b->d_step, b->d_beta, b->d_sigl, b->d_rep_of_temp, b->d_temp_of_rep, b->d_acc, b->d_att,
                  b->bmap, b->d_bar, b->d_ofs0};
    for (void *p : ps) bfree(p);
    b->A = b->Bb = b->Db = b->epart = b->E = b->sigma = b->bmap = nullptr;
    b->d_bar = nullptr;
    b->d_ofs0 = nullptr;
    b->d_sys = nullptr;
    b->d_blk2sys = b->d_ofs = nullptr;
    b->d_cf = nullptr;
    b->d_step = nullptr;
    b->d_beta = b->d_sigl = nullptr;
    b->d_rep_of_temp = b->d_temp_of_rep = nullptr;
    b->d_acc = b->d_att = nullptr;
}

// before a rebuild: copy the device states, the step counter and the PT
// assignment to the host so that the rebuilt batch continues from them
static int snapshot(mj_batch *b)
{
    if (b->dirty) return MJ_OK;
    MJ_HIP_CHECK(hipSetDevice(b->device));
    MJ_HIP_CHECK(hipStreamSynchronize(b->stream));
    for (auto &q : b->sys) {
        q.m0.resize(3 * q.n);
        for (int a = 0; a < 3; ++a)
            MJ_HIP_CHECK(hipMemcpy(q.m0.data() + a * q.n, b->A + a * b->np + q.base, sizeof(double) * q.n,
                                   hipMemcpyDeviceToHost));
    }
    if (b->pt_count) {
        const int R = b->pt_count;
        b->k_tor.resize(R);
        b->k_rot.resize(R);
        b->k_acc.resize(R);
        b->k_att.resize(R);
        MJ_HIP_CHECK(hipMemcpy(b->k_tor.data(), b->d_temp_of_rep, sizeof(int) * R, hipMemcpyDeviceToHost));
        MJ_HIP_CHECK(hipMemcpy(b->k_rot.data(), b->d_rep_of_temp, sizeof(int) * R, hipMemcpyDeviceToHost));
        MJ_HIP_CHECK(hipMemcpy(b->k_acc.data(), b->d_acc, sizeof(uint64_t) * R, hipMemcpyDeviceToHost));
        MJ_HIP_CHECK(hipMemcpy(b->k_att.data(), b->d_att, sizeof(uint64_t) * R, hipMemcpyDeviceToHost));
    }
    return MJ_OK;
}

extern "C" int mj_batch_create(mj_batch **out, int device)
{
    if (!out) BFAIL(MJ_EINVAL, "mj_batch_create: null argument");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        BFAIL(MJ_EINVAL, "mj_batch_create: no such device");
    MJ_HIP_CHECK(hipSetDevice(device));
    mj_batch *b = new mj_batch();
    b->device = device;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        delete b;
        BFAIL(MJ_EHIP, "mj_batch_create: stream creation failed");
    }
    *out = b;
    return MJ_OK;
}

extern "C" int mj_batch_destroy(mj_batch **pb)
{
    if (!pb || !*pb) return MJ_OK;
    mj_batch *b = *pb;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    free_dev(b);
    for (auto &s : b->sys)
        if (s.dipc) {
            dip_destroy(s.dipc);
            delete s.dipc;
            s.dipc = nullptr;
        }
    (void)hipStreamDestroy(b->stream);
    delete b;
    *pb = nullptr;
    return MJ_OK;
}

extern "C" int mj_batch_add(mj_batch *b, mj_ctx *c, double damping, double kT, uint64_t seed, int *index)
{
    if (!b || !c) BFAIL(MJ_EINVAL, "mj_batch_add: null argument");
    if (!(damping >= 0.0) || !(kT >= 0.0)) BFAIL(MJ_EINVAL, "mj_batch_add: damping and kT must be >= 0");
    if (c->device != b->device) BFAIL(MJ_EINVAL, "mj_batch_add: context on another device");
    if (!c->state_set) BFAIL(MJ_ESTATE, "mj_batch_add: the context has no state");
    BTRY(mj_finalize(c));
    if (c->mm_generic) BFAIL(MJ_ENOTSUP, "mj_batch_add: only translation-invariant (stencil) Hamiltonians");
    HostSys s;
    for (int d = 0; d < 3; ++d) {
        s.dim[d] = c->dim[d];
        s.periodic[d] = c->periodic[d];
    }
    s.n = c->n;
    s.rx = c->rx;
    s.ry = c->ry;
    s.rz = c->rz;
    s.n_iso = c->n_iso;
    s.n_dm = c->n_dm;
    s.n_gen = c->n_gen;
    s.bmode = c->bmode;
    s.ofs3 = c->h_ofs3;
    s.cf = c->h_cf;
    for (int a = 0; a < 3; ++a) s.zb[a] = c->zb[a];
    if (c->bmode == 2) {   // per-site B: the map (AoS on the context's host side) and the Zeeman block
        for (int k = 0; k < 9; ++k) s.z9[k] = c->zblk[k];
        s.bmap.resize(3 * (size_t)c->n);
        for (int64_t i = 0; i < c->n; ++i)
            for (int a = 0; a < 3; ++a) s.bmap[a * c->n + i] = c->bvec[3 * i + a];
    }
    s.alpha = damping;
    s.kT = kT;
    s.seed = seed;
    s.m0.resize(3 * c->n);
    MJ_HIP_CHECK(hipSetDevice(b->device));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    for (int a = 0; a < 3; ++a)
        MJ_HIP_CHECK(hipMemcpy(s.m0.data() + a * c->n, c->A + a * c->np, sizeof(double) * c->n, hipMemcpyDeviceToHost));
    BTRY(snapshot(b));
    if (c->dip) {
        // the dipolar transform: the batch's own plans and buffers on its stream, the context's operator
        mj_ctx *pc = new mj_ctx();
        pc->device = b->device;
        for (int d = 0; d < 3; ++d) {
            pc->dim[d] = c->dim[d];
            pc->periodic[d] = c->periodic[d];
        }
        pc->n = c->n;
        pc->stream = b->stream;
        const int rc = dip_clone(pc, c);
        if (rc != MJ_OK) {
            if (pc->dip) dip_destroy(pc);
            delete pc;
            return rc;
        }
        s.dipc = pc;
        b->has_dip = 1;
    }
    b->sys.push_back(std::move(s));
    b->dirty = 1;
    if (index) *index = (int)b->sys.size() - 1;
    return MJ_OK;
}

extern "C" int mj_batch_set_pt(mj_batch *b, int first, int count, const double *kT_ladder, int every, uint64_t seed)
{
    if (!b || !kT_ladder) BFAIL(MJ_EINVAL, "mj_batch_set_pt: null argument");
    if (first < 0 || count < 2 || first + count > (int)b->sys.size() || every < 1)
        BFAIL(MJ_EINVAL, "mj_batch_set_pt: replicas must be existing systems (count >= 2), every >= 1");
    const HostSys &r0 = b->sys[first];
    for (int k = 0; k < count; ++k) {
        const HostSys &r = b->sys[first + k];
        if (!(kT_ladder[k] > 0.0)) BFAIL(MJ_EINVAL, "mj_batch_set_pt: temperatures must be > 0");
        if (r.dipc) BFAIL(MJ_ENOTSUP, "mj_batch_set_pt: no dipolar replicas");
        bool same = r.n == r0.n && r.alpha == r0.alpha && r.cf == r0.cf && r.ofs3 == r0.ofs3 && r.bmode == r0.bmode &&
                    r.zb[0] == r0.zb[0] && r.zb[1] == r0.zb[1] && r.zb[2] == r0.zb[2] && r.bmap == r0.bmap &&
                    std::equal(r.z9, r.z9 + 9, r0.z9);
        for (int d = 0; d < 3; ++d) same = same && r.dim[d] == r0.dim[d] && r.periodic[d] == r0.periodic[d];
        if (!same) BFAIL(MJ_EINVAL, "mj_batch_set_pt: replicas must share lattice, Hamiltonian and damping");
    }
    BTRY(snapshot(b));
    b->pt_first = first;
    b->pt_count = count;
    b->pt_every = every;
    b->pt_seed = seed;
    b->pt_kT.assign(kT_ladder, kT_ladder + count);
    b->k_tor.clear();
    b->k_rot.clear();
    b->k_acc.clear();
    b->k_att.clear();
    for (int k = 0; k < count; ++k) b->sys[first + k].kT = kT_ladder[k];   // replica k starts at temperature k
    b->dirty = 1;
    return MJ_OK;
}

static bool tile_ok(int bx, int by) { return bx >= 1 && by >= 1 && bx * by == MJ_BATCH_NTHR; }

extern "C" int mj_batch_set_tiles(mj_batch *b, const int *tiles)
{
    if (!b || !tiles) BFAIL(MJ_EINVAL, "mj_batch_set_tiles: null argument");
    for (size_t s = 0; s < b->sys.size(); ++s) {
        int bx = tiles[2 * s], by = tiles[2 * s + 1];
        if (!(bx == 0 && by == 0) && !tile_ok(bx, by))
            BFAIL(MJ_EINVAL, "mj_batch_set_tiles: a tile must hold 256 sites (bx * by == 256) or be 0 0");
    }
    BTRY(snapshot(b));
    for (size_t s = 0; s < b->sys.size(); ++s) {
        b->sys[s].bx = tiles[2 * s];
        b->sys[s].by = tiles[2 * s + 1];
    }
    b->dirty = 1;
    return MJ_OK;
}

static void auto_tile(const HostSys &s, int &bx, int &by)
{
    if (s.bx && s.by) {
        bx = s.bx;
        by = s.by;
        return;
    }
    bx = 32;
    by = 8;
    if (s.dim[0] < 32) {
        bx = 16;
        by = 16;
    }
}

// (re)build the device tables and arrays; the states come from HostSys::m0
static int batch_finalize(mj_batch *b)
{
    if (!b->dirty) return MJ_OK;
    if (b->sys.empty()) BFAIL(MJ_ESTATE, "mj_batch: no system added");
    MJ_HIP_CHECK(hipSetDevice(b->device));
    MJ_HIP_CHECK(hipStreamSynchronize(b->stream));
    free_dev(b);
    std::vector<BatchSys> tab(b->sys.size());
    std::vector<int> ofs, ofs0, blk2sys;
    size_t shm_fused = 0;
    std::vector<double> cf;
    int64_t np = 0;
    int nb = 0;
    size_t shm = MJ_BATCH_NTHR * sizeof(double);
    for (size_t q = 0; q < b->sys.size(); ++q) {
        HostSys &s = b->sys[q];
        BatchSys &t = tab[q];
        memset(&t, 0, sizeof t);
        int bx, by;
        auto_tile(s, bx, by);
        t.nx = s.dim[0];
        t.ny = s.dim[1];
        t.nz = s.dim[2];
        t.px = s.periodic[0];
        t.py = s.periodic[1];
        t.pz = s.periodic[2];
        t.bx = bx;
        t.by = by;
        t.bz = 1;
        t.rx = s.rx;
        t.ry = s.ry;
        t.rz = s.rz;
        t.sx = bx + 2 * s.rx;
        t.sy = by + 2 * s.ry;
        t.sz = 1 + 2 * s.rz;
        t.ntx = (t.nx + bx - 1) / bx;
        t.nty = (t.ny + by - 1) / by;
        t.ntz = t.nz;
        t.nblocks = t.ntx * t.nty * t.ntz;
        t.block0 = nb;
        nb += t.nblocks;
        for (int k = 0; k < t.nblocks; ++k) blk2sys.push_back((int)q);
        t.n_iso = s.n_iso;
        t.n_dm = s.n_dm;
        t.n_gen = s.n_gen;
        t.ofs_off = (int)ofs.size();
        for (size_t k = 0; k < s.ofs3.size() / 3; ++k)
            ofs.push_back(s.ofs3[3 * k + 2] * t.sx * t.sy + s.ofs3[3 * k + 1] * t.sx + s.ofs3[3 * k]);
        // the fused step's S0 = the tile plus 2R: its offsets, and the LDS of S0 and S1
        const int s0x = bx + 4 * s.rx, s0y = by + 4 * s.ry, s0z = 1 + 4 * s.rz;
        t.ofs0_off = (int)ofs0.size();
        for (size_t k = 0; k < s.ofs3.size() / 3; ++k)
            ofs0.push_back(s.ofs3[3 * k + 2] * s0x * s0y + s.ofs3[3 * k + 1] * s0x + s.ofs3[3 * k]);
        shm_fused = std::max(shm_fused, (size_t)3 * ((size_t)s0x * s0y * s0z + (size_t)t.sx * t.sy * t.sz) * sizeof(double));
        t.cf_off = (int)cf.size();
        cf.insert(cf.end(), s.cf.begin(), s.cf.end());
        t.bmode = s.bmode;
        for (int a = 0; a < 3; ++a) t.zb[a] = s.zb[a];
        for (int k = 0; k < 9; ++k) t.z9[k] = s.z9[k];
        t.alpha = s.alpha;
        t.den = 1.0 + s.alpha * s.alpha;
        t.rden = 1.0 / t.den;
        t.seed = s.seed;
        s.base = np;
        t.base = np;
        np += (s.n + 63) / 64 * 64;
        t.hdip = nullptr;
        t.dnx = t.dny = 0;
        t.hstride = 0;
        if (s.dipc) dip_geometry(s.dipc, &t.hdip, &t.dnx, &t.dny, &t.hstride);
        size_t lds = (size_t)3 * t.sx * t.sy * t.sz * sizeof(double);
        shm = std::max(shm, lds);
    }
    if (shm > 65536) BFAIL(MJ_ENOTSUP, "mj_batch: a system's tile with its halo does not fit in LDS");
    b->np = np;
    b->nblocks = nb;
    b->shm = shm;
    for (auto &s : b->sys)
        if (s.dipc) s.dipc->np = np;   // the batch's SoA stride: the transform reads the system's slice with it
    b->shm_fused = std::max(shm, shm_fused);
    b->fused_ok = b->shm_fused <= 65536;
    // the persistent kernel's grid: as many workgroups as are resident at once
    // (a multiple of 8, so that the XCD of a virtual block is its workgroup's)
    {
        int occ = 0, occf = 0, cus = 0;
        MJ_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k_batch_heun_persist, MJ_BATCH_NTHR, shm));
        if (b->fused_ok)
            MJ_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occf, k_batch_heun_persist_fused, MJ_BATCH_NTHR,
                                                                      b->shm_fused));
        MJ_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
        auto grid = [&](int o) {
            int g = std::min(nb, std::max(o, 0) * cus);
            if (g >= 8) g -= g % 8;
            return g;
        };
        b->pgrid = grid(occ);
        b->pgrid_fused = grid(occf);
    }
    const size_t bytes = sizeof(double) * 3 * (size_t)np;
    MJ_HIP_CHECK(hipMalloc(&b->A, bytes));
    MJ_HIP_CHECK(hipMalloc(&b->Bb, bytes));
    MJ_HIP_CHECK(hipMalloc(&b->Db, bytes));
    MJ_HIP_CHECK(hipMalloc(&b->epart, sizeof(double) * nb));
    MJ_HIP_CHECK(hipMalloc(&b->E, sizeof(double) * tab.size()));
    MJ_HIP_CHECK(hipMalloc(&b->sigma, sizeof(double) * tab.size()));
    MJ_HIP_CHECK(hipMalloc(&b->d_sys, sizeof(BatchSys) * tab.size()));
    MJ_HIP_CHECK(hipMalloc(&b->d_blk2sys, sizeof(int) * blk2sys.size()));
    MJ_HIP_CHECK(hipMalloc(&b->d_ofs, sizeof(int) * std::max<size_t>(1, ofs.size())));
    MJ_HIP_CHECK(hipMalloc(&b->d_cf, sizeof(double) * std::max<size_t>(1, cf.size())));
    MJ_HIP_CHECK(hipMalloc(&b->d_ofs0, sizeof(int) * std::max<size_t>(1, ofs0.size())));
    if (!ofs0.empty())
        MJ_HIP_CHECK(hipMemcpy(b->d_ofs0, ofs0.data(), sizeof(int) * ofs0.size(), hipMemcpyHostToDevice));
    MJ_HIP_CHECK(hipMalloc(&b->d_step, sizeof(uint64_t)));
    MJ_HIP_CHECK(hipMemcpy(b->d_sys, tab.data(), sizeof(BatchSys) * tab.size(), hipMemcpyHostToDevice));
    MJ_HIP_CHECK(hipMemcpy(b->d_blk2sys, blk2sys.data(), sizeof(int) * blk2sys.size(), hipMemcpyHostToDevice));
    if (!ofs.empty()) MJ_HIP_CHECK(hipMemcpy(b->d_ofs, ofs.data(), sizeof(int) * ofs.size(), hipMemcpyHostToDevice));
    if (!cf.empty()) MJ_HIP_CHECK(hipMemcpy(b->d_cf, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice));
    MJ_HIP_CHECK(hipMemcpy(b->d_step, &b->h_step, sizeof(uint64_t), hipMemcpyHostToDevice));
    {
        const GridBar z0 = {0u, 0u, 0u, 0u};
        MJ_HIP_CHECK(hipMalloc(&b->d_bar, sizeof(GridBar)));
        MJ_HIP_CHECK(hipMemcpy(b->d_bar, &z0, sizeof(GridBar), hipMemcpyHostToDevice));
    }
    {
        std::vector<double> z(tab.size(), 0.0);
        MJ_HIP_CHECK(hipMemcpy(b->sigma, z.data(), sizeof(double) * z.size(), hipMemcpyHostToDevice));
    }
    // states (padding zero)
    std::vector<double> soa((size_t)3 * np, 0.0);
    for (auto &s : b->sys)
        for (int a = 0; a < 3; ++a)
            std::copy(s.m0.begin() + a * s.n, s.m0.begin() + (a + 1) * s.n, soa.begin() + a * np + s.base);
    MJ_HIP_CHECK(hipMemcpy(b->A, soa.data(), bytes, hipMemcpyHostToDevice));
    std::fill(soa.begin(), soa.end(), 0.0);
    MJ_HIP_CHECK(hipMemcpy(b->Bb, soa.data(), bytes, hipMemcpyHostToDevice));
    MJ_HIP_CHECK(hipMemcpy(b->Db, soa.data(), bytes, hipMemcpyHostToDevice));
    for (auto &s : b->sys) std::vector<double>().swap(s.m0);
    // per-site B maps at the systems' places (kept on the host for rebuilds)
    bool any_map = false;
    for (auto &s : b->sys) any_map = any_map || s.bmode == 2;
    if (any_map) {
        for (auto &s : b->sys)
            if (s.bmode == 2)
                for (int a = 0; a < 3; ++a)
                    std::copy(s.bmap.begin() + a * s.n, s.bmap.begin() + (a + 1) * s.n, soa.begin() + a * np + s.base);
        MJ_HIP_CHECK(hipMalloc(&b->bmap, bytes));
        MJ_HIP_CHECK(hipMemcpy(b->bmap, soa.data(), bytes, hipMemcpyHostToDevice));
    }
    // parallel tempering tables
    if (b->pt_count) {
        const int R = b->pt_count;
        std::vector<int> id(R);
        for (int k = 0; k < R; ++k) id[k] = k;
        MJ_HIP_CHECK(hipMalloc(&b->d_beta, sizeof(double) * R));
        MJ_HIP_CHECK(hipMalloc(&b->d_sigl, sizeof(double) * R));
        MJ_HIP_CHECK(hipMalloc(&b->d_rep_of_temp, sizeof(int) * R));
        MJ_HIP_CHECK(hipMalloc(&b->d_temp_of_rep, sizeof(int) * R));
        MJ_HIP_CHECK(hipMalloc(&b->d_acc, sizeof(unsigned long long) * R));
        MJ_HIP_CHECK(hipMalloc(&b->d_att, sizeof(unsigned long long) * R));
        std::vector<double> beta(R);
        for (int k = 0; k < R; ++k) beta[k] = 1.0 / b->pt_kT[k];
        MJ_HIP_CHECK(hipMemcpy(b->d_beta, beta.data(), sizeof(double) * R, hipMemcpyHostToDevice));
        const bool kept = (int)b->k_tor.size() == R;
        std::vector<uint64_t> z(R, 0);
        MJ_HIP_CHECK(hipMemcpy(b->d_rep_of_temp, kept ? b->k_rot.data() : id.data(), sizeof(int) * R,
                               hipMemcpyHostToDevice));
        MJ_HIP_CHECK(hipMemcpy(b->d_temp_of_rep, kept ? b->k_tor.data() : id.data(), sizeof(int) * R,
                               hipMemcpyHostToDevice));
        MJ_HIP_CHECK(hipMemcpy(b->d_acc, kept ? b->k_acc.data() : z.data(), sizeof(uint64_t) * R, hipMemcpyHostToDevice));
        MJ_HIP_CHECK(hipMemcpy(b->d_att, kept ? b->k_att.data() : z.data(), sizeof(uint64_t) * R, hipMemcpyHostToDevice));
    }
    b->g_timestep = -1.0;
    b->dirty = 0;
    return MJ_OK;
}

// stage timestep of the batch's integration: timestep / N_loop for
// integrations 1 and 2 (the reference's get_dt_mode), the whole step for SIB
// and Depondt (as the single-system run_steps forms it)
static double batch_dt(const mj_batch *b, double timestep)
{
    return b->integ == MJ_HEUN ? timestep / 2.0 : timestep;
}

// per-system thermal amplitudes for a time step, the single-system formulas:
// sqrt(damping * kT * hbar / dt_stage) for integrations 1 and 2,
// sqrt(2 * damping * kT * hbar / dt) for SIB and Depondt
static int set_timestep(mj_batch *b, double timestep)
{
    if (b->g_timestep == timestep && b->g_integ == b->integ) return MJ_OK;
    drop_graph(b);
    const double dt_stage = batch_dt(b, timestep);
    const double f = (b->integ == MJ_SIB || b->integ == MJ_DEPONDT) ? 2.0 : 1.0;
    std::vector<double> sig(b->sys.size());
    // the temperature a replica holds now (after exchanges) comes from the device
    std::vector<int> tor;
    if (b->pt_count) {
        tor.resize(b->pt_count);
        MJ_HIP_CHECK(hipMemcpy(tor.data(), b->d_temp_of_rep, sizeof(int) * b->pt_count, hipMemcpyDeviceToHost));
    }
    for (size_t q = 0; q < b->sys.size(); ++q) {
        double kT = b->sys[q].kT;
        int r = (int)q - b->pt_first;
        if (b->pt_count && r >= 0 && r < b->pt_count) kT = b->pt_kT[tor[r]];
        const double alpha = b->sys[q].alpha;
        sig[q] = (kT * alpha > 1.0e-10) ? sqrt(f * alpha * kT * kHbarB / dt_stage) : 0.0;
    }
    MJ_HIP_CHECK(hipMemcpy(b->sigma, sig.data(), sizeof(double) * sig.size(), hipMemcpyHostToDevice));
    if (b->pt_count) {
        std::vector<double> sl(b->pt_count);
        const double alpha = b->sys[b->pt_first].alpha;
        for (int k = 0; k < b->pt_count; ++k) sl[k] = sqrt(f * alpha * b->pt_kT[k] * kHbarB / dt_stage);
        MJ_HIP_CHECK(hipMemcpy(b->d_sigl, sl.data(), sizeof(double) * sl.size(), hipMemcpyHostToDevice));
    }
    b->g_timestep = timestep;
    b->g_integ = b->integ;
    return MJ_OK;
}

static int blaunch_err(const char *what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string(what) + ": " + hipGetErrorString(e));
        return MJ_EHIP;
    }
    return MJ_OK;
}

// one stage kernel over the whole batch
template <int MODE, bool NOISE>
static void launch_stage(const mj_batch *b, const KArgs &a)
{
    if (b->has_dip)
        hipLaunchKernelGGL((k_batch_stage<MODE, NOISE, true>), dim3(b->nblocks), dim3(MJ_BATCH_NTHR), b->shm, b->stream,
                           a, b->d_sys, b->d_blk2sys, b->sigma, b->d_step);
    else
        hipLaunchKernelGGL((k_batch_stage<MODE, NOISE, false>), dim3(b->nblocks), dim3(MJ_BATCH_NTHR), b->shm, b->stream,
                           a, b->d_sys, b->d_blk2sys, b->sigma, b->d_step);
}

static void batch_base(const mj_batch *b, KArgs &a, double timestep)
{
    memset(&a, 0, sizeof a);
    a.np = b->np;
    a.nblocks = b->nblocks;   // the batch's workgroup count (for the XCD remap)
    a.use_stencil = 1;
    a.ofs = b->d_ofs;
    a.cf = b->d_cf;
    a.bmap = b->bmap;
    a.dt = batch_dt(b, timestep);
}

// energies of the entering states per system, then the replica exchange of
// step j of the segment
// H_fft of every dipolar system for a stage evaluated on `src` (into each
// system's own grid, which the stage kernels read through BatchSys::hdip)
static int dip_fields(mj_batch *b, const double *src)
{
    if (!b->has_dip) return MJ_OK;
    for (auto &s : b->sys)
        if (s.dipc) {
            KArgs k;
            memset(&k, 0, sizeof k);
            BTRY(dip_apply(s.dipc, src + s.base, &k));
        }
    return MJ_OK;
}

static PtArgs pt_args(const mj_batch *b)
{
    PtArgs p;
    p.E = b->E;
    p.beta = b->d_beta;
    p.sig_ladder = b->d_sigl;
    p.rep_of_temp = b->d_rep_of_temp;
    p.temp_of_rep = b->d_temp_of_rep;
    p.sigma = b->sigma;
    p.acc = b->d_acc;
    p.att = b->d_att;
    p.first = b->pt_first;
    p.count = b->pt_count;
    p.every = b->pt_every > 0 ? b->pt_every : 1;
    p.seed = b->pt_seed;
    return p;
}

static int enqueue_exchange(mj_batch *b, uint64_t j)
{
    hipLaunchKernelGGL(k_batch_esum, dim3((unsigned)b->sys.size()), dim3(MJ_BATCH_NTHR), 0, b->stream, b->epart,
                       b->d_sys, b->E);
    BTRY(blaunch_err("batched energy"));
    hipLaunchKernelGGL(k_pt_exchange, dim3(1), dim3(64), 0, b->stream, pt_args(b), (const uint64_t *)b->d_step, j);
    return blaunch_err("replica exchange");
}

// one step of integration 1, 5 or 6 for every system, the launches of the
// single-system path (run_steps).  Their first stage carries the thermal
// field, so on exchange steps a field pass forms the entering energies and
// the exchange (which moves the amplitudes) runs before it.  Integration 1
// writes out of place: inside a captured segment the state buffers alternate
// (segments have an even number of steps), an eager step copies it back.
static int enqueue_other(mj_batch *b, const KArgs &a, bool with_e, bool alternate)
{
    BTRY(dip_fields(b, b->A));
    if (with_e) {
        KArgs e = a;
        e.src = b->A;
        e.energy = 1;
        e.epart = b->epart;
        launch_stage<MODE_FIELD, false>(b, e);
        BTRY(blaunch_err("batched energy pass"));
        BTRY(enqueue_exchange(b, a.step));
    }
    if (b->integ == MJ_EULER) {
        KArgs s = a;
        s.src = b->A;
        s.dst = b->Bb;
        launch_stage<MODE_EULER, true>(b, s);
        BTRY(blaunch_err("batched euler stage"));
        if (alternate) {
            std::swap(b->A, b->Bb);
            return MJ_OK;
        }
        MJ_HIP_CHECK(hipMemcpyAsync(b->A, b->Bb, sizeof(double) * 3 * (size_t)b->np, hipMemcpyDeviceToDevice,
                                    b->stream));
        return MJ_OK;
    }
    KArgs s1 = a;
    s1.src = b->A;
    s1.dst = b->Bb;
    s1.aux = b->Db;
    KArgs s2 = a;
    s2.src = b->Bb;
    s2.own = b->A;
    s2.aux = b->Db;
    s2.dst = b->A;   // own-site read then write: safe in place
    if (b->integ == MJ_SIB) {
        launch_stage<MODE_SIB1, true>(b, s1);
        BTRY(blaunch_err("batched SIB stage 1"));
        BTRY(dip_fields(b, b->Bb));
        launch_stage<MODE_SIB2, true>(b, s2);
        return blaunch_err("batched SIB stage 2");
    }
    launch_stage<MODE_DEP1, true>(b, s1);
    BTRY(blaunch_err("batched Depondt stage 1"));
    BTRY(dip_fields(b, b->Bb));
    launch_stage<MODE_DEP2, true>(b, s2);
    return blaunch_err("batched Depondt stage 2");
}

// one step for every system; with_e: energies of the entering states are
// reduced per system and the PT exchange runs before the first stage that
// carries the thermal field (stage 2 of Heun)
static int enqueue_step(mj_batch *b, double timestep, uint64_t j, bool with_e, bool alternate = false)
{
    KArgs a;
    batch_base(b, a, timestep);
    a.step = j;
    if (b->integ != MJ_HEUN) return enqueue_other(b, a, with_e, alternate);
    KArgs s1 = a;
    s1.src = b->A;
    s1.dst = b->Bb;
    s1.aux = b->Db;
    if (with_e) {
        s1.energy = 1;
        s1.epart = b->epart;
    }
    BTRY(dip_fields(b, b->A));
    launch_stage<MODE_HEUN1, false>(b, s1);
    BTRY(blaunch_err("batched stage 1"));
    if (with_e) BTRY(enqueue_exchange(b, j));
    BTRY(dip_fields(b, b->Bb));
    KArgs s2 = a;
    s2.src = b->Bb;
    s2.own = b->A;
    s2.aux = b->Db;
    s2.dst = b->A;
    launch_stage<MODE_HEUN2, true>(b, s2);
    return blaunch_err("batched stage 2");
}

// how integration 2's segments run: the requested form, or its per-stage
// counterpart when a system's fused tile does not fit in LDS
static int eff_launch(const mj_batch *b)
{
    if (b->has_dip || b->integ == MJ_EULER) return MJ_BATCH_STAGES;
    if (b->integ != MJ_HEUN)   // SIB, Depondt: the fused step unless one launch per stage is asked for
        return (b->fused_ok && b->launch != MJ_BATCH_STAGES) ? MJ_BATCH_FUSED : MJ_BATCH_STAGES;
    if (b->fused_ok) return b->launch;
    return b->launch == MJ_BATCH_FUSED ? MJ_BATCH_STAGES
                                       : (b->launch == MJ_BATCH_PERSISTENT_FUSED ? MJ_BATCH_PERSISTENT : b->launch);
}

// `len` steps starting at an absolute step that is a multiple of the segment
static int enqueue_segment(mj_batch *b, double timestep, int len)
{
    const int mode = eff_launch(b);
    if (mode == MJ_BATCH_PERSISTENT || mode == MJ_BATCH_PERSISTENT_FUSED) {
        const bool fused = mode == MJ_BATCH_PERSISTENT_FUSED;
        const int grid = fused ? b->pgrid_fused : b->pgrid;
        if (grid < 1) BFAIL(MJ_ENOTSUP, "mj_batch: the persistent kernel does not fit on the device");
        KArgs a;
        batch_base(b, a, timestep);
        PersistArgs p;
        p.bar = b->d_bar;
        p.sys = b->d_sys;
        p.blk2sys = b->d_blk2sys;
        p.A = b->A;
        p.Bb = b->Bb;
        p.Db = b->Db;
        p.epart = b->epart;
        p.stepp = b->d_step;
        p.pt = pt_args(b);
        p.len = len;
        p.nsys = (int)b->sys.size();
        p.spin = b->spin;
        p.ofs0 = b->d_ofs0;
        if (fused)
            hipLaunchKernelGGL(k_batch_heun_persist_fused, dim3((unsigned)grid), dim3(MJ_BATCH_NTHR), b->shm_fused,
                               b->stream, a, p);
        else
            hipLaunchKernelGGL(k_batch_heun_persist, dim3((unsigned)grid), dim3(MJ_BATCH_NTHR), b->shm, b->stream, a, p);
        return blaunch_err("persistent batched segment");
    }
    if (mode == MJ_BATCH_FUSED) {
        // one fused launch per step, the state alternating between A and Bb
        // (`len` is even); a field pass forms the energies an exchange needs
        for (int j = 0; j < len; ++j) {
            KArgs a;
            batch_base(b, a, timestep);
            a.step = (uint64_t)j;
            a.src = (j & 1) ? b->Bb : b->A;
            a.dst = (j & 1) ? b->A : b->Bb;
            if (b->pt_count && j % b->pt_every == 0) {
                KArgs e = a;
                e.dst = nullptr;
                e.energy = 1;
                e.epart = b->epart;
                launch_stage<MODE_FIELD, false>(b, e);
                BTRY(blaunch_err("batched energy pass"));
                BTRY(enqueue_exchange(b, (uint64_t)j));
            }
            if (b->integ == MJ_SIB)
                hipLaunchKernelGGL(k_batch_fused<MJ_SIB>, dim3(b->nblocks), dim3(MJ_BATCH_NTHR), b->shm_fused,
                                   b->stream, a, b->d_sys, b->d_blk2sys, b->sigma, b->d_step, b->d_ofs0);
            else if (b->integ == MJ_DEPONDT)
                hipLaunchKernelGGL(k_batch_fused<MJ_DEPONDT>, dim3(b->nblocks), dim3(MJ_BATCH_NTHR), b->shm_fused,
                                   b->stream, a, b->d_sys, b->d_blk2sys, b->sigma, b->d_step, b->d_ofs0);
            else
                hipLaunchKernelGGL(k_batch_fused<MJ_HEUN>, dim3(b->nblocks), dim3(MJ_BATCH_NTHR), b->shm_fused,
                                   b->stream, a, b->d_sys, b->d_blk2sys, b->sigma, b->d_step, b->d_ofs0);
            BTRY(blaunch_err("batched fused step"));
        }
        hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, b->stream, b->d_step, (uint64_t)len);
        return blaunch_err("step counter");
    }
    for (int j = 0; j < len; ++j) {
        // exchanges happen at absolute steps that are multiples of `every` (the
        // kernel skips step 0); a segment starts on such a step and holds one
        // or two exchange intervals
        const bool with_e = b->pt_count && j % b->pt_every == 0;
        BTRY(enqueue_step(b, timestep, (uint64_t)j, with_e, true));
    }
    hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, b->stream, b->d_step, (uint64_t)len);
    return blaunch_err("step counter");
}

static int batch_run(mj_batch *b, int nsteps, double timestep)
{
    BTRY(batch_finalize(b));
    BTRY(set_timestep(b, timestep));
    // integration 1 and the fused forms of integration 2 alternate the state
    // buffers within a segment: an even number of steps brings the state back to A
    const int seg0 = b->pt_count ? b->pt_every : 50;
    const int mode = eff_launch(b);
    const bool alt = b->integ == MJ_EULER || mode == MJ_BATCH_FUSED || mode == MJ_BATCH_PERSISTENT_FUSED;
    const int seg = (alt && seg0 % 2) ? 2 * seg0 : seg0;
    if (b->seg != seg) drop_graph(b);
    b->seg = seg;
    while (nsteps > 0) {
        // dipolar systems: every step eager (the transforms are not captured)
        if (!b->has_dip && b->h_step % (uint64_t)seg == 0 && nsteps >= seg) {
            if (!b->gexec) {
                MJ_HIP_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
                int rc = enqueue_segment(b, timestep, seg);
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamEndCapture(b->stream, &g);
                if (rc != MJ_OK) {
                    if (g) (void)hipGraphDestroy(g);
                    return rc;
                }
                if (e != hipSuccess) BFAIL(MJ_EHIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                b->graph = g;
                MJ_HIP_CHECK(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
            }
            MJ_HIP_CHECK(hipGraphLaunch(b->gexec, b->stream));
            if (mode == MJ_BATCH_PERSISTENT || mode == MJ_BATCH_PERSISTENT_FUSED) b->ran_persist = 1;
            b->h_step += (uint64_t)seg;
            nsteps -= seg;
        } else {
            // unaligned steps run eagerly, one at a time
            const bool with_e = b->pt_count && b->h_step % (uint64_t)b->pt_every == 0;
            BTRY(enqueue_step(b, timestep, 0, with_e, false));
            hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, b->stream, b->d_step, (uint64_t)1);
            BTRY(blaunch_err("step counter"));
            b->h_step += 1;
            nsteps -= 1;
        }
    }
    return MJ_OK;
}

// after a run (stream synchronized): a persistent segment whose grid barrier
// gave up has left the states undefined; the barrier is reset for later runs
static int check_barrier(mj_batch *b)
{
    if (!b->d_bar || !b->ran_persist) return MJ_OK;
    b->ran_persist = 0;
    GridBar g;
    MJ_HIP_CHECK(hipMemcpy(&g, b->d_bar, sizeof g, hipMemcpyDeviceToHost));
    if (!g.abort) return MJ_OK;
    const GridBar z0 = {0u, 0u, 0u, 0u};
    MJ_HIP_CHECK(hipMemcpy(b->d_bar, &z0, sizeof z0, hipMemcpyHostToDevice));
    BFAIL(MJ_EHIP, "mj_batch: a persistent segment's grid barrier gave up (workgroups not all resident); "
                   "the batch's states are undefined");
}

extern "C" int mj_batch_set_launch(mj_batch *b, int mode, int64_t spin)
{
    if (!b) BFAIL(MJ_EINVAL, "mj_batch_set_launch: null batch");
    if (mode < MJ_BATCH_STAGES || mode > MJ_BATCH_PERSISTENT_FUSED) BFAIL(MJ_EINVAL, "mj_batch_set_launch: no such form");
    if (spin > (int64_t)0xFFFFFFFFu) BFAIL(MJ_EINVAL, "mj_batch_set_launch: spin limit above 2^32 - 1");
    const unsigned sp = spin < 0 ? (1u << 22) : (unsigned)spin;
    if (mode != b->launch || sp != b->spin) {
        MJ_HIP_CHECK(hipSetDevice(b->device));
        MJ_HIP_CHECK(hipStreamSynchronize(b->stream));
        drop_graph(b);
    }
    b->launch = mode;
    b->spin = sp;
    return MJ_OK;
}

extern "C" int mj_batch_steps(mj_batch *b, int nsteps, int integrator, double timestep)
{
    if (!b) BFAIL(MJ_EINVAL, "mj_batch_steps: null batch");
    if (integrator != MJ_EULER && integrator != MJ_HEUN && integrator != MJ_SIB && integrator != MJ_DEPONDT)
        BFAIL(MJ_ENOTSUP, "mj_batch_steps: batches run integrations 1, 2, 5 and 6");
    if (nsteps < 0 || !(timestep > 0.0)) BFAIL(MJ_EINVAL, "mj_batch_steps: nsteps >= 0 and timestep > 0");
    MJ_HIP_CHECK(hipSetDevice(b->device));
    b->integ = integrator;
    BTRY(batch_run(b, nsteps, timestep));
    MJ_HIP_CHECK(hipStreamSynchronize(b->stream));
    return check_barrier(b);
}

extern "C" int mj_batch_profile(mj_batch *b, int nsteps, int integrator, double timestep, double *total_ms)
{
    if (!b || !total_ms) BFAIL(MJ_EINVAL, "mj_batch_profile: null argument");
    if (integrator != MJ_EULER && integrator != MJ_HEUN && integrator != MJ_SIB && integrator != MJ_DEPONDT)
        BFAIL(MJ_ENOTSUP, "mj_batch_profile: batches run integrations 1, 2, 5 and 6");
    if (nsteps < 0 || !(timestep > 0.0)) BFAIL(MJ_EINVAL, "mj_batch_profile: nsteps >= 0 and timestep > 0");
    MJ_HIP_CHECK(hipSetDevice(b->device));
    b->integ = integrator;
    BTRY(batch_finalize(b));
    BTRY(set_timestep(b, timestep));
    MJ_HIP_CHECK(hipStreamSynchronize(b->stream));
    hipEvent_t e0, e1;
    MJ_HIP_CHECK(hipEventCreate(&e0));
    MJ_HIP_CHECK(hipEventCreate(&e1));
    MJ_HIP_CHECK(hipEventRecord(e0, b->stream));
    int rc = batch_run(b, nsteps, timestep);
    MJ_HIP_CHECK(hipEventRecord(e1, b->stream));
    MJ_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.0f;
    MJ_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *total_ms = ms;
    BTRY(rc);
    return check_barrier(b);
}

extern "C" int mj_batch_get_state(mj_batch *b, int sys, double *m, int64_t n3)
{
    if (!b || !m) BFAIL(MJ_EINVAL, "mj_batch_get_state: null argument");
    if (sys < 0 || sys >= (int)b->sys.size()) BFAIL(MJ_EINVAL, "mj_batch_get_state: no such system");
    const HostSys &s = b->sys[sys];
    if (n3 != 3 * s.n) BFAIL(MJ_EINVAL, "mj_batch_get_state: size must be 3 * sites");
    if (b->dirty) {
        for (int64_t i = 0; i < s.n; ++i)
            for (int a = 0; a < 3; ++a) m[3 * i + a] = s.m0[a * s.n + i];
        return MJ_OK;
    }
    MJ_HIP_CHECK(hipSetDevice(b->device));
    double *tmp = nullptr;
    MJ_HIP_CHECK(hipMalloc(&tmp, sizeof(double) * 3 * s.n));
    hipLaunchKernelGGL(k_batch_soa2aos, dim3((unsigned)((s.n + 255) / 256)), dim3(256), 0, b->stream, b->A + s.base,
                       tmp, s.n, b->np);
    hipError_t e = hipMemcpyAsync(m, tmp, sizeof(double) * 3 * s.n, hipMemcpyDeviceToHost, b->stream);
    hipError_t e2 = hipStreamSynchronize(b->stream);
    (void)hipFree(tmp);
    if (e != hipSuccess || e2 != hipSuccess) BFAIL(MJ_EHIP, "mj_batch_get_state: copy failed");
    return MJ_OK;
}

// energy per site of every system's current state (the single-system mj_energy)
extern "C" int mj_batch_energies(mj_batch *b, double *E)
{
    if (!b || !E) BFAIL(MJ_EINVAL, "mj_batch_energies: null argument");
    MJ_HIP_CHECK(hipSetDevice(b->device));
    BTRY(batch_finalize(b));
    KArgs a;
    batch_base(b, a, 1.0);
    a.src = b->A;
    a.energy = 1;
    a.epart = b->epart;
    BTRY(dip_fields(b, b->A));
    launch_stage<MODE_FIELD, false>(b, a);
    BTRY(blaunch_err("batched energy pass"));
    hipLaunchKernelGGL(k_batch_esum, dim3((unsigned)b->sys.size()), dim3(256), 0, b->stream, b->epart, b->d_sys, b->E);
    BTRY(blaunch_err("batched energy sum"));
    std::vector<double> t(b->sys.size());
    MJ_HIP_CHECK(hipMemcpyAsync(t.data(), b->E, sizeof(double) * t.size(), hipMemcpyDeviceToHost, b->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(b->stream));
    for (size_t q = 0; q < t.size(); ++q) E[q] = t[q] / (double)b->sys[q].n;
    return MJ_OK;
}

extern "C" int mj_batch_pt_state(mj_batch *b, int *temp_of_rep, uint64_t *accepted, uint64_t *attempted)
{
    if (!b || !temp_of_rep) BFAIL(MJ_EINVAL, "mj_batch_pt_state: null argument");
    if (!b->pt_count) BFAIL(MJ_ESTATE, "mj_batch_pt_state: no parallel-tempering group");
    MJ_HIP_CHECK(hipSetDevice(b->device));
    BTRY(batch_finalize(b));
    MJ_HIP_CHECK(hipStreamSynchronize(b->stream));
    MJ_HIP_CHECK(hipMemcpy(temp_of_rep, b->d_temp_of_rep, sizeof(int) * b->pt_count, hipMemcpyDeviceToHost));
    if (accepted) MJ_HIP_CHECK(hipMemcpy(accepted, b->d_acc, sizeof(uint64_t) * b->pt_count, hipMemcpyDeviceToHost));
    if (attempted) MJ_HIP_CHECK(hipMemcpy(attempted, b->d_att, sizeof(uint64_t) * b->pt_count, hipMemcpyDeviceToHost));
    return MJ_OK;
}

extern "C" int mj_batch_info(mj_batch *b, int64_t *sites, int *workgroups, int *tiles)
{
    if (!b) BFAIL(MJ_EINVAL, "mj_batch_info: null argument");
    MJ_HIP_CHECK(hipSetDevice(b->device));
    BTRY(batch_finalize(b));
    int64_t n = 0;
    for (auto &s : b->sys) n += s.n;
    if (sites) *sites = n;
    if (workgroups) *workgroups = b->nblocks;
    if (tiles)
        for (size_t q = 0; q < b->sys.size(); ++q) {
            int bx, by;
            auto_tile(b->sys[q], bx, by);
            tiles[2 * q] = bx;
            tiles[2 * q + 1] = by;
        }
    return MJ_OK;
}
