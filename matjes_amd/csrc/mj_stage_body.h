// Body of the fused field + integrator-stage kernel, shared by the
// ahead-of-time kernels (runtime geometry, mj_kernels.hip) and the kernels
// compiled per Hamiltonian at set-up (compile-time geometry and unrolled
// stencil, mj_jit.hip).  The geometry policy G supplies tile/lattice sizes and
// the stencil sum; both policies sum the stencil in the same order, so the two
// paths give identical results.
//
// One kernel evaluates B_eff = -dH/dm at every site of a tile and applies one
// integrator stage in registers, so a stage is one pass over HBM:
//   - the state (SoA mx|my|mz, FP64) of the tile plus a halo of the stencil
//     reach is staged in LDS (open boundaries zero-filled, periodic wrapped),
//     then every site sums its translation-invariant 3x3 blocks from LDS;
//   - Zeeman (uniform or per-site B order parameter) and the dipolar H_fft
//     (read from the FFT output grid) are added in the epilogue;
//   - the stage (reference Euler / two-half-step Heun, SIB, Depondt) and the
//     in-register Philox thermal field finish the site;
//   - optionally the energy of the stencil state is reduced per block, and
//     block 0 folds the previous kernel's per-block partials into E.
// Reference: field hamiltonian_collection.f90:433-454; stage loop
// spindynamics.f90:260-323; energy hamiltonian_collection.f90:365-400.
#pragma once

__device__ __forceinline__ bool mj_wrap_idx(int &g, int n, int periodic)
{
    if (g >= 0 && g < n) return true;
    if (!periodic) return false;
    g %= n;
    if (g < 0) g += n;
    return true;
}

// workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous band of tiles so halos are shared inside one L2
__device__ __forceinline__ int mj_xcd_remap(int bid, int nb)
{
    if ((nb & 7) != 0) return bid;
    return (bid & 7) * (nb >> 3) + (bid >> 3);
}

// Deterministic block reductions.  Full wavefronts: xor butterfly inside the
// wave (every lane ends with the same value: each step adds the same two
// operands in either order), then the per-wave values in wave order; one
// barrier instead of a log2(nthr)-step tree.  Blocks that are not a multiple
// of 64 threads (tiny lattices) use a tree over LDS.  Every thread gets the
// result; `red` (nthr doubles) may be reused after the call.
template <class OP>
__device__ __forceinline__ double mj_block_reduce(double v, double *red, int nthr, OP op)
{
    const int t = threadIdx.x;
    if ((nthr & 63) == 0) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = op(v, __shfl_xor(v, o, 64));
        if ((t & 63) == 0) red[t >> 6] = v;
        __syncthreads();
        double r = red[0];
        for (int w = 1; w < (nthr >> 6); ++w) r = op(r, red[w]);
        __syncthreads();
        return r;
    }
    red[t] = v;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s && t + s < nthr) red[t] = op(red[t], red[t + s]);
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

struct MjAdd {
    __device__ __forceinline__ double operator()(double a, double b) const { return a + b; }
};
struct MjMax {
    __device__ __forceinline__ double operator()(double a, double b) const { return fmax(a, b); }
};

__device__ __forceinline__ double mj_block_sum(double v, double *red, int nthr)
{
    return mj_block_reduce(v, red, nthr, MjAdd());
}

__device__ __forceinline__ double mj_block_max(double v, double *red, int nthr)
{
    return mj_block_reduce(v, red, nthr, MjMax());
}

// runtime-geometry policy (ahead-of-time kernels)
struct RtGeom {
    static constexpr int kNit = 0;         // trips of the tile fill: known only at run time
    static constexpr bool kDma = false;    // direct global -> LDS fill (compile-time geometry only)
    // LDS stride between the x, y, z components of the tile (the tile size here)
    static __device__ __forceinline__ int tstride(const KArgs &, int T) { return T; }
    static __device__ __forceinline__ int bx(const KArgs &a) { return a.bx; }
    static __device__ __forceinline__ int by(const KArgs &a) { return a.by; }
    static __device__ __forceinline__ int bz(const KArgs &a) { return a.bz; }
    static __device__ __forceinline__ int rx(const KArgs &a) { return a.rx; }
    static __device__ __forceinline__ int ry(const KArgs &a) { return a.ry; }
    static __device__ __forceinline__ int rz(const KArgs &a) { return a.rz; }
    static __device__ __forceinline__ int nx(const KArgs &a) { return a.nx; }
    static __device__ __forceinline__ int ny(const KArgs &a) { return a.ny; }
    static __device__ __forceinline__ int nz(const KArgs &a) { return a.nz; }
    static __device__ __forceinline__ int ntx(const KArgs &a) { return a.ntx; }
    static __device__ __forceinline__ int nty(const KArgs &a) { return a.nty; }
    static __device__ __forceinline__ int nblocks(const KArgs &a) { return a.nblocks; }
    // tile of this workgroup, and the slot of its energy partial (the tile's
    // index: partials are summed in tile order whichever workgroup made them)
    static __device__ __forceinline__ int block(const KArgs &a) { return mj_xcd_remap(blockIdx.x, a.nblocks); }
    static __device__ __forceinline__ int part(const KArgs &a) { return mj_xcd_remap(blockIdx.x, a.nblocks); }
    static __device__ __forceinline__ void stencil(const KArgs &a, const double *L, int T, int c, double &ax,
                                                   double &ay, double &az)
    {
        for (int k = 0; k < a.n_iso; ++k) {
            const int j = c + a.ofs[k];
            const double s = a.cf[k];
            ax = fma(s, L[j], ax);
            ay = fma(s, L[T + j], ay);
            az = fma(s, L[2 * T + j], az);
        }
        const double *cd = a.cf + a.n_iso;
        for (int k = 0; k < a.n_dm; ++k) {
            const int j = c + a.ofs[a.n_iso + k];
            const double s = cd[4 * k], wx = cd[4 * k + 1], wy = cd[4 * k + 2], wz = cd[4 * k + 3];
            const double mx = L[j], my = L[T + j], mz = L[2 * T + j];
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
            const int j = c + a.ofs[a.n_iso + a.n_dm + k];
            const double *q = cg + 9 * k;
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
    }
};


// The thread's own site in a tile, and the own-site operands of the epilogue
// (issued early: their latency overlaps the tile fill or the stencil).
struct SiteCtx {
    int lx, ly, lz, gx, gy, gz;
    bool active;
    int64_t gi;
    v3 pre_aux, pre_own;
    v3 th;            // NOISE stages: the thermal field, when computed ahead of the epilogue
    uint64_t nstep;   // NOISE stages with a compile-time geometry: the step of the Philox counter
};

// The Philox thermal field of a NOISE stage needs only (seed, step, site), not
// the tile: it is computed while the tile loads are in flight, and its
// Langevin term on m0 (HEUN2, thermal_term) is formed in the epilogue.  A field
// supplied by the caller (a.noise) is read in the epilogue too
// (site_thermal_final): loads here would make the Philox path wait for the
// tile loads.
template <int MODE>
__device__ __forceinline__ v3 thermal_term(const KArgs &a, const SiteCtx &s, v3 bt)
{
    if constexpr (MODE == MODE_HEUN2) return llg_rhs(s.pre_own, bt, a.alpha, a.den, a.rden);
    return bt;
}

template <int MODE, bool NOISE>
__device__ __forceinline__ void site_thermal(const KArgs &a, SiteCtx &s)
{
    if constexpr (NOISE) {
        if (s.active && !a.noise) {
            const v3 g = normals3(a.seed, s.nstep, (uint32_t)s.gi);
            s.th = mk3(a.sigma * g.x, a.sigma * g.y, a.sigma * g.z);
        }
    }
}

// Epilogue: the early result (kernels with a compile-time tile fill), else the
// whole thermal part here, as the ahead-of-time kernels have no fill to overlap.
template <int MODE, class G>
__device__ __forceinline__ v3 site_thermal_final(const KArgs &a, const SiteCtx &s)
{
    const int64_t np = a.np;
    if (a.noise) return thermal_term<MODE>(a, s, mk3(a.noise[s.gi], a.noise[s.gi + np], a.noise[s.gi + 2 * np]));
    if constexpr (G::kNit > 0) {
        return thermal_term<MODE>(a, s, s.th);
    } else {
        const v3 g = normals3(a.seed, a.stepp ? *a.stepp + a.step : a.step, (uint32_t)s.gi);
        return thermal_term<MODE>(a, s, mk3(a.sigma * g.x, a.sigma * g.y, a.sigma * g.z));
    }
}

// Interior tile straight from global memory into LDS: each lane moves 16 B,
// two neighbouring sites of a tile row (tile rows and the lattice rows have
// even length and start at even sites, so a pair never straddles a row and is
// 16-byte aligned), and lands at its wave's LDS base + 16 * lane.  Every lane
// of every wave issues (the components are padded to whole trips of the
// workgroup: surplus lanes reload the last pair into the padding), so the
// code is branch-free: no join after which the compiler's waits for other
// loads would become full waits.  The loads bypass the registers; the caller
// waits for them (vmcnt) before the barrier.
template <class G>
__device__ __forceinline__ void tile_fill_direct(const KArgs &a, double *lds, int64_t base)
{
    constexpr int BX = G::kBx, BY = G::kBy, BZ = G::kBz, RX = G::kRx, RY = G::kRy, RZ = G::kRz;
    constexpr int SX = BX + 2 * RX, SY = BY + 2 * RY, SZ = BZ + 2 * RZ;
    constexpr int T = SX * SY * SZ, P = T / 2, NTHR = BX * BY * BZ, TS = G::kT;
    constexpr int NX = G::kNx, NY = G::kNy;
    static_assert(TS >= 2 * NTHR * ((P + NTHR - 1) / NTHR), "component stride must hold whole trips");
    const int64_t np = a.np;
    const int wave = (int)threadIdx.x >> 6;
#pragma unroll
    for (int p0 = 0; p0 < P; p0 += NTHR) {
        const int q = min(p0 + (int)threadIdx.x, P - 1);
        const int t = 2 * q;
        const int lx = t % SX;
        const int r = t / SX;
        const int ly = r % SY;
        const int lz = r / SY;
        const int64_t g = base + lx + (int64_t)NX * ((int64_t)ly + (int64_t)NY * lz);
        const int wfirst = __builtin_amdgcn_readfirstlane(p0 + 64 * wave);
#pragma unroll
        for (int c = 0; c < 3; ++c)
            __builtin_amdgcn_global_load_lds((const void *)(a.src + g + c * np),
                                             (__attribute__((address_space(3))) void *)(lds + c * TS + 2 * wfirst), 16,
                                             0, 0);
    }
}

// One tile of the state plus its halo into LDS: interior tiles take a
// branch-free path, boundary tiles wrap (periodic) or zero-fill (open).  With
// a compile-time trip count the loads go to registers first and the stage's
// thermal part is computed before they are stored.
template <int MODE, bool NOISE, class G>
__device__ __forceinline__ void tile_fill(const KArgs &a, double *lds, int x0, int y0, int z0, SiteCtx &st)
{
    const int BX = G::bx(a), BY = G::by(a), BZ = G::bz(a);
    const int RX = G::rx(a), RY = G::ry(a), RZ = G::rz(a);
    const int NX = G::nx(a), NY = G::ny(a), NZ = G::nz(a);
    const int SX = BX + 2 * RX, SY = BY + 2 * RY, SZ = BZ + 2 * RZ;
    const int TT = SX * SY * SZ;          // tile sites
    const int T = G::tstride(a, TT);      // LDS component stride
    const int nthr = BX * BY * BZ;
    const int64_t np = a.np;
    const bool interior = x0 - RX >= 0 && x0 + BX + RX <= NX && y0 - RY >= 0 && y0 + BY + RY <= NY &&
                          z0 - RZ >= 0 && z0 + BZ + RZ <= NZ;
    if (interior) {
        const int64_t base = (int64_t)(x0 - RX) + (int64_t)NX * ((int64_t)(y0 - RY) + (int64_t)NY * (z0 - RZ));
        if constexpr (G::kNit > 0) {
            constexpr int NIT = G::kNit;
            double vx[NIT], vy[NIT], vz[NIT];
            // unconditional loads (the last trip's surplus threads reload element
            // T-1 and store nothing): straight-line code keeps the compiler's
            // waits on the tile loads exact, so the thermal part overlaps them
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int t = min(threadIdx.x + it * nthr, TT - 1);
                const int lx = t % SX;
                const int r = t / SX;
                const int ly = r % SY;
                const int lz = r / SY;
                const int64_t g = base + lx + (int64_t)NX * ((int64_t)ly + (int64_t)NY * lz);
                vx[it] = a.src[g];
                vy[it] = a.src[g + np];
                vz[it] = a.src[g + 2 * np];
            }
            site_thermal<MODE, NOISE>(a, st);
#pragma unroll
            for (int it = 0; it < NIT; ++it) {
                const int t = threadIdx.x + it * nthr;
                if (t < TT) {
                    lds[t] = vx[it];
                    lds[T + t] = vy[it];
                    lds[2 * T + t] = vz[it];
                }
            }
            return;
        }
        const int nit = (TT + nthr - 1) / nthr;
        for (int it = 0; it < nit; ++it) {
            const int t = threadIdx.x + it * nthr;
            if (t >= TT) break;
            const int lx = t % SX;
            const int r = t / SX;
            const int ly = r % SY;
            const int lz = r / SY;
            const int64_t g = base + lx + (int64_t)NX * ((int64_t)ly + (int64_t)NY * lz);
            lds[t] = a.src[g];
            lds[T + t] = a.src[g + np];
            lds[2 * T + t] = a.src[g + 2 * np];
        }
    } else {
        if constexpr (G::kNit > 0) site_thermal<MODE, NOISE>(a, st);
        for (int t = threadIdx.x; t < TT; t += nthr) {
            const int lx = t % SX;
            const int r = t / SX;
            const int ly = r % SY;
            const int lz = r / SY;
            int gx = x0 + lx - RX, gy = y0 + ly - RY, gz = z0 + lz - RZ;
            double vx = 0.0, vy = 0.0, vz = 0.0;
            if (mj_wrap_idx(gx, NX, a.px) && mj_wrap_idx(gy, NY, a.py) && mj_wrap_idx(gz, NZ, a.pz)) {
                const int64_t g = (int64_t)gx + (int64_t)NX * ((int64_t)gy + (int64_t)NY * gz);
                vx = a.src[g];
                vy = a.src[g + np];
                vz = a.src[g + 2 * np];
            }
            lds[t] = vx;
            lds[T + t] = vy;
            lds[2 * T + t] = vz;
        }
    }
}

template <int MODE, bool NOISE, class G>
__device__ __forceinline__ SiteCtx site_begin(const KArgs &a, int x0, int y0, int z0, uint64_t nstep)
{
    SiteCtx s;
    const int BX = G::bx(a), BY = G::by(a);
    const int64_t np = a.np;
    s.lx = threadIdx.x % BX;
    s.ly = (threadIdx.x / BX) % BY;
    s.lz = threadIdx.x / (BX * BY);
    s.gx = x0 + s.lx;
    s.gy = y0 + s.ly;
    s.gz = z0 + s.lz;
    s.active = s.gx < G::nx(a) && s.gy < G::ny(a) && s.gz < G::nz(a);
    s.gi = s.active ? (int64_t)s.gx + (int64_t)G::nx(a) * ((int64_t)s.gy + (int64_t)G::ny(a) * s.gz) : 0;
    constexpr bool kAux = MODE == MODE_HEUN2 || MODE == MODE_DEP2;
    constexpr bool kOwn = (MODE == MODE_HEUN2 && NOISE) || MODE == MODE_SIB2 || MODE == MODE_DEP2;
    s.pre_aux = mk3(0.0, 0.0, 0.0);
    s.pre_own = mk3(0.0, 0.0, 0.0);
    s.th = mk3(0.0, 0.0, 0.0);
    s.nstep = nstep;
    // compile-time geometries load unconditionally (gi = 0 off the lattice): no
    // join that would make the thermal part wait for them
    constexpr bool kAll = G::kNit > 0;
    if constexpr (kAux && !(NOISE && G::kNit > 0)) {
        if (kAll || s.active) s.pre_aux = mk3(a.aux[s.gi], a.aux[s.gi + np], a.aux[s.gi + 2 * np]);
    }
    if constexpr (kOwn) {
        if (kAll || s.active) s.pre_own = mk3(a.own[s.gi], a.own[s.gi + np], a.own[s.gi + 2 * np]);
    }
    return s;
}

// Field at the own site from the tile in LDS, the block reductions (energy,
// max |B|: slot `part`; `fold`: this block folds the previous kernel's
// energy partials) and the integrator epilogue.
template <int MODE, bool NOISE, class G>
__device__ __forceinline__ void site_end(const KArgs &a, double *lds, const SiteCtx &st, int part, bool fold)
{
    const int BX = G::bx(a), BY = G::by(a), BZ = G::bz(a);
    const int RX = G::rx(a), RY = G::ry(a), RZ = G::rz(a);
    const int SX = BX + 2 * RX, SY = BY + 2 * RY, SZ = BZ + 2 * RZ;
    const int T = G::tstride(a, SX * SY * SZ);   // LDS component stride
    const int nthr = BX * BY * BZ;
    const int64_t np = a.np;
    const int lx = st.lx, ly = st.ly, lz = st.lz, gx = st.gx, gy = st.gy, gz = st.gz;
    const bool active = st.active;
    const int64_t gi = st.gi;
    const v3 pre_aux = st.pre_aux, pre_own = st.pre_own;
    (void)pre_aux;
    (void)pre_own;
    (void)SZ;

    double ax = 0.0, ay = 0.0, az = 0.0;   // sum_d C_d m_{i+d}  (= (H + H^T) m at i)
    v3 m;                                  // state at the own site
    if (a.use_stencil) {
        const int c = (lz + RZ) * SX * SY + (ly + RY) * SX + (lx + RX);
        m = mk3(lds[c], lds[T + c], lds[2 * T + c]);
        G::stencil(a, lds, T, c, ax, ay, az);
    } else {
        m = active ? mk3(a.src[gi], a.src[gi + np], a.src[gi + 2 * np]) : mk3(0.0, 0.0, 0.0);
        if (active) {
            ax = a.fld_in[gi];
            ay = a.fld_in[gi + np];
            az = a.fld_in[gi + 2 * np];
        }
    }

    // Zeeman: Z^T b (uniform or per site)
    double zx = 0.0, zy = 0.0, zz = 0.0;
    if (a.bmode == 1) {
        zx = a.zb[0];
        zy = a.zb[1];
        zz = a.zb[2];
    } else if (a.bmode == 3 && active) {
        double bxv = 0.0, byv = 0.0, bzv = 0.0;
        // unrolled over the at most 4 maps: constant indices into wmap keep a
        // kernel's local copy of the arguments in registers (no scratch)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (k < a.nmaps) {
                const double *f = a.maps + (int64_t)k * 3 * np;
                bxv = bxv + a.wmap[3 * k] * f[gi];
                byv = byv + a.wmap[3 * k + 1] * f[gi + np];
                bzv = bzv + a.wmap[3 * k + 2] * f[gi + 2 * np];
            }
        }
        zx = (bxv * a.z[0] + byv * a.z[3]) + bzv * a.z[6];
        zy = (bxv * a.z[1] + byv * a.z[4]) + bzv * a.z[7];
        zz = (bxv * a.z[2] + byv * a.z[5]) + bzv * a.z[8];
    } else if (a.bmode == 2 && active) {
        const double bxv = a.bmap[gi], byv = a.bmap[gi + np], bzv = a.bmap[gi + 2 * np];
        zx = (bxv * a.z[0] + byv * a.z[3]) + bzv * a.z[6];
        zy = (bxv * a.z[1] + byv * a.z[4]) + bzv * a.z[7];
        zz = (bxv * a.z[2] + byv * a.z[5]) + bzv * a.z[8];
    }
    // dipolar H_fft at this site (the padded grid keeps the lattice at its origin)
    double hx = 0.0, hy = 0.0, hz = 0.0;
    if (a.hdip && active) {
        const int64_t h = (int64_t)gx + (int64_t)a.dnx * ((int64_t)gy + (int64_t)a.dny * gz);
        hx = a.hdip[h];
        hy = a.hdip[h + a.hstride];
        hz = a.hdip[h + 2 * a.hstride];
    }
    // B = -(H m + H^T m + Z^T b) - 2 H_fft      (hamiltonian_collection.f90:442-453)
    v3 B = mk3(-(ax + zx) - 2.0 * hx, -(ay + zy) - 2.0 * hy, -(az + zz) - 2.0 * hz);

    // max |B| over the block (update_time.f90:58-68: the field of the last stage, no thermal part)
    if (a.bmax_part) {
        const double nb = active ? norm3(B) : 0.0;
        __syncthreads();
        const double s = mj_block_max(nb, lds, nthr);
        if (threadIdx.x == 0) a.bmax_part[part] = s;
    }
    // energy of the stencil state: m.(H m) + b.(Z m) + m.H_fft = m.(acc/2 + Z^T b + h)
    if (a.energy) {
        double e = 0.0;
        if (active)
            e = m.x * ((0.5 * ax + zx) + hx) + m.y * ((0.5 * ay + zy) + hy) + m.z * ((0.5 * az + zz) + hz);
        __syncthreads();
        double s = mj_block_sum(e, lds, nthr);
        if (threadIdx.x == 0) a.epart[part] = s;
    }
    // block 0 folds the previous kernel's partials into E (stream order makes them complete)
    if (a.ered_in && fold) {
        __syncthreads();
        double v = 0.0;
        for (int i = threadIdx.x; i < a.n_ered; i += nthr) v += a.ered_in[i];
        double s = mj_block_sum(v, lds, nthr);
        if (threadIdx.x == 0) *a.e_out = s * a.e_scale;
    }
    if (!active) return;

    const double alpha = a.alpha;
    // the stage's effective field (no thermal term): MODE_FIELD's product, or
    // the field of a step's last stage when the host keeps it
    if (a.fld_out) {
        a.fld_out[gi] = B.x;
        a.fld_out[gi + np] = B.y;
        a.fld_out[gi + 2 * np] = B.z;
    }
    if constexpr (MODE == MODE_FIELD) {
    } else if constexpr (MODE == MODE_EULER) {
        v3 D = llg_rhs(m, B, alpha, a.den, a.rden);
        // Dint = 0 + 1 D (get_Dmode_int); the leading zero changes only the sign of an exact zero
        v3 dint = D;
        v3 o;
        if constexpr (NOISE) o = euler_step(m, dint, llg_rhs(m, site_thermal_final<MODE, G>(a, st), alpha, a.den, a.rden), a.dt);
        else o = euler_step0(m, dint, a.dt);
        a.dst[gi] = o.x;
        a.dst[gi + np] = o.y;
        a.dst[gi + 2 * np] = o.z;
    } else if constexpr (MODE == MODE_RKS) {
        // stage i of integration 4 (spindynamics.f90:294-318 with get_Dmode_int,
        // solver_order.f90:46-58): D_i = LLG(B(m), m); Dint = (0 + w_1 D_1) + ... + w_i D_i;
        // m' = euler(m, Dint, 0, dt_i).  Unrolled over the at most 7 earlier stages
        // (constant indices keep the argument block in registers).
        v3 D = llg_rhs(m, B, alpha, a.den, a.rden);
        double dx = 0.0, dy = 0.0, dz = 0.0;
#pragma unroll
        for (int k = 0; k < 7; ++k) {
            if (k < a.nprev) {
                const double *p = a.dprev[k];
                dx = dx + a.wrk[k] * p[gi];
                dy = dy + a.wrk[k] * p[gi + np];
                dz = dz + a.wrk[k] * p[gi + 2 * np];
            }
        }
        double wi = 0.0;
#pragma unroll
        for (int k = 0; k < 8; ++k)
            if (k == a.nprev) wi = a.wrk[k];
        dx = dx + wi * D.x;
        dy = dy + wi * D.y;
        dz = dz + wi * D.z;
        v3 o = euler_step0(m, mk3(dx, dy, dz), a.dt);
        if (a.dout) {
            a.dout[gi] = D.x;
            a.dout[gi + np] = D.y;
            a.dout[gi + 2 * np] = D.z;
        }
        a.dst[gi] = o.x;
        a.dst[gi + np] = o.y;
        a.dst[gi + 2 * np] = o.z;
    } else if constexpr (MODE == MODE_HEUN1) {
        // stage 1 of integration 2: D1 = LLG(B(m0), m0); m1 = euler(m0, 1*D1 + 0*D2, 0, dt/2)
        v3 D = llg_rhs(m, B, alpha, a.den, a.rden);
        // Dint = 0 + 1 D (get_Dmode_int); the leading zero changes only the sign of an exact zero
        v3 dint = D;
        v3 o = euler_step0(m, dint, a.dt);
        a.dst[gi] = o.x;
        a.dst[gi + np] = o.y;
        a.dst[gi + 2 * np] = o.z;
        a.aux[gi] = D.x;
        a.aux[gi + np] = D.y;
        a.aux[gi + 2 * np] = D.z;
    } else if constexpr (MODE == MODE_HEUN2) {
        // stage 2: D2 = LLG(B(m1), m1); m2 = euler(m1, 0.5 D1 + 0.5 D2, DT(m0), dt/2)
        v3 D2 = llg_rhs(m, B, alpha, a.den, a.rden);
        const v3 D1 = pre_aux;
        // Dint = (0 + D1/2) + D2/2 (get_Dmode_int) without the exact-zero start
        v3 dint = mk3(0.5 * D1.x + 0.5 * D2.x, 0.5 * D1.y + 0.5 * D2.y, 0.5 * D1.z + 0.5 * D2.z);
        v3 o;
        if constexpr (NOISE) {
            // Langevin term of the last stage on m0 (pre_own, site_thermal), spindynamics.f90:269-271,311
            o = euler_step(m, dint, site_thermal_final<MODE, G>(a, st), a.dt);
        } else {
            o = euler_step0(m, dint, a.dt);
        }
        a.dst[gi] = o.x;
        a.dst[gi + np] = o.y;
        a.dst[gi + 2 * np] = o.z;
    } else {
        if constexpr (NOISE) {
            const v3 bt = site_thermal_final<MODE, G>(a, st);
            B = mk3(B.x + bt.x, B.y + bt.y, B.z + bt.z);
        }
        if constexpr (MODE == MODE_SIB1) {
            v3 av = sib_vec(m, B, alpha, a.dt);
            v3 mt = cayley(m, av);
            a.dst[gi] = 0.5 * (m.x + mt.x);
            a.dst[gi + np] = 0.5 * (m.y + mt.y);
            a.dst[gi + 2 * np] = 0.5 * (m.z + mt.z);
        } else if constexpr (MODE == MODE_SIB2) {
            v3 av = sib_vec(m, B, alpha, a.dt);   // m is the midpoint p here
            v3 o = cayley(pre_own, av);
            a.dst[gi] = o.x;
            a.dst[gi + np] = o.y;
            a.dst[gi + 2 * np] = o.z;
        } else if constexpr (MODE == MODE_DEP1) {
            v3 b1 = dep_field(m, B, alpha, a.den, a.rden);
            v3 mt = dep_rotate(m, b1, a.dt);
            a.dst[gi] = mt.x;
            a.dst[gi + np] = mt.y;
            a.dst[gi + 2 * np] = mt.z;
            a.aux[gi] = b1.x;
            a.aux[gi + np] = b1.y;
            a.aux[gi + 2 * np] = b1.z;
        } else if constexpr (MODE == MODE_DEP2) {
            v3 b2 = dep_field(m, B, alpha, a.den, a.rden);   // m is the predictor here
            const v3 b1 = pre_aux;
            v3 ba = mk3(0.5 * (b1.x + b2.x), 0.5 * (b1.y + b2.y), 0.5 * (b1.z + b2.z));
            v3 o = dep_rotate(pre_own, ba, a.dt);
            a.dst[gi] = o.x;
            a.dst[gi + np] = o.y;
            a.dst[gi + 2 * np] = o.z;
        }
    }
}

template <int MODE, bool NOISE, class G>
__device__ __forceinline__ void stage_body(const KArgs &a, double *lds)
{
    const int bid = G::block(a);
    const int tx = bid % G::ntx(a);
    const int tr = bid / G::ntx(a);
    const int ty = tr % G::nty(a);
    const int tz = tr / G::nty(a);
    const int x0 = tx * G::bx(a), y0 = ty * G::by(a), z0 = tz * G::bz(a);
    // interior tiles of the compile-time geometries: the tile goes straight to
    // LDS first, before any other load (the compiler does not count these loads
    // among the ones in flight, so a later wait on a register load would also
    // wait for them); the thermal field is computed while they land
    // the Philox step, read before anything is written (a scalar load)
    uint64_t nstep = 0;
    if constexpr (NOISE && G::kNit > 0) nstep = a.stepp ? *a.stepp + a.step : a.step;
    bool direct = false;
    if constexpr (G::kDma) {
        constexpr int RX = G::kRx, RY = G::kRy, RZ = G::kRz, NX = G::kNx, NY = G::kNy;
        direct = a.use_stencil && (((unsigned long long)a.src) & 15ull) == 0 && x0 - RX >= 0 &&
                 x0 + G::kBx + RX <= NX && y0 - RY >= 0 && y0 + G::kBy + RY <= NY && z0 - RZ >= 0 &&
                 z0 + G::kBz + RZ <= G::nz(a);
        if (direct)
            tile_fill_direct<G>(a, lds,
                                (int64_t)(x0 - RX) + (int64_t)NX * ((int64_t)(y0 - RY) + (int64_t)NY * (z0 - RZ)));
    }
    SiteCtx st = site_begin<MODE, NOISE, G>(a, x0, y0, z0, nstep);
    if (direct) {
        site_thermal<MODE, NOISE>(a, st);
        __builtin_amdgcn_s_waitcnt(0);   // this wave's tile loads are in LDS
    } else if (a.use_stencil) {
        tile_fill<MODE, NOISE, G>(a, lds, x0, y0, z0, st);
    } else if constexpr (G::kNit > 0) {
        site_thermal<MODE, NOISE>(a, st);
    }
    // with the thermal part computed early, the epilogue's aux operand is
    // loaded after it (fewer registers live across it)
    constexpr bool kAux = MODE == MODE_HEUN2 || MODE == MODE_DEP2;
    if constexpr (kAux && NOISE && G::kNit > 0) st.pre_aux = mk3(a.aux[st.gi], a.aux[st.gi + a.np], a.aux[st.gi + 2 * a.np]);
    if (a.use_stencil) __syncthreads();
    site_end<MODE, NOISE, G>(a, lds, st, G::part(a), blockIdx.x == 0);
}


// Persistent form of stage_body for stages that read nothing but their tile
// (stage 1 of the two-stage integrators; compile-time geometry with the direct
// global -> LDS fill, a whole number of tiles per XCD band).  Each workgroup
// walks its XCD's band of tiles; the next interior tile goes from HBM straight
// into the second of two LDS buffers while the current one is computed, so the
// fill latency of all but the first tile is hidden.  Same arithmetic and
// energy-partial slots (the tile index) as stage_body.
template <int MODE, class G>
__device__ __forceinline__ void stage_fill_any(const KArgs &a, double *buf, int t)
{
    constexpr int RX = G::kRx, RY = G::kRy, RZ = G::kRz, NX = G::kNx, NY = G::kNy;
    const int tx = t % G::ntx(a), tr = t / G::ntx(a), ty = tr % G::nty(a), tz = tr / G::nty(a);
    const int x0 = tx * G::kBx, y0 = ty * G::kBy, z0 = tz * G::kBz;
    const bool direct = (((unsigned long long)a.src) & 15ull) == 0 && x0 - RX >= 0 && x0 + G::kBx + RX <= NX &&
                        y0 - RY >= 0 && y0 + G::kBy + RY <= NY && z0 - RZ >= 0 && z0 + G::kBz + RZ <= G::nz(a);
    if (direct) {
        tile_fill_direct<G>(a, buf, (int64_t)(x0 - RX) + (int64_t)NX * ((int64_t)(y0 - RY) + (int64_t)NY * (z0 - RZ)));
    } else {
        SiteCtx dummy;
        dummy.active = false;
        tile_fill<MODE, false, G>(a, buf, x0, y0, z0, dummy);
    }
}

template <int MODE, class G>
__device__ __forceinline__ void stage_body_persistent(const KArgs &a, double *lds)
{
    static_assert(G::kDma, "the persistent form needs the direct LDS fill");
    static_assert(MODE == MODE_HEUN1 || MODE == MODE_EULER || MODE == MODE_SIB1 || MODE == MODE_DEP1,
                  "stages with own-site operands are not persistent");
    constexpr int TS = G::kT;
    const int per_xcd = (int)gridDim.x >> 3;          // gridDim.x is a multiple of 8
    const int band = G::nblocks(a) >> 3;              // tiles per XCD band
    const int xcd = (int)blockIdx.x & 7, slot = (int)blockIdx.x >> 3;
    double *buf0 = lds, *buf1 = lds + 3 * TS;
    int k = 0;
    int t = slot < band ? xcd * band + slot : -1;
    if (t < 0) return;                                // uniform: the whole workgroup
    stage_fill_any<MODE, G>(a, buf0, t);
    for (;;) {
        const int jn = slot + (k + 1) * per_xcd;
        const int tn = jn < band ? xcd * band + jn : -1;
        __builtin_amdgcn_s_waitcnt(0);                // this wave's tile loads (and earlier stores) done
        __syncthreads();                              // every wave's: tile t is in LDS, the other buffer is free
        double *cur = (k & 1) ? buf1 : buf0;
        double *nxt = (k & 1) ? buf0 : buf1;
        if (tn >= 0) stage_fill_any<MODE, G>(a, nxt, tn);
        const int tx = t % G::ntx(a), tr = t / G::ntx(a), ty = tr % G::nty(a), tz = tr / G::nty(a);
        SiteCtx st = site_begin<MODE, false, G>(a, tx * G::kBx, ty * G::kBy, tz * G::kBz, 0);
        site_end<MODE, false, G>(a, cur, st, t, false);
        if (tn < 0) break;
        t = tn;
        ++k;
    }
}

// ---------------------------------------------------------------------------
// Both stages of the reference's integration 2 in one kernel (compile-time
// geometry F, mj_jit.hip).  Regions of a tile (BX x BY x BZ sites, reach R):
//   S0 = the tile plus 2R on every side: m0, staged in LDS;
//   S1 = the tile plus R: stage 1 (HEUN1) runs on every S1 site from S0 and
//        its m1 goes to a second LDS region;
// then stage 2 (HEUN2) runs on the tile from S1 and writes m2.  S1 sites
// outside the tile repeat stage 1 of the neighbouring tiles with the same
// arithmetic, and every operation is the two-kernel path's, so m2 equals that
// path's bit for bit while HBM sees m0 in and m2 out (48 B per site-step
// instead of 144).  A thread takes kSpt sites, one in each sub-tile of
// BY/kSpt rows, so that the tile (and with it the share of S1 outside it)
// grows without more threads per workgroup.  Energy partials of m0 are
// formed per tile of the two-kernel path (the sub-tiles of BY_BASE rows,
// whole wavefronts, in that path's lane order);
// workgroup 0 of the next launch folds them in that path's order (a fold by
// the last workgroup to finish would need agent-scope fences in every
// workgroup, and those write back the whole L2 on this multi-XCD part).
// ---------------------------------------------------------------------------
// Zeeman of the fused kernel: the field mode is a compile-time constant of the
// geometry (no branch between the stencil's LDS reads and their sums)
template <class F>
__device__ __forceinline__ v3 fused_zeeman(const KArgs &a, bool live, int64_t gi)
{
    double zx = 0.0, zy = 0.0, zz = 0.0;
    if constexpr (F::kBmode == 1) {
        zx = a.zb[0];
        zy = a.zb[1];
        zz = a.zb[2];
    } else if (F::kBmode == 2 && live) {
        const int64_t np = a.np;
        const double bxv = a.bmap[gi], byv = a.bmap[gi + np], bzv = a.bmap[gi + 2 * np];
        zx = (bxv * a.z[0] + byv * a.z[3]) + bzv * a.z[6];
        zy = (bxv * a.z[1] + byv * a.z[4]) + bzv * a.z[7];
        zz = (bxv * a.z[2] + byv * a.z[5]) + bzv * a.z[8];
    }
    return mk3(zx, zy, zz);
}

// global site of a region position (lattice coordinates g*), wrapped on
// periodic axes; false off an open edge
template <class F>
__device__ __forceinline__ bool fused_site(const KArgs &a, int gx, int gy, int gz, int64_t &gi)
{
    if (!mj_wrap_idx(gx, F::kNx, a.px) || !mj_wrap_idx(gy, F::kNy, a.py) || !mj_wrap_idx(gz, F::kNz, a.pz))
        return false;
    gi = (int64_t)gx + (int64_t)F::kNx * ((int64_t)gy + (int64_t)F::kNy * gz);
    return true;
}

// stage 1 of integration 2 at one site from S0 (center c0): D1 and m1
// (site_end, MODE_HEUN1), the energy term of m0 when wanted
template <class F>
__device__ __forceinline__ void fused_stage1(const KArgs &a, const double *L0, int c0, bool live, int64_t gi,
                                             v3 &m0, v3 &D1, v3 &m1, double &e)
{
    constexpr int T0 = F::kT0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    m0 = mk3(L0[c0], L0[T0 + c0], L0[2 * T0 + c0]);
    F::stencil0(L0, c0, ax, ay, az);
    const v3 z = fused_zeeman<F>(a, live, gi);
    const double hx = 0.0, hy = 0.0, hz = 0.0;
    const v3 B = mk3(-(ax + z.x) - 2.0 * hx, -(ay + z.y) - 2.0 * hy, -(az + z.z) - 2.0 * hz);
    e = m0.x * ((0.5 * ax + z.x) + hx) + m0.y * ((0.5 * ay + z.y) + hy) + m0.z * ((0.5 * az + z.z) + hz);
    D1 = llg_rhs(m0, B, a.alpha, a.den, a.rden);
    m1 = euler_step0(m0, D1, a.dt);   // Dint = 1 D1, as in stage 1 of the two-kernel path
}

template <class F>
__device__ __forceinline__ void fused_s0_fill(const KArgs &a, double *L0, int x0, int y0, int z0, bool direct)
{
    constexpr int SX = F::kS0x, SY = F::kS0y, SZ = F::kS0z, T0 = F::kT0, NTHR = F::kNthr;
    constexpr int RX2 = 2 * F::kRx, RY2 = 2 * F::kRy, RZ2 = 2 * F::kRz;
    const int64_t np = a.np;
    if constexpr (F::kDma) {
      if (direct) {
        // interior: 16-byte pairs along x straight into LDS (tile_fill_direct)
        constexpr int P = SX * SY * SZ / 2;
        static_assert(T0 >= 2 * NTHR * ((P + NTHR - 1) / NTHR), "S0 stride must hold whole trips");
        const int64_t base =
            (int64_t)(x0 - RX2) + (int64_t)F::kNx * ((int64_t)(y0 - RY2) + (int64_t)F::kNy * (z0 - RZ2));
        const int wave = (int)threadIdx.x >> 6;
#pragma unroll
        for (int p0 = 0; p0 < P; p0 += NTHR) {
            const int q = min(p0 + (int)threadIdx.x, P - 1);
            const int t = 2 * q;
            const int lx = t % SX;
            const int r = t / SX;
            const int ly = r % SY;
            const int lz = r / SY;
            const int64_t g = base + lx + (int64_t)F::kNx * ((int64_t)ly + (int64_t)F::kNy * lz);
            const int wfirst = __builtin_amdgcn_readfirstlane(p0 + 64 * wave);
#pragma unroll
            for (int c = 0; c < 3; ++c)
                __builtin_amdgcn_global_load_lds((const void *)(a.src + g + c * np),
                                                 (__attribute__((address_space(3))) void *)(L0 + c * T0 + 2 * wfirst),
                                                 16, 0, 0);
        }
        return;
      }
    }
    (void)direct;
    for (int t = threadIdx.x; t < SX * SY * SZ; t += NTHR) {
        const int lx = t % SX;
        const int r = t / SX;
        const int ly = r % SY;
        const int lz = r / SY;
        int64_t g = 0;
        double vx = 0.0, vy = 0.0, vz = 0.0;
        if (fused_site<F>(a, x0 + lx - RX2, y0 + ly - RY2, z0 + lz - RZ2, g)) {
            vx = a.src[g];
            vy = a.src[g + np];
            vz = a.src[g + 2 * np];
        }
        L0[t] = vx;
        L0[T0 + t] = vy;
        L0[2 * T0 + t] = vz;
    }
}

// S1 position of ring site k (S1 minus the tile): z slabs, then y slabs of
// the tile's planes, then x slabs of its rows
template <class F>
__device__ __forceinline__ void fused_ring_pos(int k, int &sx, int &sy, int &sz)
{
    constexpr int BX = F::kBx, BY = F::kBy, BZ = F::kBz, RX = F::kRx, RY = F::kRy, RZ = F::kRz;
    constexpr int S1X = BX + 2 * RX, S1Y = BY + 2 * RY;
    constexpr int NZS = 2 * RZ * S1X * S1Y, NYS = 2 * RY * BZ * S1X;
    if (k < NZS) {
        const int plane = k / (S1X * S1Y), rem = k % (S1X * S1Y);
        sz = plane < RZ ? plane : plane + BZ;
        sy = rem / S1X;
        sx = rem % S1X;
    } else if (k < NZS + NYS) {
        const int q = k - NZS, row = q / S1X;
        const int yy = row % (2 * RY);
        sz = RZ + row / (2 * RY);
        sy = yy < RY ? yy : yy + BY;
        sx = q % S1X;
    } else {
        const int q = k - NZS - NYS, col = q % (2 * RX), row = q / (2 * RX);
        sy = RY + row % BY;
        sz = RZ + row / BY;
        sx = col < RX ? col : col + BX;
    }
}

template <bool NOISE, class F>
__device__ __forceinline__ void fused_heun_body(const KArgs &a, double *lds)
{
    constexpr int BX = F::kBx, BY = F::kBy, BZ = F::kBz, RX = F::kRx, RY = F::kRy, RZ = F::kRz;
    constexpr int SPT = F::kSpt, NTHR = F::kNthr, NB = F::kNthrBase, NBW = NB / 64;
    constexpr int BYS = BY / SPT;                 // rows of one sub-tile of the thread's sites
    constexpr int S0X = F::kS0x, S0Y = F::kS0y, T0 = F::kT0;
    constexpr int S1X = BX + 2 * RX, S1Y = BY + 2 * RY, S1Z = BZ + 2 * RZ, T1 = S1X * S1Y * S1Z;
    constexpr int NRING = T1 - BX * BY * BZ;
    static_assert(NB % 64 == 0 && NTHR * SPT == BX * BY * BZ, "whole wavefronts, every tile site once");
    static_assert(SPT == 1 || (SPT == F::kNsub && NTHR == NB && BZ == 1), "one site per energy sub-tile");
    double *L0 = lds;                 // S0: 3 x T0
    double *L1 = lds + 3 * T0;        // S1: 3 x T1
    double *red = L1 + 3 * T1;        // per-wave sums (BX BY BZ / 64)
    const int64_t np = a.np;

    const int bid = mj_xcd_remap(blockIdx.x, F::kNblocks);
    const int tx = bid % F::kNtx, tr = bid / F::kNtx, ty = tr % F::kNty, tz = tr / F::kNty;
    const int x0 = tx * BX, y0 = ty * BY, z0 = tz * BZ;
    uint64_t nstep = 0;
    if constexpr (NOISE) nstep = a.stepp ? *a.stepp + a.step : a.step;
    const bool direct = F::kDma && (((unsigned long long)a.src) & 15ull) == 0 && x0 - 2 * RX >= 0 &&
                        x0 + BX + 2 * RX <= F::kNx && y0 - 2 * RY >= 0 && y0 + BY + 2 * RY <= F::kNy &&
                        z0 - 2 * RZ >= 0 && z0 + BZ + 2 * RZ <= F::kNz;
    fused_s0_fill<F>(a, L0, x0, y0, z0, direct);

    // the thread's own sites: row ly + j BYS of the tile for j < SPT
    const int lx = threadIdx.x % BX, ly = (threadIdx.x / BX) % BYS, lz = threadIdx.x / (BX * BYS);
    const int gx = x0 + lx, gz = z0 + lz;
    bool active[SPT];
    int64_t gia[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int gy = y0 + ly + j * BYS;
        active[j] = gx < F::kNx && gy < F::kNy && gz < F::kNz;
        gia[j] = active[j] ? (int64_t)gx + (int64_t)F::kNx * ((int64_t)gy + (int64_t)F::kNy * gz) : 0;
    }
    // the Philox thermal field of the own sites while the tile lands
    v3 th[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        th[j] = mk3(0.0, 0.0, 0.0);
        if constexpr (NOISE) {
            if (active[j]) {
                const v3 g = normals3(a.seed, nstep, (uint32_t)gia[j]);
                th[j] = mk3(a.sigma * g.x, a.sigma * g.y, a.sigma * g.z);
            }
        }
    }
    if (direct) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // pass A: stage 1 at the own positions (wrapped where the tile passes a periodic edge)
    v3 D1[SPT], q[SPT];
    double e[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int lyj = ly + j * BYS;
        D1[j] = mk3(0.0, 0.0, 0.0);
        q[j] = mk3(0.0, 0.0, 0.0);
        e[j] = 0.0;
        const int s1 = (lz + RZ) * S1X * S1Y + (lyj + RY) * S1X + (lx + RX);
        int64_t gw = 0;
        v3 m1 = mk3(0.0, 0.0, 0.0);
        if (fused_site<F>(a, gx, y0 + lyj, gz, gw)) {
            const int c0 = (lz + 2 * RZ) * S0X * S0Y + (lyj + 2 * RY) * S0X + (lx + 2 * RX);
            v3 m0;
            double ee;
            fused_stage1<F>(a, L0, c0, true, gw, m0, D1[j], m1, ee);
            if (active[j]) e[j] = ee;
            if constexpr (NOISE) {
                // Langevin term of the last stage on m0 (HEUN2: thermal_term, euler_step)
                if (active[j]) q[j] = thermal_increment(llg_rhs(m0, th[j], a.alpha, a.den, a.rden), a.dt);
            }
        }
        L1[s1] = m1.x;
        L1[T1 + s1] = m1.y;
        L1[2 * T1 + s1] = m1.z;
    }
    // energy partials of m0, one per sub-tile of NB sites (mj_block_reduce's
    // order: the sub-tile's wavefronts in lane order, then in wave order)
    if (a.energy) {
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            double v = e[j];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
            if ((threadIdx.x & 63) == 0) red[j * (NTHR / 64) + (threadIdx.x >> 6)] = v;
        }
    }
    // pass B: stage 1 on the ring of S1 (kept apart from pass A in the
    // schedule: overlapping them doubles the live registers)
    __builtin_amdgcn_sched_barrier(0);
    for (int k = threadIdx.x; k < NRING; k += NTHR) {
        int sx, sy, sz;
        fused_ring_pos<F>(k, sx, sy, sz);
        const int s1 = sz * S1X * S1Y + sy * S1X + sx;
        int64_t gw = 0;
        v3 m1 = mk3(0.0, 0.0, 0.0);
        if (fused_site<F>(a, x0 - RX + sx, y0 - RY + sy, z0 - RZ + sz, gw)) {
            const int c0 = (sz + RZ) * S0X * S0Y + (sy + RY) * S0X + (sx + RX);
            v3 m0, d;
            double ee;
            fused_stage1<F>(a, L0, c0, true, gw, m0, d, m1, ee);
        }
        L1[s1] = m1.x;
        L1[T1 + s1] = m1.y;
        L1[2 * T1 + s1] = m1.z;
    }
    __syncthreads();
    // sub-tile h (the two-kernel path's tile (tx, ty * NSUB + h, tz)): its
    // waves' sums are red[h NBW ...]; one thread per sub-tile adds them in order
    const int hsub = SPT == 1 ? (int)threadIdx.x / NB : (int)threadIdx.x;
    if (a.energy && (SPT == 1 ? (threadIdx.x % NB) == 0 : hsub < SPT)) {
        int sub_ty, sub_tz;
        if constexpr (F::kSubAxis == 1) {
            sub_ty = ty * F::kNsub + hsub;
            sub_tz = tz;
        } else {
            sub_ty = ty;
            sub_tz = tz * F::kNsub + hsub;
        }
        if (sub_ty < F::kNtyBase && sub_tz < F::kNtzBase) {
            double r = red[hsub * NBW];
            for (int w = 1; w < NBW; ++w) r = r + red[hsub * NBW + w];
            a.epart[tx + F::kNtx * (sub_ty + F::kNtyBase * sub_tz)] = r;
        }
    }

    // stage 2 at the own sites from S1
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        if (active[j]) {
            const int s1 = (lz + RZ) * S1X * S1Y + (ly + j * BYS + RY) * S1X + (lx + RX);
            const v3 m = mk3(L1[s1], L1[T1 + s1], L1[2 * T1 + s1]);
            double ax = 0.0, ay = 0.0, az = 0.0;
            F::stencil1(L1, s1, ax, ay, az);
            const v3 z = fused_zeeman<F>(a, true, gia[j]);
            const double hx = 0.0, hy = 0.0, hz = 0.0;
            const v3 B = mk3(-(ax + z.x) - 2.0 * hx, -(ay + z.y) - 2.0 * hy, -(az + z.z) - 2.0 * hz);
            if (a.fld_out) {   // the last stage's field, kept for the host
                a.fld_out[gia[j]] = B.x;
                a.fld_out[gia[j] + np] = B.y;
                a.fld_out[gia[j] + 2 * np] = B.z;
            }
            const v3 D2 = llg_rhs(m, B, a.alpha, a.den, a.rden);
            const v3 dint = mk3(0.5 * D1[j].x + 0.5 * D2.x, 0.5 * D1[j].y + 0.5 * D2.y, 0.5 * D1[j].z + 0.5 * D2.z);
            v3 o;
            if constexpr (NOISE) o = euler_step_q(m, dint, q[j], a.dt);
            else o = euler_step0(m, dint, a.dt);
            const int64_t g = gia[j];
            a.dst[g] = o.x;
            a.dst[g + np] = o.y;
            a.dst[g + 2 * np] = o.z;
        }
    }

    // workgroup 0 folds the previous launch's partials (the energy of the state
    // entering the previous step) as the stage-2 kernel's block 0 folds them:
    // strided sums over its first NB threads, then their wavefronts in order
    if (a.ered_in && blockIdx.x == 0) {
        __syncthreads();
        double v = 0.0;
        if (threadIdx.x < NB)
            for (int i = threadIdx.x; i < a.n_ered; i += NB) v += a.ered_in[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = red[0];
            for (int w = 1; w < NBW; ++w) r = r + red[w];
            *a.e_out = r * a.e_scale;
        }
    }
}

// Both stages of SIB (integration 5) or of Depondt (integration 6) in one
// kernel, the same regions as fused_heun_body.  SIB: stage 1 (SIB1, the
// midpoint p = (m + cay(m, a1(m)))/2) runs on every S1 site from S0 and its p
// goes to LDS; stage 2 (SIB2, m' = cay(m, a2(p))) runs on the tile from S1.
// Depondt: stage 1 (DEP1) rotates m about its field b1 into the predictor,
// which goes to LDS (the own sites keep b1); stage 2 (DEP2) rotates m about
// (b1 + b2)/2, b2 the predictor's field.  Both stages add the same thermal
// field (Philox of the site and step), so the S1 ring draws it too.  Every
// operation is the two-kernel path's, so m' equals it bit for bit.
template <bool NOISE, class F>
__device__ __forceinline__ v3 fused_sib_thermal(const KArgs &a, uint64_t nstep, int64_t gi)
{
    if constexpr (NOISE) {
        const v3 g = normals3(a.seed, nstep, (uint32_t)gi);
        return mk3(a.sigma * g.x, a.sigma * g.y, a.sigma * g.z);
    }
    return mk3(0.0, 0.0, 0.0);
}

// stage 1 at one site from S0: SIB's midpoint or Depondt's predictor p (and
// its field b1), the energy term of m0
template <bool DEP, bool NOISE, class F>
__device__ __forceinline__ void fused_sib_stage1(const KArgs &a, const double *L0, int c0, int64_t gi, v3 bt, v3 &m0,
                                                 v3 &p, v3 &b1, double &e)
{
    constexpr int T0 = F::kT0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    m0 = mk3(L0[c0], L0[T0 + c0], L0[2 * T0 + c0]);
    F::stencil0(L0, c0, ax, ay, az);
    const v3 z = fused_zeeman<F>(a, true, gi);
    const double hx = 0.0, hy = 0.0, hz = 0.0;
    v3 B = mk3(-(ax + z.x) - 2.0 * hx, -(ay + z.y) - 2.0 * hy, -(az + z.z) - 2.0 * hz);
    e = m0.x * ((0.5 * ax + z.x) + hx) + m0.y * ((0.5 * ay + z.y) + hy) + m0.z * ((0.5 * az + z.z) + hz);
    if constexpr (NOISE) B = mk3(B.x + bt.x, B.y + bt.y, B.z + bt.z);
    if constexpr (DEP) {
        b1 = dep_field(m0, B, a.alpha, a.den, a.rden);
        p = dep_rotate(m0, b1, a.dt);
    } else {
        const v3 av = sib_vec(m0, B, a.alpha, a.dt);
        const v3 mt = cayley(m0, av);
        p = mk3(0.5 * (m0.x + mt.x), 0.5 * (m0.y + mt.y), 0.5 * (m0.z + mt.z));
        b1 = mk3(0.0, 0.0, 0.0);
    }
}

template <bool DEP, bool NOISE, class F>
__device__ __forceinline__ void fused_sib_body(const KArgs &a, double *lds)
{
    constexpr int BX = F::kBx, BY = F::kBy, BZ = F::kBz, RX = F::kRx, RY = F::kRy, RZ = F::kRz;
    constexpr int SPT = F::kSpt, NTHR = F::kNthr, NB = F::kNthrBase, NBW = NB / 64;
    constexpr int BYS = BY / SPT;
    constexpr int S0X = F::kS0x, S0Y = F::kS0y;
    constexpr int S1X = BX + 2 * RX, S1Y = BY + 2 * RY, S1Z = BZ + 2 * RZ, T1 = S1X * S1Y * S1Z;
    constexpr int NRING = T1 - BX * BY * BZ;
    static_assert(NB % 64 == 0 && NTHR * SPT == BX * BY * BZ, "whole wavefronts, every tile site once");
    static_assert(SPT == 1 || (SPT == F::kNsub && NTHR == NB && BZ == 1), "one site per energy sub-tile");
    constexpr int T0 = F::kT0;
    double *L0 = lds;
    double *L1 = lds + 3 * T0;
    double *red = L1 + 3 * T1;
    const int64_t np = a.np;

    const int bid = mj_xcd_remap(blockIdx.x, F::kNblocks);
    const int tx = bid % F::kNtx, tr = bid / F::kNtx, ty = tr % F::kNty, tz = tr / F::kNty;
    const int x0 = tx * BX, y0 = ty * BY, z0 = tz * BZ;
    uint64_t nstep = 0;
    if constexpr (NOISE) nstep = a.stepp ? *a.stepp + a.step : a.step;
    const bool direct = F::kDma && (((unsigned long long)a.src) & 15ull) == 0 && x0 - 2 * RX >= 0 &&
                        x0 + BX + 2 * RX <= F::kNx && y0 - 2 * RY >= 0 && y0 + BY + 2 * RY <= F::kNy &&
                        z0 - 2 * RZ >= 0 && z0 + BZ + 2 * RZ <= F::kNz;
    fused_s0_fill<F>(a, L0, x0, y0, z0, direct);

    const int lx = threadIdx.x % BX, ly = (threadIdx.x / BX) % BYS, lz = threadIdx.x / (BX * BYS);
    const int gx = x0 + lx, gz = z0 + lz;
    bool active[SPT];
    int64_t gia[SPT];
    v3 th[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int gy = y0 + ly + j * BYS;
        active[j] = gx < F::kNx && gy < F::kNy && gz < F::kNz;
        gia[j] = active[j] ? (int64_t)gx + (int64_t)F::kNx * ((int64_t)gy + (int64_t)F::kNy * gz) : 0;
        th[j] = mk3(0.0, 0.0, 0.0);
    }
    if (direct) __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    // stage 1 at the own positions; the own sites keep m0, the thermal field
    // and (Depondt) the predictor's field b1
    v3 m0s[SPT], b1s[SPT];
    double e[SPT];
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        const int lyj = ly + j * BYS;
        m0s[j] = mk3(0.0, 0.0, 0.0);
        b1s[j] = mk3(0.0, 0.0, 0.0);
        e[j] = 0.0;
        const int s1 = (lz + RZ) * S1X * S1Y + (lyj + RY) * S1X + (lx + RX);
        int64_t gw = 0;
        v3 p = mk3(0.0, 0.0, 0.0);
        if (fused_site<F>(a, gx, y0 + lyj, gz, gw)) {
            const int c0 = (lz + 2 * RZ) * S0X * S0Y + (lyj + 2 * RY) * S0X + (lx + 2 * RX);
            const v3 bt = fused_sib_thermal<NOISE, F>(a, nstep, gw);
            double ee;
            v3 m0, b1;
            fused_sib_stage1<DEP, NOISE, F>(a, L0, c0, gw, bt, m0, p, b1, ee);
            if (active[j]) {
                e[j] = ee;
                m0s[j] = m0;
                b1s[j] = b1;
                th[j] = bt;
            }
        }
        L1[s1] = p.x;
        L1[T1 + s1] = p.y;
        L1[2 * T1 + s1] = p.z;
    }
    if (a.energy) {
#pragma unroll
        for (int j = 0; j < SPT; ++j) {
            double v = e[j];
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
            if ((threadIdx.x & 63) == 0) red[j * (NTHR / 64) + (threadIdx.x >> 6)] = v;
        }
    }
    // stage 1 on the ring of S1
    __builtin_amdgcn_sched_barrier(0);
    for (int k = threadIdx.x; k < NRING; k += NTHR) {
        int sx, sy, sz;
        fused_ring_pos<F>(k, sx, sy, sz);
        const int s1 = sz * S1X * S1Y + sy * S1X + sx;
        int64_t gw = 0;
        v3 p = mk3(0.0, 0.0, 0.0);
        if (fused_site<F>(a, x0 - RX + sx, y0 - RY + sy, z0 - RZ + sz, gw)) {
            const int c0 = (sz + RZ) * S0X * S0Y + (sy + RY) * S0X + (sx + RX);
            const v3 bt = fused_sib_thermal<NOISE, F>(a, nstep, gw);
            v3 m0, b1;
            double ee;
            fused_sib_stage1<DEP, NOISE, F>(a, L0, c0, gw, bt, m0, p, b1, ee);
        }
        L1[s1] = p.x;
        L1[T1 + s1] = p.y;
        L1[2 * T1 + s1] = p.z;
    }
    __syncthreads();
    const int hsub = SPT == 1 ? (int)threadIdx.x / NB : (int)threadIdx.x;
    if (a.energy && (SPT == 1 ? (threadIdx.x % NB) == 0 : hsub < SPT)) {
        int sub_ty, sub_tz;
        if constexpr (F::kSubAxis == 1) {
            sub_ty = ty * F::kNsub + hsub;
            sub_tz = tz;
        } else {
            sub_ty = ty;
            sub_tz = tz * F::kNsub + hsub;
        }
        if (sub_ty < F::kNtyBase && sub_tz < F::kNtzBase) {
            double r = red[hsub * NBW];
            for (int w = 1; w < NBW; ++w) r = r + red[hsub * NBW + w];
            a.epart[tx + F::kNtx * (sub_ty + F::kNtyBase * sub_tz)] = r;
        }
    }

    // stage 2 at the own sites from S1
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < SPT; ++j) {
        if (active[j]) {
            const int s1 = (lz + RZ) * S1X * S1Y + (ly + j * BYS + RY) * S1X + (lx + RX);
            const v3 p = mk3(L1[s1], L1[T1 + s1], L1[2 * T1 + s1]);
            double ax = 0.0, ay = 0.0, az = 0.0;
            F::stencil1(L1, s1, ax, ay, az);
            const v3 z = fused_zeeman<F>(a, true, gia[j]);
            const double hx = 0.0, hy = 0.0, hz = 0.0;
            v3 B = mk3(-(ax + z.x) - 2.0 * hx, -(ay + z.y) - 2.0 * hy, -(az + z.z) - 2.0 * hz);
            if (a.fld_out) {   // the last stage's field, kept for the host
                a.fld_out[gia[j]] = B.x;
                a.fld_out[gia[j] + np] = B.y;
                a.fld_out[gia[j] + 2 * np] = B.z;
            }
            if constexpr (NOISE) B = mk3(B.x + th[j].x, B.y + th[j].y, B.z + th[j].z);
            v3 o;
            if constexpr (DEP) {
                const v3 b2 = dep_field(p, B, a.alpha, a.den, a.rden);
                const v3 b1 = b1s[j];
                const v3 ba = mk3(0.5 * (b1.x + b2.x), 0.5 * (b1.y + b2.y), 0.5 * (b1.z + b2.z));
                o = dep_rotate(m0s[j], ba, a.dt);
            } else {
                const v3 av = sib_vec(p, B, a.alpha, a.dt);
                o = cayley(m0s[j], av);
            }
            const int64_t g = gia[j];
            a.dst[g] = o.x;
            a.dst[g + np] = o.y;
            a.dst[g + 2 * np] = o.z;
        }
    }

    // workgroup 0 folds the previous launch's partials (fused_heun_body's fold)
    if (a.ered_in && blockIdx.x == 0) {
        __syncthreads();
        double v = 0.0;
        if (threadIdx.x < NB)
            for (int i = threadIdx.x; i < a.n_ered; i += NB) v += a.ered_in[i];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v = v + __shfl_xor(v, o, 64);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = red[0];
            for (int w = 1; w < NBW; ++w) r = r + red[w];
            *a.e_out = r * a.e_scale;
        }
    }
}
