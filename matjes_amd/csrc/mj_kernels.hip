// Stage kernels of the LLG hot path for gfx950.
//
// One kernel evaluates B_eff = -dH/dm at every site of a tile and applies one
// integrator stage in registers, so a stage is one pass over HBM:
//   - the state (SoA mx|my|mz, FP64) of the tile plus a halo of the stencil
//     reach is staged in LDS (open boundaries are zero-filled, periodic ones
//     wrapped), then every site sums its translation-invariant 3x3 blocks
//     from LDS (isotropic / isotropic+antisymmetric / general groups);
//   - Zeeman (uniform or per-site B order parameter) and the dipolar H_fft
//     (read from the FFT output grid) are added in the epilogue;
//   - the stage (reference Euler / two-half-step Heun, SIB, Depondt) and the
//     in-register Philox thermal field finish the site;
//   - optionally the energy of the stencil state is reduced per block, and
//     block 0 folds the previous kernel's per-block partials into E.
// Reference: field hamiltonian_collection.f90:433-454; stage loop
// spindynamics.f90:260-323; energy hamiltonian_collection.f90:365-400.
#include "mj_device.h"
#include "mj_internal.h"

namespace {

__device__ __forceinline__ bool wrap_idx(int &g, int n, int periodic)
{
    if (g >= 0 && g < n) return true;
    if (!periodic) return false;
    g %= n;
    if (g < 0) g += n;
    return true;
}

// workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous band of tiles so halos are shared inside one L2
__device__ __forceinline__ int xcd_remap(int bid, int nb)
{
    if ((nb & 7) != 0) return bid;
    return (bid & 7) * (nb >> 3) + (bid >> 3);
}

__device__ __forceinline__ double block_sum(double v, double *red)
{
    // deterministic tree over the block (blockDim.x <= 1024, multiple of 64 or full block)
    int t = threadIdx.x;
    red[t] = v;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (s < (int)blockDim.x && t < s && t + s < (int)blockDim.x) red[t] = red[t] + red[t + s];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

}  // namespace

template <int MODE, bool NOISE>
__global__ void __launch_bounds__(1024) k_stage(const KArgs a)
{
    extern __shared__ double lds[];
    const int nthr = blockDim.x;
    const int bid = xcd_remap(blockIdx.x, a.nblocks);
    const int tx = bid % a.ntx;
    const int tr = bid / a.ntx;
    const int ty = tr % a.nty;
    const int tz = tr / a.nty;
    const int x0 = tx * a.bx, y0 = ty * a.by, z0 = tz * a.bz;
    const int T = a.sx * a.sy * a.sz;
    const int64_t np = a.np;

    if (a.use_stencil) {
        for (int t = threadIdx.x; t < T; t += nthr) {
            int lx = t % a.sx;
            int r = t / a.sx;
            int ly = r % a.sy;
            int lz = r / a.sy;
            int gx = x0 + lx - a.rx, gy = y0 + ly - a.ry, gz = z0 + lz - a.rz;
            double vx = 0.0, vy = 0.0, vz = 0.0;
            if (wrap_idx(gx, a.nx, a.px) && wrap_idx(gy, a.ny, a.py) && wrap_idx(gz, a.nz, a.pz)) {
                int64_t g = (int64_t)gx + (int64_t)a.nx * ((int64_t)gy + (int64_t)a.ny * gz);
                vx = a.src[g];
                vy = a.src[g + np];
                vz = a.src[g + 2 * np];
            }
            lds[t] = vx;
            lds[T + t] = vy;
            lds[2 * T + t] = vz;
        }
        __syncthreads();
    }

    const int lx = threadIdx.x % a.bx;
    const int ly = (threadIdx.x / a.bx) % a.by;
    const int lz = threadIdx.x / (a.bx * a.by);
    const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
    const bool active = gx < a.nx && gy < a.ny && gz < a.nz;
    const int64_t gi = active ? (int64_t)gx + (int64_t)a.nx * ((int64_t)gy + (int64_t)a.ny * gz) : 0;

    double ax = 0.0, ay = 0.0, az = 0.0;   // sum_d C_d m_{i+d}  (= (H + H^T) m at i)
    v3 m;                                  // state at the own site
    if (a.use_stencil) {
        const int c = (lz + a.rz) * a.sx * a.sy + (ly + a.ry) * a.sx + (lx + a.rx);
        m = mk3(lds[c], lds[T + c], lds[2 * T + c]);
        for (int k = 0; k < a.n_iso; ++k) {
            const int j = c + a.ofs[k];
            const double s = a.cf[k];
            ax = fma(s, lds[j], ax);
            ay = fma(s, lds[T + j], ay);
            az = fma(s, lds[2 * T + j], az);
        }
        const double *cd = a.cf + a.n_iso;
        for (int k = 0; k < a.n_dm; ++k) {
            const int j = c + a.ofs[a.n_iso + k];
            const double s = cd[4 * k], wx = cd[4 * k + 1], wy = cd[4 * k + 2], wz = cd[4 * k + 3];
            const double mx = lds[j], my = lds[T + j], mz = lds[2 * T + j];
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
            const double mx = lds[j], my = lds[T + j], mz = lds[2 * T + j];
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
    } else if (a.bmode == 2 && active) {
        const double bxv = a.bmap[gi], byv = a.bmap[gi + np], bzv = a.bmap[gi + 2 * np];
        zx = (bxv * a.z[0] + byv * a.z[3]) + bzv * a.z[6];
        zy = (bxv * a.z[1] + byv * a.z[4]) + bzv * a.z[7];
        zz = (bxv * a.z[2] + byv * a.z[5]) + bzv * a.z[8];
    }
    // dipolar H_fft at this site (padded grid keeps the lattice at its origin)
    double hx = 0.0, hy = 0.0, hz = 0.0;
    if (a.hdip && active) {
        const int64_t h = (int64_t)gx + (int64_t)a.dnx * ((int64_t)gy + (int64_t)a.dny * gz);
        hx = a.hdip[h];
        hy = a.hdip[h + a.hstride];
        hz = a.hdip[h + 2 * a.hstride];
    }
    // B = -(H m + H^T m + Z^T b) - 2 H_fft      (hamiltonian_collection.f90:442-453)
    v3 B = mk3(-(ax + zx) - 2.0 * hx, -(ay + zy) - 2.0 * hy, -(az + zz) - 2.0 * hz);

    // energy of the stencil state: m.(H m) + b.(Z m) + m.H_fft = m.(acc/2 + Z^T b + h)
    if (a.energy) {
        double e = 0.0;
        if (active)
            e = m.x * ((0.5 * ax + zx) + hx) + m.y * ((0.5 * ay + zy) + hy) + m.z * ((0.5 * az + zz) + hz);
        __syncthreads();
        double s = block_sum(e, lds);
        if (threadIdx.x == 0) a.epart[blockIdx.x] = s;
    }
    // block 0 folds the previous kernel's partials into E (stream order makes them complete)
    if (a.ered_in && blockIdx.x == 0) {
        __syncthreads();
        double v = 0.0;
        for (int i = threadIdx.x; i < a.n_ered; i += nthr) v += a.ered_in[i];
        double s = block_sum(v, lds);
        if (threadIdx.x == 0) *a.e_out = s * a.e_scale;
    }
    if (!active) return;

    const double alpha = a.alpha;
    if constexpr (MODE == MODE_FIELD) {
        if (a.fld_out) {
            a.fld_out[gi] = B.x;
            a.fld_out[gi + np] = B.y;
            a.fld_out[gi + 2 * np] = B.z;
        }
    } else if constexpr (MODE == MODE_EULER) {
        v3 D = llg_rhs(m, B, alpha, a.den);
        v3 dint = mk3(0.0 + 1.0 * D.x, 0.0 + 1.0 * D.y, 0.0 + 1.0 * D.z);
        v3 DT = mk3(0.0, 0.0, 0.0);
        if constexpr (NOISE) {
            v3 bt;
            if (a.noise) bt = mk3(a.noise[gi], a.noise[gi + np], a.noise[gi + 2 * np]);
            else {
                v3 g = normals3(a.seed, a.step, (uint32_t)gi);
                bt = mk3(a.sigma * g.x, a.sigma * g.y, a.sigma * g.z);
            }
            DT = llg_rhs(m, bt, alpha, a.den);
        }
        v3 o = euler_step(m, dint, DT, a.dt);
        a.dst[gi] = o.x;
        a.dst[gi + np] = o.y;
        a.dst[gi + 2 * np] = o.z;
    } else if constexpr (MODE == MODE_HEUN1) {
        // stage 1 of integration 2: D1 = LLG(B(m0), m0); m1 = euler(m0, 1*D1 + 0*D2, 0, dt/2)
        v3 D = llg_rhs(m, B, alpha, a.den);
        v3 dint = mk3(0.0 + 1.0 * D.x, 0.0 + 1.0 * D.y, 0.0 + 1.0 * D.z);
        v3 o = euler_step(m, dint, mk3(0.0, 0.0, 0.0), a.dt);
        a.dst[gi] = o.x;
        a.dst[gi + np] = o.y;
        a.dst[gi + 2 * np] = o.z;
        a.aux[gi] = D.x;
        a.aux[gi + np] = D.y;
        a.aux[gi + 2 * np] = D.z;
    } else if constexpr (MODE == MODE_HEUN2) {
        // stage 2: D2 = LLG(B(m1), m1); m2 = euler(m1, 0.5 D1 + 0.5 D2, DT(m0), dt/2)
        v3 D2 = llg_rhs(m, B, alpha, a.den);
        v3 D1 = mk3(a.aux[gi], a.aux[gi + np], a.aux[gi + 2 * np]);
        v3 dint = mk3((0.0 + 0.5 * D1.x) + 0.5 * D2.x, (0.0 + 0.5 * D1.y) + 0.5 * D2.y,
                      (0.0 + 0.5 * D1.z) + 0.5 * D2.z);
        v3 DT = mk3(0.0, 0.0, 0.0);
        if constexpr (NOISE) {
            v3 m0 = mk3(a.own[gi], a.own[gi + np], a.own[gi + 2 * np]);
            v3 bt;
            if (a.noise) bt = mk3(a.noise[gi], a.noise[gi + np], a.noise[gi + 2 * np]);
            else {
                v3 g = normals3(a.seed, a.step, (uint32_t)gi);
                bt = mk3(a.sigma * g.x, a.sigma * g.y, a.sigma * g.z);
            }
            DT = llg_rhs(m0, bt, alpha, a.den);
        }
        v3 o = euler_step(m, dint, DT, a.dt);
        a.dst[gi] = o.x;
        a.dst[gi + np] = o.y;
        a.dst[gi + 2 * np] = o.z;
    } else if constexpr (MODE == MODE_SIB1 || MODE == MODE_SIB2 || MODE == MODE_DEP1 ||
                         MODE == MODE_DEP2) {
        if constexpr (NOISE) {
            v3 bt;
            if (a.noise) bt = mk3(a.noise[gi], a.noise[gi + np], a.noise[gi + 2 * np]);
            else {
                v3 g = normals3(a.seed, a.step, (uint32_t)gi);
                bt = mk3(a.sigma * g.x, a.sigma * g.y, a.sigma * g.z);
            }
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
            v3 m0 = mk3(a.own[gi], a.own[gi + np], a.own[gi + 2 * np]);
            v3 o = cayley(m0, av);
            a.dst[gi] = o.x;
            a.dst[gi + np] = o.y;
            a.dst[gi + 2 * np] = o.z;
        } else if constexpr (MODE == MODE_DEP1) {
            v3 b1 = dep_field(m, B, alpha);
            v3 mt = dep_rotate(m, b1, a.dt);
            a.dst[gi] = mt.x;
            a.dst[gi + np] = mt.y;
            a.dst[gi + 2 * np] = mt.z;
            a.aux[gi] = b1.x;
            a.aux[gi + np] = b1.y;
            a.aux[gi + 2 * np] = b1.z;
        } else {
            v3 b2 = dep_field(m, B, alpha);   // m is the predictor here
            v3 b1 = mk3(a.aux[gi], a.aux[gi + np], a.aux[gi + 2 * np]);
            v3 ba = mk3(0.5 * (b1.x + b2.x), 0.5 * (b1.y + b2.y), 0.5 * (b1.z + b2.z));
            v3 m0 = mk3(a.own[gi], a.own[gi + np], a.own[gi + 2 * np]);
            v3 o = dep_rotate(m0, ba, a.dt);
            a.dst[gi] = o.x;
            a.dst[gi + np] = o.y;
            a.dst[gi + 2 * np] = o.z;
        }
    }
}

// energy partials -> E (one block) when no following kernel folds them
__global__ void k_reduce(const double *part, int n, double *out, double scale)
{
    __shared__ double red[1024];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
    double s = block_sum(v, red);
    if (threadIdx.x == 0) *out = s * scale;
}

// generic path: acc = (H + H^T) m, CSR over 3N rows in AoS numbering, SoA in/out
__global__ void k_spmv_sym(int64_t n3, const int *rp, const int *ci, const double *v, const double *m,
                           int64_t np, double *out)
{
    int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n3) return;
    double acc = 0.0;
    for (int k = rp[r]; k < rp[r + 1]; ++k) {
        int c = ci[k];
        acc = fma(v[k], m[(int64_t)(c / 3) + (c % 3) * np], acc);
    }
    out[(r / 3) + (r % 3) * np] = acc;
}

__global__ void k_aos2soa(const double *in, double *out, int64_t n, int64_t np)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = in[3 * i];
    out[i + np] = in[3 * i + 1];
    out[i + 2 * np] = in[3 * i + 2];
}

__global__ void k_soa2aos(const double *in, double *out, int64_t n, int64_t np, double eps)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // truncate on output (src/vector_utils/precision.f90:39-51); the carried state is untruncated
    double x = in[i], y = in[i + np], z = in[i + 2 * np];
    out[3 * i] = fabs(x) < eps ? 0.0 : x;
    out[3 * i + 1] = fabs(y) < eps ? 0.0 : y;
    out[3 * i + 2] = fabs(z) < eps ? 0.0 : z;
}

__global__ void k_normals(uint64_t seed, uint64_t step, int64_t n, double *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    v3 g = normals3(seed, step, (uint32_t)i);
    out[3 * i] = g.x;
    out[3 * i + 1] = g.y;
    out[3 * i + 2] = g.z;
}

// ---------------------------------------------------------------------------
// host-side launchers
// ---------------------------------------------------------------------------
typedef void (*stage_fn)(const KArgs);

static stage_fn pick(int mode, bool noise)
{
#define MJ_PICK(M)                                                 \
    case M:                                                        \
        return noise ? (stage_fn)k_stage<M, true> : (stage_fn)k_stage<M, false>;
    switch (mode) {
        MJ_PICK(MODE_FIELD)
        MJ_PICK(MODE_EULER)
        MJ_PICK(MODE_HEUN1)
        MJ_PICK(MODE_HEUN2)
        MJ_PICK(MODE_SIB1)
        MJ_PICK(MODE_SIB2)
        MJ_PICK(MODE_DEP1)
        MJ_PICK(MODE_DEP2)
    }
#undef MJ_PICK
    return nullptr;
}

int mj_launch_stage(int mode, bool noise, const KArgs &a, hipStream_t s)
{
    stage_fn f = pick(mode, noise);
    int nthr = a.bx * a.by * a.bz;
    size_t T = (size_t)a.sx * a.sy * a.sz;
    size_t shm = a.use_stencil ? 3 * T * sizeof(double) : 0;
    size_t need = (size_t)nthr * sizeof(double);
    if (shm < need) shm = need;  // block reduction scratch
    hipLaunchKernelGGL(f, dim3(a.nblocks), dim3(nthr), shm, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("stage kernel launch: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    return MJ_OK;
}

int mj_launch_reduce(const double *part, int n, double *out, double scale, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, s, part, n, out, scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("reduce launch: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    return MJ_OK;
}

int mj_launch_spmv(const mj_ctx *c, const double *m, double *out, hipStream_t s)
{
    int64_t n3 = 3 * c->n;
    hipLaunchKernelGGL(k_spmv_sym, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, s, n3, c->d_rp, c->d_ci,
                       c->d_v, m, c->np, out);
    return hipGetLastError() == hipSuccess ? MJ_OK : MJ_EHIP;
}

int mj_launch_aos2soa(const double *in, double *out, int64_t n, int64_t np, hipStream_t s)
{
    hipLaunchKernelGGL(k_aos2soa, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n, np);
    return hipGetLastError() == hipSuccess ? MJ_OK : MJ_EHIP;
}

int mj_launch_soa2aos(const double *in, double *out, int64_t n, int64_t np, double eps, hipStream_t s)
{
    hipLaunchKernelGGL(k_soa2aos, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, in, out, n, np, eps);
    return hipGetLastError() == hipSuccess ? MJ_OK : MJ_EHIP;
}

int mj_launch_normals(uint64_t seed, uint64_t step, int64_t n, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_normals, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, seed, step, n, out);
    return hipGetLastError() == hipSuccess ? MJ_OK : MJ_EHIP;
}
