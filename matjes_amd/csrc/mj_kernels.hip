// Ahead-of-time kernels of the LLG hot path: the fused stage kernel with
// runtime geometry (used for small lattices and whenever no per-Hamiltonian
// kernel is compiled), plus the small helper kernels.
#include "mj_device.h"
#include "mj_internal.h"
#include "mj_stage_body.h"

template <int MODE, bool NOISE>
__global__ void __launch_bounds__(1024) k_stage(const KArgs a)
{
    extern __shared__ double lds[];
    stage_body<MODE, NOISE, RtGeom>(a, lds);
}

// energy partials -> E (one block) when no following kernel folds them
__global__ void k_reduce(const double *part, int n, double *out, double scale)
{
    __shared__ double red[1024];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += part[i];
    double s = mj_block_sum(v, red, blockDim.x);
    if (threadIdx.x == 0) *out = s * scale;
}

// the device's own arithmetic against the library operations (mj_math_probe):
// fn 0 mj_sqrt_rn, 1 sqrt, 2 mj_rcp_rn, 3 1.0 / x
__global__ void k_math_probe(int fn, int64_t n, const double *x, double *y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i];
    double r;
    switch (fn) {
    case 0: r = mj_sqrt_rn(v); break;
    case 1: r = sqrt(v); break;
    case 2: r = mj_rcp_rn(v); break;
    default: r = 1.0 / v; break;
    }
    y[i] = r;
}

int mj_launch_math_probe(int fn, int64_t n, const double *x, double *y, hipStream_t s)
{
    hipLaunchKernelGGL(k_math_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, fn, n, x, y);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("math probe launch: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    return MJ_OK;
}

// energy partials -> E with the fold of a stage kernel's block 0 (blockDim =
// the stage kernels' threads: the same strided sums and block reduction)
__global__ void k_fold(const double *part, int n, double *out, double scale)
{
    extern __shared__ double red[];
    // the same sums in the same order, eight loads in flight at a time
    double v = 0.0;
    const int nt = blockDim.x;
    int i = threadIdx.x;
    for (; i + 7 * nt < n; i += 8 * nt) {
        double q[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] = part[i + k * nt];
#pragma unroll
        for (int k = 0; k < 8; ++k) v += q[k];
    }
    for (; i < n; i += nt) v += part[i];
    double s = mj_block_sum(v, red, blockDim.x);
    if (threadIdx.x == 0) *out = s * scale;
}

// per-block maxima -> max (one block)
__global__ void k_reduce_max(const double *part, int n, double *out)
{
    __shared__ double red[1024];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v = fmax(v, part[i]);
    double s = mj_block_max(v, red, blockDim.x);
    if (threadIdx.x == 0) *out = s;
}

// device step counter of graph-replayed steps: set, or advance by n
__global__ void k_step_counter(uint64_t *p, uint64_t v, int add)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = add ? *p + v : v;
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
        MJ_PICK(MODE_RKS)
    }
#undef MJ_PICK
    return nullptr;
}

// energy of one term at every site, summed per block:
//   mode 0  m_i . sum_k A_k m_{i+d_k}      (bonds leaving an open boundary dropped)
//   mode 1  m_i . (Z^T b_i)
//   mode 2  m_i . h_i                       (H_fft m on the padded grid)
__global__ void __launch_bounds__(256) k_term_energy(const TermArgs t)
{
    __shared__ double red[256];
    const int64_t n = (int64_t)t.dim[0] * t.dim[1] * t.dim[2];
    double e = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t np = t.np;
        const int x = (int)(i % t.dim[0]);
        const int64_t r = i / t.dim[0];
        const int y = (int)(r % t.dim[1]);
        const int z = (int)(r / t.dim[1]);
        const double mx = t.m[i], my = t.m[i + np], mz = t.m[i + 2 * np];
        double fx = 0.0, fy = 0.0, fz = 0.0;
        if (t.mode == 0) {
            for (int k = 0; k < t.nofs; ++k) {
                int q[3] = {x + t.ofs3[3 * k], y + t.ofs3[3 * k + 1], z + t.ofs3[3 * k + 2]};
                bool in = true;
                for (int d = 0; d < 3; ++d) {
                    if (q[d] < 0 || q[d] >= t.dim[d]) {
                        if (!t.per[d]) in = false;
                        q[d] = ((q[d] % t.dim[d]) + t.dim[d]) % t.dim[d];
                    }
                }
                if (!in) continue;
                const int64_t j = (int64_t)q[0] + (int64_t)t.dim[0] * ((int64_t)q[1] + (int64_t)t.dim[1] * q[2]);
                const double nx = t.m[j], ny = t.m[j + np], nz = t.m[j + 2 * np];
                const double *A = t.blk9 + 9 * k;
                fx += (A[0] * nx + A[1] * ny) + A[2] * nz;
                fy += (A[3] * nx + A[4] * ny) + A[5] * nz;
                fz += (A[6] * nx + A[7] * ny) + A[8] * nz;
            }
        } else if (t.mode == 1) {
            double bx = t.b[0], by = t.b[1], bz = t.b[2];
            if (t.bmode == 2) {
                bx = t.bmap[i];
                by = t.bmap[i + np];
                bz = t.bmap[i + 2 * np];
            }
            fx = (bx * t.z[0] + by * t.z[3]) + bz * t.z[6];
            fy = (bx * t.z[1] + by * t.z[4]) + bz * t.z[7];
            fz = (bx * t.z[2] + by * t.z[5]) + bz * t.z[8];
        } else {
            const int64_t h = (int64_t)x + (int64_t)t.dnx * ((int64_t)y + (int64_t)t.dny * z);
            fx = t.hd[h];
            fy = t.hd[h + t.hstride];
            fz = t.hd[h + 2 * t.hstride];
        }
        e += (mx * fx + my * fy) + mz * fz;
    }
    const double s = mj_block_sum(e, red, blockDim.x);
    if (threadIdx.x == 0) t.part[blockIdx.x] = s;
}

// product of one registered term with a vector (AoS), the backend mult_r / mult_l
//   H v:    out_i = sum_k A_k v_{i+d_k}        H^T v:  out_i = sum_k A_k^T v_{i-d_k}
// (bonds leaving an open boundary absent); Zeeman: out_i = Z v_i or Z^T v_i
__global__ void __launch_bounds__(256) k_term_mult(const TermArgs t)
{
    const int64_t n = (int64_t)t.dim[0] * t.dim[1] * t.dim[2];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double r[3] = {0.0, 0.0, 0.0};
    if (t.mode == 3) {
        const int x = (int)(i % t.dim[0]);
        const int64_t rr = i / t.dim[0];
        const int y = (int)(rr % t.dim[1]);
        const int z = (int)(rr / t.dim[1]);
        const int sg = t.side == 0 ? 1 : -1;
        for (int k = 0; k < t.nofs; ++k) {
            int q[3] = {x + sg * t.ofs3[3 * k], y + sg * t.ofs3[3 * k + 1], z + sg * t.ofs3[3 * k + 2]};
            bool in = true;
            for (int d = 0; d < 3; ++d) {
                if (q[d] < 0 || q[d] >= t.dim[d]) {
                    if (!t.per[d]) in = false;
                    q[d] = ((q[d] % t.dim[d]) + t.dim[d]) % t.dim[d];
                }
            }
            if (!in) continue;
            const int64_t j = (int64_t)q[0] + (int64_t)t.dim[0] * ((int64_t)q[1] + (int64_t)t.dim[1] * q[2]);
            const double *A = t.blk9 + 9 * k;
            const double v0 = t.vin[3 * j], v1 = t.vin[3 * j + 1], v2 = t.vin[3 * j + 2];
            for (int a = 0; a < 3; ++a) {
                if (t.side == 0)
                    r[a] += (A[3 * a] * v0 + A[3 * a + 1] * v1) + A[3 * a + 2] * v2;
                else
                    r[a] += (A[a] * v0 + A[3 + a] * v1) + A[6 + a] * v2;
            }
        }
    } else {
        const double v0 = t.vin[3 * i], v1 = t.vin[3 * i + 1], v2 = t.vin[3 * i + 2];
        for (int a = 0; a < 3; ++a) {
            if (t.side == 0)
                r[a] = (t.z[3 * a] * v0 + t.z[3 * a + 1] * v1) + t.z[3 * a + 2] * v2;
            else
                r[a] = (t.z[a] * v0 + t.z[3 + a] * v1) + t.z[6 + a] * v2;
        }
    }
    for (int a = 0; a < 3; ++a)
        t.vout[3 * i + a] = t.beta == 0.0 ? t.alpha * r[a] : t.alpha * r[a] + t.beta * t.vout[3 * i + a];
}

int mj_launch_term_mult(const TermArgs &t, int64_t n, hipStream_t s)
{
    hipLaunchKernelGGL(k_term_mult, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("term product launch: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    return MJ_OK;
}

int mj_launch_term_energy(const TermArgs &t, int nblocks, hipStream_t s)
{
    hipLaunchKernelGGL(k_term_energy, dim3(nblocks), dim3(256), 0, s, t);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("term energy launch: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    return MJ_OK;
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

int mj_launch_fold(const double *part, int n, double *out, double scale, int nthr, hipStream_t s)
{
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(nthr), sizeof(double) * nthr, s, part, n, out, scale);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("fold launch: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    return MJ_OK;
}

int mj_launch_reduce_max(const double *part, int n, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(k_reduce_max, dim3(1), dim3(1024), 0, s, part, n, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("reduce launch: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    return MJ_OK;
}

int mj_launch_step_counter(uint64_t *p, uint64_t v, int add, hipStream_t s)
{
    hipLaunchKernelGGL(k_step_counter, dim3(1), dim3(64), 0, s, p, v, add);
    return hipGetLastError() == hipSuccess ? MJ_OK : MJ_EHIP;
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
