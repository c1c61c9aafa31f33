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
