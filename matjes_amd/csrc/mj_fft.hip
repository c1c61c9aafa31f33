// Dipolar (FFT) Hamiltonian on the device: rocFFT transforms around a fused
// k-space 3x3 tensor product.
//
// Same contract as the reference's fft_H (fft_ham_base.f90:191-267,
// fft_ham_fftw.f90:134-262, cuda_fft.cu:38-70):
//   K_F   = R2C(K_n) / prod(N_rep)                      (once, at set-up)
//   H_fft = C2R( H_F ),  H_F(i,k) = sum_a K_F(a,i,k) M_F(a,k),  M_F = R2C(pad(M))
// with the lattice at the origin of the N_rep grid and zero padding along
// open directions (fft_ham_internal.f90:26-124).  The field contribution
// -2 H_fft is added inside the stage kernel that consumes the grid.
// Layout: three component grids (SoA), x fastest; half spectrum along x.
//
// Films (at most 32 layers along z) take a cheaper, equal route: the
// convolution along z is a short Toeplitz sum, so only the in-plane
// directions are transformed (2D R2C/C2R of the Lz real layers, no padding
// along z) and the k-space kernel applies, for every in-plane k,
//   H_F(i, z) = sum_z' sum_a K2_F(a, i, (z - z') mod N_rep_z) M_F(a, z'),
//   K2_F(a, i, dz) = R2C_2D(K_n(a, i, :, :, dz)) / (N_rep_x N_rep_y),
// which is the same circular convolution as the 3D transform (the padded
// layers of M are zero).  The k-space kernel stages the block's M_F and
// K2_F columns in LDS.
#include <rocfft/rocfft.h>

#include <stdlib.h>

#include <algorithm>
#include <string>
#include <vector>

#include "mj_internal.h"

struct DipolarFFT {
    int nrep[3];
    int64_t nk = 0;   // real points
    int64_t nc = 0;   // complex points (half spectrum)
    int sym = 1;      // 1: six-component symmetric K_F
    double *grid = nullptr;     // H on the N_rep grid (3 * nk), the inverse transform's output
    double *grid_in = nullptr;  // M on the N_rep grid (3 * nk): zero outside the lattice from set-up on, the
                                // lattice's sites rewritten per call; none where the state is the input itself
    double2 *spec = nullptr;    // M_F / H_F (3 * nc)
    double2 *kf = nullptr;      // K_F: 6 or 9 components (nc each)
    void *work = nullptr;
    size_t work_size = 0;
    rocfft_plan fwd = nullptr, inv = nullptr;
    rocfft_execution_info info = nullptr;
    // film form (2D transforms + layer mixing)
    int layer = 0;
    int lz = 0, nzr = 0;
    int64_t np2 = 0, nc2 = 0;   // in-plane real / complex points
    double2 *spec_h = nullptr;  // H_F (3 * lz * nc2)
    int kpb = 0;                // in-plane k points per workgroup of the layer kernel
    int nzt = 0;                // layer offsets stored per component: nzr, or lz when K2_F has a z-parity
    int odd = 0;                // z-parity form: bit q set where K(q)(-dz) = -K(q)(dz) (else +K(q)(dz))
};

static int rf_check(rocfft_status s, const char *what)
{
    if (s != rocfft_status_success) {
        mj_set_error(std::string("rocFFT ") + what + " failed (status " + std::to_string((int)s) + ")");
        return MJ_EHIP;
    }
    return MJ_OK;
}
#define RF_TRY(call, what)                       \
    do {                                         \
        int r_ = rf_check((call), (what));       \
        if (r_ != MJ_OK) return r_;              \
    } while (0)

// M into the N_rep grid at the lattice's own positions (the rest of grid_in
// stays zero from set-up: the forward transform leaves its input as it is);
// one thread per lattice site (ns = nx ny nz), np the state's SoA stride
__global__ void k_place(const double *m, int64_t np, int64_t ns, int nx, int ny, int rx, int ry, int64_t nk,
                        double *grid)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    const int x = (int)(s % nx);
    const int64_t r = s / nx;
    const int y = (int)(r % ny);
    const int64_t z = r / ny;
    const int64_t p = (int64_t)x + (int64_t)rx * ((int64_t)y + (int64_t)ry * z);
    grid[p] = m[s];
    grid[p + nk] = m[s + np];
    grid[p + 2 * nk] = m[s + 2 * np];
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    double2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b)
{
    double2 r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}

// h += k m (complex)
__device__ __forceinline__ void cmac(double2 &h, double2 k, double2 m)
{
    h.x = fma(k.x, m.x, h.x);
    h.x = fma(-k.y, m.y, h.x);
    h.y = fma(k.x, m.y, h.y);
    h.y = fma(k.y, m.x, h.y);
}

// H_F(i) = sum_a K_F(a,i) M_F(a), in place over the spectrum (one thread per k)
template <bool SYM>
__global__ void k_kmul(double2 *spec, const double2 *kf, int64_t nc)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nc) return;
    double2 m0 = spec[k], m1 = spec[k + nc], m2 = spec[k + 2 * nc];
    double2 h0, h1, h2;
    if constexpr (SYM) {
        // kf: xx, xy, xz, yy, yz, zz
        double2 kxx = kf[k], kxy = kf[k + nc], kxz = kf[k + 2 * nc];
        double2 kyy = kf[k + 3 * nc], kyz = kf[k + 4 * nc], kzz = kf[k + 5 * nc];
        h0 = cadd(cadd(cmul(kxx, m0), cmul(kxy, m1)), cmul(kxz, m2));
        h1 = cadd(cadd(cmul(kxy, m0), cmul(kyy, m1)), cmul(kyz, m2));
        h2 = cadd(cadd(cmul(kxz, m0), cmul(kyz, m1)), cmul(kzz, m2));
    } else {
        // kf component a + 3*i holds K_F(a, i)
        h0 = cadd(cadd(cmul(kf[k], m0), cmul(kf[k + nc], m1)), cmul(kf[k + 2 * nc], m2));
        h1 = cadd(cadd(cmul(kf[k + 3 * nc], m0), cmul(kf[k + 4 * nc], m1)), cmul(kf[k + 5 * nc], m2));
        h2 = cadd(cadd(cmul(kf[k + 6 * nc], m0), cmul(kf[k + 7 * nc], m1)), cmul(kf[k + 8 * nc], m2));
    }
    spec[k] = h0;
    spec[k + nc] = h1;
    spec[k + 2 * nc] = h2;
}

// H_F(i, z) = sum_z' sum_a K2_F(a, i, (z - z') mod nzr) M_F(a, z') for a
// block of KPB in-plane k points; thread (kl, z), kl fastest.  The block's
// K2_F columns (ncomp x nzt) are staged in LDS; M_F is read from memory (the
// lz threads of one k share each value through L1, and the LDS left to the
// tensor alone holds more workgroups per CU: 3% faster than staging both).
// k2b holds K2_F tile by tile, [block][q][dz][kl] (k_layer_tiles), so that a
// block's K2_F columns are one contiguous run; M_F rows are KPB points long.
// PAR: the tensor has a z-parity per component (the dipolar one: xz and yz
// odd, the others even), only dz = 0 .. lz-1 is stored and a negative offset
// reads K(q)(|dz|), negated for the odd components: the operands, and so the
// field, are those of the full tensor, with half of its traffic.
template <bool SYM, bool PAR>
__global__ void k_layer_mix(const double2 *__restrict__ mf, const double2 *__restrict__ k2b, double2 *__restrict__ hf,
                            int64_t nc2, int lz, int nzr, int kpb, int odd)
{
    extern __shared__ double2 sh[];
    constexpr int NC = SYM ? 6 : 9;
    const int nzt = PAR ? lz : nzr;
    double2 *Ks = sh;                        // [NC][nzt][kpb]
    const int64_t k0 = (int64_t)blockIdx.x * kpb;
    const int nthr = kpb * lz;
    const double2 *kt = k2b + (int64_t)blockIdx.x * NC * nzt * kpb;
    for (int t = threadIdx.x; t < NC * nzt * kpb; t += nthr) Ks[t] = kt[t];
    __syncthreads();
    const int kl = threadIdx.x % kpb, z = threadIdx.x / kpb;
    const int64_t k = k0 + kl;
    if (z >= lz || k >= nc2) return;
    // complex multiply-accumulates as four FMAs each (half the operations of
    // a product then a sum; the field changes in its last bits only)
    double2 h0 = make_double2(0.0, 0.0), h1 = h0, h2 = h0;
    for (int zp = 0; zp < lz; ++zp) {
        int dz = z - zp;
        int flip = 0;   // PAR: components whose value is negated at this offset
        if (dz < 0) {
            if constexpr (PAR) {
                dz = -dz;
                flip = odd;
            } else {
                dz += nzr;
            }
        }
        const double2 m0 = mf[(int64_t)(0 * lz + zp) * nc2 + k], m1 = mf[(int64_t)(1 * lz + zp) * nc2 + k],
                      m2 = mf[(int64_t)(2 * lz + zp) * nc2 + k];
        auto kc = [&](int q) {
            double2 v = Ks[(q * nzt + dz) * kpb + kl];
            if constexpr (PAR) {
                if ((flip >> q) & 1) {
                    v.x = -v.x;
                    v.y = -v.y;
                }
            }
            return v;
        };
        if constexpr (SYM) {
            const double2 kxx = kc(0), kxy = kc(1), kxz = kc(2), kyy = kc(3), kyz = kc(4), kzz = kc(5);
            cmac(h0, kxx, m0), cmac(h0, kxy, m1), cmac(h0, kxz, m2);
            cmac(h1, kxy, m0), cmac(h1, kyy, m1), cmac(h1, kyz, m2);
            cmac(h2, kxz, m0), cmac(h2, kyz, m1), cmac(h2, kzz, m2);
        } else {
            cmac(h0, kc(0), m0), cmac(h0, kc(1), m1), cmac(h0, kc(2), m2);
            cmac(h1, kc(3), m0), cmac(h1, kc(4), m1), cmac(h1, kc(5), m2);
            cmac(h2, kc(6), m0), cmac(h2, kc(7), m1), cmac(h2, kc(8), m2);
        }
    }
    hf[(int64_t)(0 * lz + z) * nc2 + k] = h0;
    hf[(int64_t)(1 * lz + z) * nc2 + k] = h1;
    hf[(int64_t)(2 * lz + z) * nc2 + k] = h2;
}

// The dipolar form of k_layer_mix (symmetric tensor whose xz and yz
// components are odd in dz, the others even) at a layer count known when
// compiling; the operands and their order of summation are those of
// k_layer_mix (the same field bit for bit).  Thread (kl, z), kl fastest, 8
// k points per block of k.  Persistent: workgroup g mixes k blocks g, g + G,
// g + 2G, ... (G workgroups, all resident); the block's K2_F columns and M_F
// rows are both staged in LDS, and the next block's are fetched into
// registers while the current one is mixed; a negative offset costs two sign
// flips per source layer instead of a test per component.
template <int LZ>
__global__ void __launch_bounds__(8 * LZ) k_layer_mix_dip(const double2 *__restrict__ mf,
                                                          const double2 *__restrict__ k2b, double2 *__restrict__ hf,
                                                          int64_t nc2, int nblk)
{
    constexpr int KPB = 8, NC = 6, NT = KPB * LZ;
    extern __shared__ double2 sh[];
    double2 *Ks = sh;                    // [NC][LZ][KPB]
    double2 *Ms = sh + NC * LZ * KPB;    // [3 LZ][KPB]
    double2 rk0, rk1, rk2, rk3, rk4, rk5, rm0, rm1, rm2;   // the next block's share of this thread
#define MJ_MIX_FETCH(B)                                                                            \
    do {                                                                                           \
        const double2 *kt_ = k2b + (int64_t)(B) * NC * LZ * KPB + threadIdx.x;                     \
        rk0 = kt_[0 * NT], rk1 = kt_[1 * NT], rk2 = kt_[2 * NT];                                   \
        rk3 = kt_[3 * NT], rk4 = kt_[4 * NT], rk5 = kt_[5 * NT];                                   \
        const int64_t kk_ = (int64_t)(B) * KPB + threadIdx.x % KPB;                                \
        const int r_ = threadIdx.x / KPB;                                                          \
        const bool in_ = kk_ < nc2;                                                                \
        rm0 = in_ ? mf[(int64_t)(r_ + 0 * LZ) * nc2 + kk_] : make_double2(0.0, 0.0);               \
        rm1 = in_ ? mf[(int64_t)(r_ + 1 * LZ) * nc2 + kk_] : make_double2(0.0, 0.0);               \
        rm2 = in_ ? mf[(int64_t)(r_ + 2 * LZ) * nc2 + kk_] : make_double2(0.0, 0.0);               \
    } while (0)
    const int kl = threadIdx.x % KPB, z = threadIdx.x / KPB;
    int b = blockIdx.x;
    MJ_MIX_FETCH(b);
    for (; b < nblk; b += gridDim.x) {
        __syncthreads();   // the previous block's products are done with LDS
        Ks[threadIdx.x + 0 * NT] = rk0, Ks[threadIdx.x + 1 * NT] = rk1, Ks[threadIdx.x + 2 * NT] = rk2;
        Ks[threadIdx.x + 3 * NT] = rk3, Ks[threadIdx.x + 4 * NT] = rk4, Ks[threadIdx.x + 5 * NT] = rk5;
        Ms[threadIdx.x + 0 * NT] = rm0, Ms[threadIdx.x + 1 * NT] = rm1, Ms[threadIdx.x + 2 * NT] = rm2;
        __syncthreads();
        if (b + (int)gridDim.x < nblk) MJ_MIX_FETCH(b + gridDim.x);
        // k past nc2: zero tensor and M_F columns, no store
        double2 h0 = make_double2(0.0, 0.0), h1 = h0, h2 = h0;
#pragma unroll 1
        for (int zp = 0; zp < LZ; ++zp) {
            const bool neg = z < zp;
            const int dz = neg ? zp - z : z - zp;
            const double2 m0 = Ms[(0 * LZ + zp) * KPB + kl], m1 = Ms[(1 * LZ + zp) * KPB + kl],
                          m2 = Ms[(2 * LZ + zp) * KPB + kl];
            const double2 kxx = Ks[(0 * LZ + dz) * KPB + kl], kxy = Ks[(1 * LZ + dz) * KPB + kl];
            double2 kxz = Ks[(2 * LZ + dz) * KPB + kl];
            const double2 kyy = Ks[(3 * LZ + dz) * KPB + kl];
            double2 kyz = Ks[(4 * LZ + dz) * KPB + kl];
            const double2 kzz = Ks[(5 * LZ + dz) * KPB + kl];
            if (neg) {
                kxz.x = -kxz.x, kxz.y = -kxz.y;
                kyz.x = -kyz.x, kyz.y = -kyz.y;
            }
            cmac(h0, kxx, m0), cmac(h0, kxy, m1), cmac(h0, kxz, m2);
            cmac(h1, kxy, m0), cmac(h1, kyy, m1), cmac(h1, kyz, m2);
            cmac(h2, kxz, m0), cmac(h2, kyz, m1), cmac(h2, kzz, m2);
        }
        const int64_t k = (int64_t)b * KPB + kl;
        if (k < nc2) {
            hf[(int64_t)(0 * LZ + z) * nc2 + k] = h0;
            hf[(int64_t)(1 * LZ + z) * nc2 + k] = h1;
            hf[(int64_t)(2 * LZ + z) * nc2 + k] = h2;
        }
    }
#undef MJ_MIX_FETCH
}

// launches k_layer_mix_dip for the film's layer count (2 .. 32)
template <int LZ>
static bool launch_layer_mix_dip(int lz, unsigned nb, hipStream_t st, const double2 *mf, const double2 *k2b,
                                 double2 *hf, int64_t nc2)
{
    if constexpr (LZ > 32) {
        return false;
    } else {
        if (lz != LZ) return launch_layer_mix_dip<LZ + 1>(lz, nb, st, mf, k2b, hf, nc2);
        const size_t shm = sizeof(double2) * 9 * LZ * 8;
        // as many workgroups as are resident at once, at most one per k block
        static const int per_cu = [] {
            int n = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_layer_mix_dip<LZ>, 8 * LZ,
                                                             sizeof(double2) * 9 * LZ * 8) != hipSuccess)
                n = 1;
            return n > 0 ? n : 1;
        }();
        int cus = 0, dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
            cus = 256;
        const unsigned grid = std::min<unsigned>(nb, (unsigned)(per_cu * cus));
        hipLaunchKernelGGL(k_layer_mix_dip<LZ>, dim3(grid), dim3(8 * LZ), shm, st, mf, k2b, hf, nc2, (int)nb);
        return true;
    }
}

// K2_F, [q][dz][k] with k fastest (nzs offsets per component), into tiles of
// kpb points: [block][q][dz][kl] with the first nzt offsets of each
// component, zero past nc2
__global__ void k_layer_tiles(const double2 *__restrict__ k2, double2 *__restrict__ k2b, int64_t nc2, int ncomp,
                              int nzs, int nzt, int kpb, int64_t total)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int nq = ncomp * nzt;
    const int kl = (int)(t % kpb);
    const int64_t r = t / kpb;                 // (block * nq + row)
    const int64_t row = r % nq, blk = r / nq;  // row = q * nzt + dz
    const int64_t srow = (row / nzt) * nzs + row % nzt;
    const int64_t k = blk * kpb + kl;
    k2b[t] = k < nc2 ? k2[srow * nc2 + k] : make_double2(0.0, 0.0);
}

// z-parity of the transformed tensor: counts the (q, dz, k) with 0 < dz < lz
// where K2_F(q)(nzr - dz) differs from +-K2_F(q)(dz) (sign by bit q of odd)
__global__ void k_parity_check(const double2 *__restrict__ k2, int64_t nc2, int nzr, int lz, int odd, int64_t total,
                               unsigned long long *bad)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t k = t % nc2, r = t / nc2;
    const int dz = 1 + (int)(r % (lz - 1)), q = (int)(r / (lz - 1));
    const double2 a = k2[((int64_t)q * nzr + dz) * nc2 + k], b = k2[((int64_t)q * nzr + nzr - dz) * nc2 + k];
    const double s = ((odd >> q) & 1) ? -1.0 : 1.0;
    if (!(b.x == s * a.x && b.y == s * a.y)) atomicAdd(bad, 1ull);
}

__global__ void k_scale(double2 *x, int64_t n, double s)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i].x *= s;
    x[i].y *= s;
}

static int make_plan(rocfft_plan *p, rocfft_transform_type t, const int nrep[3], size_t batch, int dims = 3)
{
    size_t len[3] = {(size_t)nrep[0], (size_t)nrep[1], (size_t)nrep[2]};   // fastest first
    RF_TRY(rocfft_plan_create(p, rocfft_placement_notinplace, t, rocfft_precision_double, (size_t)dims, len, batch,
                              nullptr),
           "plan_create");
    return MJ_OK;
}

int dip_destroy(mj_ctx *c)
{
    DipolarFFT *d = c->dip;
    if (!d) return MJ_OK;
    if (d->fwd) rocfft_plan_destroy(d->fwd);
    if (d->inv) rocfft_plan_destroy(d->inv);
    if (d->info) rocfft_execution_info_destroy(d->info);
    if (d->grid) hipFree(d->grid);
    if (d->grid_in) hipFree(d->grid_in);
    if (d->spec) hipFree(d->spec);
    if (d->kf) hipFree(d->kf);
    if (d->spec_h) hipFree(d->spec_h);
    if (d->work) hipFree(d->work);
    delete d;
    c->dip = nullptr;
    return MJ_OK;
}

// the zeroed input grid of M where the state cannot be the transform's input
// itself (the lattice does not fill N_rep, or its SoA stride is not the site
// count, or, for a batch's copy, the stride is the batch's, set later)
static int alloc_grid_in(mj_ctx *c, DipolarFFT *d, bool allow_direct)
{
    const bool direct = allow_direct && d->layer && d->nrep[0] == c->dim[0] && d->nrep[1] == c->dim[1] &&
                        c->np == (int64_t)d->lz * d->np2;
    if (direct) return MJ_OK;
    const int64_t nk = d->layer ? (int64_t)d->lz * d->np2 : d->nk;
    MJ_HIP_CHECK(hipMalloc(&d->grid_in, sizeof(double) * 3 * nk));
    // on the context's stream, which does not wait for the null stream
    MJ_HIP_CHECK(hipMemsetAsync(d->grid_in, 0, sizeof(double) * 3 * nk, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

// film form: K2_F per layer offset, 2D plans, LDS-staged layer kernel
static int dip_create_layers(mj_ctx *c, DipolarFFT *d, const double *K_n, int ncomp)
{
    d->layer = 1;
    d->lz = c->dim[2];
    d->nzr = d->nrep[2];
    d->np2 = (int64_t)d->nrep[0] * d->nrep[1];
    d->nc2 = (int64_t)(d->nrep[0] / 2 + 1) * d->nrep[1];
    const int64_t np2 = d->np2, nc2 = d->nc2;
    const int lz = d->lz, nzr = d->nzr;
    MJ_HIP_CHECK(hipMalloc(&d->grid, sizeof(double) * 3 * lz * np2));
    if (int rc_ = alloc_grid_in(c, d, true)) return rc_;
    MJ_HIP_CHECK(hipMalloc(&d->spec, sizeof(double2) * 3 * lz * nc2));
    MJ_HIP_CHECK(hipMalloc(&d->spec_h, sizeof(double2) * 3 * lz * nc2));
    MJ_HIP_CHECK(hipMalloc(&d->kf, sizeof(double2) * ncomp * nzr * nc2));
    // real-space K, component-major then layer offset: kr[(q * nzr + dz) * np2 + x + nx * y]
    std::vector<double> kr((size_t)ncomp * nzr * np2);
    static const int sa[6] = {0, 0, 0, 1, 1, 2}, sb[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < ncomp; ++q) {
        const int a = d->sym ? sa[q] : q % 3, b = d->sym ? sb[q] : q / 3;
        for (int dz = 0; dz < nzr; ++dz)
            for (int64_t p = 0; p < np2; ++p)
                kr[((size_t)q * nzr + dz) * np2 + p] = K_n[a + 3 * b + 9 * (p + np2 * dz)];
    }
    double *kdev = nullptr;
    MJ_HIP_CHECK(hipMalloc(&kdev, sizeof(double) * kr.size()));
    MJ_HIP_CHECK(hipMemcpy(kdev, kr.data(), sizeof(double) * kr.size(), hipMemcpyHostToDevice));
    rocfft_plan pk = nullptr;
    int rc = make_plan(&pk, rocfft_transform_type_real_forward, d->nrep, (size_t)ncomp * nzr, 2);
    if (rc == MJ_OK) rc = make_plan(&d->fwd, rocfft_transform_type_real_forward, d->nrep, (size_t)3 * lz, 2);
    if (rc == MJ_OK) rc = make_plan(&d->inv, rocfft_transform_type_real_inverse, d->nrep, (size_t)3 * lz, 2);
    if (rc) {
        if (pk) rocfft_plan_destroy(pk);
        hipFree(kdev);
        return rc;
    }
    size_t ws = 0, w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(pk, &ws);
    rocfft_plan_get_work_buffer_size(d->fwd, &w1);
    rocfft_plan_get_work_buffer_size(d->inv, &w2);
    d->work_size = std::max(ws, std::max(w1, w2));
    if (d->work_size) MJ_HIP_CHECK(hipMalloc(&d->work, d->work_size));
    RF_TRY(rocfft_execution_info_create(&d->info), "execution_info_create");
    if (d->work_size) RF_TRY(rocfft_execution_info_set_work_buffer(d->info, d->work, d->work_size), "set_work_buffer");
    RF_TRY(rocfft_execution_info_set_stream(d->info, c->stream), "set_stream");
    void *in[1] = {kdev};
    void *out[1] = {d->kf};
    RF_TRY(rocfft_execute(pk, in, out, d->info), "execute K");
    const int64_t tot = (int64_t)ncomp * nzr * nc2;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, d->kf, tot,
                       1.0 / (double)np2);
    rocfft_plan_destroy(pk);
    // z-parity: K(q)(-dz) = +-K(q)(dz) exactly for every component (the
    // dipolar tensor: xz, yz odd), checked on K_n and again on K2_F (the
    // in-plane transform of a negated layer is the negated transform, if the
    // library runs every layer alike); offsets -dz sit at nzr - dz, apart from
    // 0 .. lz-1 when nzr >= 2 lz - 1
    d->nzt = nzr;
    d->odd = 0;
    const char *penv = getenv("MJ_DIP_PARITY");   // "0": keep every offset (tests)
    if (lz > 1 && nzr >= 2 * lz - 1 && !(penv && penv[0] == '0')) {
        int odd = 0;
        bool par = true;
        for (int q = 0; q < ncomp && par; ++q) {
            bool even = true, oddq = true;
            for (int dz = 1; dz < lz && (even || oddq); ++dz)
                for (int64_t p = 0; p < np2; ++p) {
                    const double a = kr[((size_t)q * nzr + dz) * np2 + p], b = kr[((size_t)q * nzr + nzr - dz) * np2 + p];
                    even = even && b == a;
                    oddq = oddq && b == -a;
                }
            if (even) continue;
            if (oddq) odd |= 1 << q;
            else par = false;
        }
        if (par) {
            unsigned long long *bad = nullptr, nbad = 1;
            MJ_HIP_CHECK(hipMalloc(&bad, sizeof(unsigned long long)));
            MJ_HIP_CHECK(hipMemsetAsync(bad, 0, sizeof(unsigned long long), c->stream));
            const int64_t tc = (int64_t)ncomp * (lz - 1) * nc2;
            hipLaunchKernelGGL(k_parity_check, dim3((unsigned)((tc + 255) / 256)), dim3(256), 0, c->stream, d->kf, nc2,
                               nzr, lz, odd, tc, bad);
            MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
            MJ_HIP_CHECK(hipMemcpy(&nbad, bad, sizeof nbad, hipMemcpyDeviceToHost));
            hipFree(bad);
            if (nbad == 0) {
                d->nzt = lz;
                d->odd = odd;
            }
        }
    }
    const int nzt = d->nzt;
    // k points per workgroup: 8 (128-byte rows of M_F and H_F; on the 16-layer
    // film 128 threads and 12 KiB of LDS), fewer where the block would pass
    // 1024 threads or 64 KiB of LDS
    const int per_k = ncomp * nzt * (int)sizeof(double2);
    d->kpb = std::max(1, std::min(std::min(8, 1024 / lz), 65536 / per_k));
    // K2_F regrouped tile by tile for the layer kernel's contiguous loads
    const int64_t nblk = (nc2 + d->kpb - 1) / d->kpb, tiles = nblk * ncomp * nzt * d->kpb;
    double2 *kb = nullptr;
    if (hipMalloc(&kb, sizeof(double2) * tiles) != hipSuccess) {
        hipFree(kdev);
        mj_set_error("mj_set_fft_operator: device allocation failed");
        return MJ_ENOMEM;
    }
    hipLaunchKernelGGL(k_layer_tiles, dim3((unsigned)((tiles + 255) / 256)), dim3(256), 0, c->stream, d->kf, kb, nc2,
                       ncomp, nzr, nzt, d->kpb, tiles);
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    hipFree(d->kf);
    d->kf = kb;
    hipFree(kdev);
    return MJ_OK;
}

int dip_create(mj_ctx *c, const int N_rep[3], const int periodic[3], const double *K_n)
{
    static bool setup_done = false;
    if (!setup_done) {
        RF_TRY(rocfft_setup(), "setup");
        setup_done = true;
    }
    for (int d = 0; d < 3; ++d) {
        int want = periodic[d] ? c->dim[d] : 2 * c->dim[d];
        if (N_rep[d] != want) {
            mj_set_error("mj_set_fft_operator: N_rep must equal dim_lat (periodic) or 2*dim_lat (open)");
            return MJ_EINVAL;
        }
    }
    if (c->dip) dip_destroy(c);
    DipolarFFT *d = new DipolarFFT();
    c->dip = d;
    for (int q = 0; q < 3; ++q) d->nrep[q] = N_rep[q];
    d->nk = (int64_t)N_rep[0] * N_rep[1] * N_rep[2];
    d->nc = (int64_t)(N_rep[0] / 2 + 1) * N_rep[1] * N_rep[2];
    const int64_t nk = d->nk, nc = d->nc;

    // symmetric operator? (K(a,b) == K(b,a) for every k, up to rounding)
    double kmax = 0.0, asym = 0.0;
    for (int64_t k = 0; k < nk; ++k)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                double v = K_n[a + 3 * b + 9 * k];
                kmax = std::max(kmax, fabs(v));
                if (a < b) asym = std::max(asym, fabs(v - K_n[b + 3 * a + 9 * k]));
            }
    d->sym = asym <= 1.0e-13 * kmax ? 1 : 0;
    const int ncomp = d->sym ? 6 : 9;
    const char *lenv = getenv("MJ_DIP_LAYERS");
    if (c->dim[2] <= 32 && !(lenv && lenv[0] == '0')) return dip_create_layers(c, d, K_n, ncomp);

    MJ_HIP_CHECK(hipMalloc(&d->grid, sizeof(double) * 3 * nk));
    if (int rc_ = alloc_grid_in(c, d, true)) return rc_;
    MJ_HIP_CHECK(hipMalloc(&d->spec, sizeof(double2) * 3 * nc));
    MJ_HIP_CHECK(hipMalloc(&d->kf, sizeof(double2) * ncomp * nc));

    // K_F on the device: batched R2C of the real-space components, / prod(N_rep)
    std::vector<double> kr((size_t)ncomp * nk);
    static const int sa[6] = {0, 0, 0, 1, 1, 2}, sb[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < ncomp; ++q) {
        int a = d->sym ? sa[q] : q % 3, b = d->sym ? sb[q] : q / 3;
        for (int64_t k = 0; k < nk; ++k) kr[(size_t)q * nk + k] = K_n[a + 3 * b + 9 * k];
    }
    double *kdev = nullptr;
    MJ_HIP_CHECK(hipMalloc(&kdev, sizeof(double) * kr.size()));
    MJ_HIP_CHECK(hipMemcpy(kdev, kr.data(), sizeof(double) * kr.size(), hipMemcpyHostToDevice));
    rocfft_plan pk = nullptr;
    int rc = make_plan(&pk, rocfft_transform_type_real_forward, N_rep, (size_t)ncomp);
    if (rc) {
        hipFree(kdev);
        return rc;
    }
    size_t ws = 0;
    rocfft_plan_get_work_buffer_size(pk, &ws);
    rc = make_plan(&d->fwd, rocfft_transform_type_real_forward, N_rep, 3);
    if (rc == MJ_OK) rc = make_plan(&d->inv, rocfft_transform_type_real_inverse, N_rep, 3);
    if (rc) {
        rocfft_plan_destroy(pk);
        hipFree(kdev);
        return rc;
    }
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(d->fwd, &w1);
    rocfft_plan_get_work_buffer_size(d->inv, &w2);
    d->work_size = std::max(ws, std::max(w1, w2));
    if (d->work_size) MJ_HIP_CHECK(hipMalloc(&d->work, d->work_size));
    RF_TRY(rocfft_execution_info_create(&d->info), "execution_info_create");
    if (d->work_size) RF_TRY(rocfft_execution_info_set_work_buffer(d->info, d->work, d->work_size), "set_work_buffer");
    RF_TRY(rocfft_execution_info_set_stream(d->info, c->stream), "set_stream");
    void *in[1] = {kdev};
    void *out[1] = {d->kf};
    RF_TRY(rocfft_execute(pk, in, out, d->info), "execute K");
    int64_t tot = (int64_t)ncomp * nc;
    hipLaunchKernelGGL(k_scale, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, c->stream, d->kf, tot,
                       1.0 / (double)nk);
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    rocfft_plan_destroy(pk);
    hipFree(kdev);
    return MJ_OK;
}

// dst->dip: a second transform with src's operator (same lattice and grid):
// plans, work and field buffers of its own on dst->stream, the transformed
// tensor copied on the device
int dip_clone(mj_ctx *dst, const mj_ctx *src)
{
    const DipolarFFT *s = src->dip;
    if (!s) {
        mj_set_error("dip_clone: no dipolar term");
        return MJ_EINVAL;
    }
    if (dst->dip) dip_destroy(dst);
    DipolarFFT *d = new DipolarFFT();
    dst->dip = d;
    for (int q = 0; q < 3; ++q) d->nrep[q] = s->nrep[q];
    d->nk = s->nk;
    d->nc = s->nc;
    d->sym = s->sym;
    d->layer = s->layer;
    d->lz = s->lz;
    d->nzr = s->nzr;
    d->np2 = s->np2;
    d->nc2 = s->nc2;
    d->kpb = s->kpb;
    d->nzt = s->nzt;
    d->odd = s->odd;
    const int ncomp = d->sym ? 6 : 9;
    size_t kbytes;
    if (d->layer) {
        const int64_t nblk = (d->nc2 + d->kpb - 1) / d->kpb;
        kbytes = sizeof(double2) * (size_t)(nblk * ncomp * d->nzt * d->kpb);
        MJ_HIP_CHECK(hipMalloc(&d->grid, sizeof(double) * 3 * d->lz * d->np2));
        if (int rc_ = alloc_grid_in(dst, d, false)) return rc_;
        MJ_HIP_CHECK(hipMalloc(&d->spec, sizeof(double2) * 3 * d->lz * d->nc2));
        MJ_HIP_CHECK(hipMalloc(&d->spec_h, sizeof(double2) * 3 * d->lz * d->nc2));
        int rc = make_plan(&d->fwd, rocfft_transform_type_real_forward, d->nrep, (size_t)3 * d->lz, 2);
        if (rc == MJ_OK) rc = make_plan(&d->inv, rocfft_transform_type_real_inverse, d->nrep, (size_t)3 * d->lz, 2);
        if (rc != MJ_OK) return rc;
    } else {
        kbytes = sizeof(double2) * (size_t)ncomp * d->nc;
        MJ_HIP_CHECK(hipMalloc(&d->grid, sizeof(double) * 3 * d->nk));
        if (int rc_ = alloc_grid_in(dst, d, false)) return rc_;
        MJ_HIP_CHECK(hipMalloc(&d->spec, sizeof(double2) * 3 * d->nc));
        int rc = make_plan(&d->fwd, rocfft_transform_type_real_forward, d->nrep, 3);
        if (rc == MJ_OK) rc = make_plan(&d->inv, rocfft_transform_type_real_inverse, d->nrep, 3);
        if (rc != MJ_OK) return rc;
    }
    MJ_HIP_CHECK(hipMalloc(&d->kf, kbytes));
    MJ_HIP_CHECK(hipMemcpy(d->kf, s->kf, kbytes, hipMemcpyDeviceToDevice));
    size_t w1 = 0, w2 = 0;
    rocfft_plan_get_work_buffer_size(d->fwd, &w1);
    rocfft_plan_get_work_buffer_size(d->inv, &w2);
    d->work_size = std::max(w1, w2);
    if (d->work_size) MJ_HIP_CHECK(hipMalloc(&d->work, d->work_size));
    RF_TRY(rocfft_execution_info_create(&d->info), "execution_info_create");
    if (d->work_size) RF_TRY(rocfft_execution_info_set_work_buffer(d->info, d->work, d->work_size), "set_work_buffer");
    RF_TRY(rocfft_execution_info_set_stream(d->info, dst->stream), "set_stream");
    return MJ_OK;
}

void dip_geometry(const mj_ctx *c, const double **grid, int *dnx, int *dny, int64_t *hstride)
{
    const DipolarFFT *d = c->dip;
    *grid = d->grid;
    *dnx = d->nrep[0];
    *dny = d->nrep[1];
    *hstride = d->layer ? (int64_t)d->lz * d->np2 : d->nk;
}

static int dip_apply_layers(mj_ctx *c, DipolarFFT *d, const double *m, KArgs *ka)
{
    const int64_t n = (int64_t)d->lz * d->np2;
    // periodic in plane with the SoA stride equal to the site count: the state
    // is already the transform's input (component a, layer z at (a lz + z)
    // N_rep_x N_rep_y), and the out-of-place real forward transform leaves it
    // as it is; otherwise its sites are placed into the zero-padded grid
    const bool direct = !d->grid_in;
    if (!direct) {
        const int64_t ns = (int64_t)c->dim[0] * c->dim[1] * c->dim[2];
        hipLaunchKernelGGL(k_place, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream, m, c->np, ns,
                           c->dim[0], c->dim[1], d->nrep[0], d->nrep[1], n, d->grid_in);
    }
    void *in[1] = {direct ? (void *)m : (void *)d->grid_in};
    void *out[1] = {d->spec};
    RF_TRY(rocfft_execute(d->fwd, in, out, d->info), "execute forward (2D)");
    const unsigned nb = (unsigned)((d->nc2 + d->kpb - 1) / d->kpb);
    const int ncomp = d->sym ? 6 : 9;
    const size_t shm = sizeof(double2) * (size_t)(ncomp * d->nzt) * d->kpb;
    const bool par = d->nzt != d->nzr;
    const dim3 g(nb), t(d->kpb * d->lz);
    // the dipolar tensor's parity (xz and yz odd): the kernel of that form
    const bool dipf = d->sym && par && d->odd == ((1 << 2) | (1 << 4)) && d->kpb == 8;
    const bool done = dipf && launch_layer_mix_dip<2>(d->lz, nb, c->stream, d->spec, d->kf, d->spec_h, d->nc2);
    if (done)
        ;   // launched above
    else if (d->sym && par)
        hipLaunchKernelGGL((k_layer_mix<true, true>), g, t, shm, c->stream, d->spec, d->kf, d->spec_h, d->nc2, d->lz,
                           d->nzr, d->kpb, d->odd);
    else if (d->sym)
        hipLaunchKernelGGL((k_layer_mix<true, false>), g, t, shm, c->stream, d->spec, d->kf, d->spec_h, d->nc2, d->lz,
                           d->nzr, d->kpb, 0);
    else if (par)
        hipLaunchKernelGGL((k_layer_mix<false, true>), g, t, shm, c->stream, d->spec, d->kf, d->spec_h, d->nc2, d->lz,
                           d->nzr, d->kpb, d->odd);
    else
        hipLaunchKernelGGL((k_layer_mix<false, false>), g, t, shm, c->stream, d->spec, d->kf, d->spec_h, d->nc2, d->lz,
                           d->nzr, d->kpb, 0);
    void *in2[1] = {d->spec_h};
    void *out2[1] = {d->grid};
    RF_TRY(rocfft_execute(d->inv, in2, out2, d->info), "execute inverse (2D)");
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("dipolar kernels: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    ka->hdip = d->grid;
    ka->dnx = d->nrep[0];
    ka->dny = d->nrep[1];
    ka->hstride = n;
    return MJ_OK;
}

int dip_apply(mj_ctx *c, const double *m, KArgs *ka)
{
    DipolarFFT *d = c->dip;
    if (d->layer) return dip_apply_layers(c, d, m, ka);
    const int64_t nk = d->nk, nc = d->nc;
    const int64_t ns = (int64_t)c->dim[0] * c->dim[1] * c->dim[2];
    hipLaunchKernelGGL(k_place, dim3((unsigned)((ns + 255) / 256)), dim3(256), 0, c->stream, m, c->np, ns, c->dim[0],
                       c->dim[1], d->nrep[0], d->nrep[1], nk, d->grid_in);
    void *in[1] = {d->grid_in};
    void *out[1] = {d->spec};
    RF_TRY(rocfft_execute(d->fwd, in, out, d->info), "execute forward");
    if (d->sym)
        hipLaunchKernelGGL(k_kmul<true>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->stream, d->spec, d->kf, nc);
    else
        hipLaunchKernelGGL(k_kmul<false>, dim3((unsigned)((nc + 255) / 256)), dim3(256), 0, c->stream, d->spec, d->kf, nc);
    void *in2[1] = {d->spec};
    void *out2[1] = {d->grid};
    RF_TRY(rocfft_execute(d->inv, in2, out2, d->info), "execute inverse");
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        mj_set_error(std::string("dipolar kernels: ") + hipGetErrorString(e));
        return MJ_EHIP;
    }
    ka->hdip = d->grid;
    ka->dnx = d->nrep[0];
    ka->dny = d->nrep[1];
    ka->hstride = nk;
    return MJ_OK;
}

extern "C" int mj_set_fft_operator(mj_ctx *c, int dim_mode, const int N_rep[3], const int periodic[3],
                                   const double *K_n)
{
    if (!c || !N_rep || !periodic || !K_n) {
        mj_set_error("mj_set_fft_operator: null argument");
        return MJ_EINVAL;
    }
    if (dim_mode != 3) {
        mj_set_error("mj_set_fft_operator: only dim_mode == 3");
        return MJ_ENOTSUP;
    }
    MJ_HIP_CHECK(hipSetDevice(c->device));
    int rc = dip_create(c, N_rep, periodic, K_n);
    if (rc) {
        dip_destroy(c);
        return rc;
    }
    c->dirty = 1;
    return MJ_OK;
}
