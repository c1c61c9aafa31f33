// C-ABI of the LLG hot path: context, Hamiltonian intake, state transfer and
// the step loop (host side of src/dynamics/spindynamics.f90:260-323).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <array>
#include <map>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "mj_internal.h"

int mj_launch_stage(int mode, bool noise, const KArgs &a, hipStream_t s);
int mj_launch_reduce(const double *part, int n, double *out, double scale, hipStream_t s);
int mj_launch_reduce_max(const double *part, int n, double *out, hipStream_t s);
int mj_launch_fold(const double *part, int n, double *out, double scale, int nthr, hipStream_t s);
int mj_launch_step_counter(uint64_t *p, uint64_t v, int add, hipStream_t s);

// steps per captured graph (even: the Euler path swaps its two state buffers every step)
static const int kGraphChunk = 32;

static void drop_graph(mj_ctx *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    c->gkey.clear();
}
int mj_launch_spmv(const mj_ctx *c, const double *m, double *out, hipStream_t s);
int mj_launch_aos2soa(const double *in, double *out, int64_t n, int64_t np, hipStream_t s);
int mj_launch_soa2aos(const double *in, double *out, int64_t n, int64_t np, double eps, hipStream_t s);
int mj_launch_normals(uint64_t seed, uint64_t step, int64_t n, double *out, hipStream_t s);
static int launch(mj_ctx *c, int mode, bool noise, const KArgs &a);

static thread_local std::string g_err;
void mj_set_error(const std::string &msg) { g_err = msg; }

#define MJ_FAIL(code, msg)        \
    do {                          \
        mj_set_error(msg);        \
        return (code);            \
    } while (0)
#define MJ_TRY(x)               \
    do {                        \
        int r_ = (x);           \
        if (r_ != MJ_OK) return r_; \
    } while (0)

static const double kHbar = 6.58211928e-1;

extern "C" int mj_last_error(char *buf, int len)
{
    if (!buf || len <= 0) return MJ_EINVAL;
    snprintf(buf, (size_t)len, "%s", g_err.c_str());
    return MJ_OK;
}

static int set_dev(mj_ctx *c)
{
    MJ_HIP_CHECK(hipSetDevice(c->device));
    return MJ_OK;
}

extern "C" int mj_create(mj_ctx **out, int device, const int dim_lat[3], const int periodic[3], int dim_mode)
{
    if (!out || !dim_lat || !periodic) MJ_FAIL(MJ_EINVAL, "mj_create: null argument");
    *out = nullptr;
    for (int d = 0; d < 3; ++d)
        if (dim_lat[d] < 1) MJ_FAIL(MJ_EINVAL, "mj_create: dim_lat must be >= 1");
    if (dim_mode != 3) MJ_FAIL(MJ_ENOTSUP, "mj_create: only dim_mode == 3 (one magnetic atom per cell)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) MJ_FAIL(MJ_EHIP, "mj_create: no HIP device");
    if (device < 0 || device >= ndev) MJ_FAIL(MJ_EINVAL, "mj_create: bad device index");
    mj_ctx *c = new mj_ctx();
    c->device = device;
    for (int d = 0; d < 3; ++d) {
        c->dim[d] = dim_lat[d];
        c->periodic[d] = periodic[d] ? 1 : 0;
    }
    c->dim_mode = dim_mode;
    c->n = (int64_t)dim_lat[0] * dim_lat[1] * dim_lat[2];
    c->np = (c->n + 63) / 64 * 64;
    int rc = set_dev(c);
    if (rc) {
        delete c;
        return rc;
    }
    size_t bytes = sizeof(double) * 3 * (size_t)c->np;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&c->A, bytes) != hipSuccess || hipMalloc(&c->Bb, bytes) != hipSuccess ||
        hipMalloc(&c->Db, bytes) != hipSuccess || hipMalloc(&c->Fb, bytes) != hipSuccess ||
        hipMalloc(&c->Nb, bytes) != hipSuccess) {
        mj_destroy(&c);
        MJ_FAIL(MJ_ENOMEM, "mj_create: device allocation failed");
    }
    hipMemsetAsync(c->A, 0, bytes, c->stream);
    hipMemsetAsync(c->Bb, 0, bytes, c->stream);
    hipMemsetAsync(c->Db, 0, bytes, c->stream);
    hipStreamSynchronize(c->stream);
    *out = c;
    return MJ_OK;
}

extern "C" int mj_destroy(mj_ctx **pc)
{
    if (!pc || !*pc) return MJ_OK;
    mj_ctx *c = *pc;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->dip) dip_destroy(c);
    if (c->jit) mj_jit_destroy(c);
    if (c->adj) mj_adj_destroy(c);
    if (c->mc) mj_mc_destroy(c);
    drop_graph(c);
    if (c->d_step) hipFree(c->d_step);
    if (c->estage) hipFree(c->estage);
    double *bufs[] = {c->A, c->Bb, c->Db, c->Fb, c->Nb, c->epart, c->evals, c->d_cf, c->d_bmap, c->d_v, c->d_maps};
    for (double *p : bufs)
        if (p) hipFree(p);
    for (double *p : c->Drk)
        if (p) hipFree(p);
    for (int *p : c->term_ofs)
        if (p) hipFree(p);
    for (double *p : c->term_blk)
        if (p) hipFree(p);
    if (c->tm_in) hipFree(c->tm_in);
    if (c->tm_out) hipFree(c->tm_out);
    if (c->Lf) hipFree(c->Lf);
    if (c->d_ofs) hipFree(c->d_ofs);
    if (c->d_rp) hipFree(c->d_rp);
    if (c->d_ci) hipFree(c->d_ci);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    *pc = nullptr;
    return MJ_OK;
}

// ---------------------------------------------------------------------------
// Hamiltonian intake
// ---------------------------------------------------------------------------
static inline void site_xyz(const mj_ctx *c, int64_t s, int &x, int &y, int &z)
{
    x = (int)(s % c->dim[0]);
    int64_t r = s / c->dim[0];
    y = (int)(r % c->dim[1]);
    z = (int)(r / c->dim[1]);
}

static inline int canon(int d, int n, int periodic)
{
    if (!periodic) return d;
    d %= n;
    if (d < 0) d += n;
    if (2 * d > n) d -= n;
    return d;
}

static void add_block(mj_ctx *c, int dx, int dy, int dz, const double *blk)
{
    for (size_t k = 0; k < c->st_blk.size() / 9; ++k) {
        if (c->st_ofs[3 * k] == dx && c->st_ofs[3 * k + 1] == dy && c->st_ofs[3 * k + 2] == dz) {
            for (int q = 0; q < 9; ++q) c->st_blk[9 * k + q] += blk[q];
            return;
        }
    }
    c->st_ofs.push_back(dx);
    c->st_ofs.push_back(dy);
    c->st_ofs.push_back(dz);
    for (int q = 0; q < 9; ++q) c->st_blk.push_back(blk[q]);
}

// site hash for the coverage check of try_compress (splitmix64 finaliser)
static inline uint64_t site_hash(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// Compress an M-M COO to translation-invariant 3x3 blocks if it is one: every
// (offset, component pair) key must carry one value, at every site that has
// that neighbour, the same number of times at each (duplicated entries add up,
// as the reference's sparse backends sum them).  Coverage is checked with a
// sum of site hashes per key against the sum over the sites that have the
// neighbour.
static int try_compress(mj_ctx *c, int64_t nnz, const int *row1, const int *col1, const double *val,
                        std::vector<std::tuple<int, int, int, double *>> *unused)
{
    (void)unused;
    struct Acc {
        double vmin, vmax;
        int64_t count;
        uint64_t hsum;
    };
    std::unordered_map<uint64_t, Acc> acc;
    acc.reserve(1024);
    int64_t n3 = 3 * c->n;
    for (int64_t k = 0; k < nnz; ++k) {
        int64_t r = (int64_t)row1[k] - 1, q = (int64_t)col1[k] - 1;
        if (r < 0 || r >= n3 || q < 0 || q >= n3) return -2;
        int ix, iy, iz, jx, jy, jz;
        site_xyz(c, r / 3, ix, iy, iz);
        site_xyz(c, q / 3, jx, jy, jz);
        int dx = canon(jx - ix, c->dim[0], c->periodic[0]);
        int dy = canon(jy - iy, c->dim[1], c->periodic[1]);
        int dz = canon(jz - iz, c->dim[2], c->periodic[2]);
        uint64_t key = ((uint64_t)(uint16_t)(int16_t)dx) | ((uint64_t)(uint16_t)(int16_t)dy << 16) |
                       ((uint64_t)(uint16_t)(int16_t)dz << 32) | ((uint64_t)(r % 3) << 48) |
                       ((uint64_t)(q % 3) << 52);
        const uint64_t hs = site_hash((uint64_t)(r / 3));
        auto it = acc.find(key);
        if (it == acc.end()) acc.emplace(key, Acc{val[k], val[k], 1, hs});
        else {
            it->second.vmin = std::min(it->second.vmin, val[k]);
            it->second.vmax = std::max(it->second.vmax, val[k]);
            it->second.count += 1;
            it->second.hsum += hs;
        }
        if (acc.size() > 4096) return -1;   // not a short-ranged stencil
    }
    std::vector<std::pair<std::array<int, 3>, std::array<double, 9>>> blocks;
    for (auto &kv : acc) {
        int dx = (int16_t)(kv.first & 0xffff), dy = (int16_t)((kv.first >> 16) & 0xffff);
        int dz = (int16_t)((kv.first >> 32) & 0xffff);
        int a = (int)((kv.first >> 48) & 0xf), b = (int)((kv.first >> 52) & 0xf);
        if (kv.second.vmin != kv.second.vmax) return -1;
        int64_t expect = 1;
        int d3[3] = {dx, dy, dz};
        for (int d = 0; d < 3; ++d)
            expect *= c->periodic[d] ? c->dim[d] : std::max(0, c->dim[d] - std::abs(d3[d]));
        if (expect <= 0 || kv.second.count % expect != 0) return -1;
        const int64_t mult = kv.second.count / expect;
        // every site with the neighbour d carries the entry mult times
        uint64_t want = 0;
        for (int64_t i = 0; i < c->n; ++i) {
            int x, y, z;
            site_xyz(c, i, x, y, z);
            const int g[3] = {x + dx, y + dy, z + dz};
            bool ok = true;
            for (int q = 0; q < 3; ++q)
                if (!c->periodic[q] && (g[q] < 0 || g[q] >= c->dim[q])) ok = false;
            if (ok) want += site_hash((uint64_t)i);
        }
        if (kv.second.hsum != want * (uint64_t)mult) return -1;
        double v = kv.second.vmin * (double)mult;
        std::array<int, 3> o = {dx, dy, dz};
        auto it = std::find_if(blocks.begin(), blocks.end(), [&](auto &p) { return p.first == o; });
        if (it == blocks.end()) {
            blocks.push_back({o, {0, 0, 0, 0, 0, 0, 0, 0, 0}});
            it = blocks.end() - 1;
        }
        it->second[a * 3 + b] = v;
    }
    mj_ctx::MMTerm term;
    for (auto &b : blocks) {
        add_block(c, b.first[0], b.first[1], b.first[2], b.second.data());
        term.ofs3.insert(term.ofs3.end(), b.first.begin(), b.first.end());
        term.blk9.insert(term.blk9.end(), b.second.begin(), b.second.end());
    }
    c->mm_terms.push_back(std::move(term));
    return 0;
}

extern "C" int mj_add_hamiltonian_coo(mj_ctx *c, int op_l, int op_r, int64_t nnz, const int *row1,
                                      const int *col1, const double *val)
{
    if (!c || nnz < 0 || (nnz > 0 && (!row1 || !col1 || !val))) MJ_FAIL(MJ_EINVAL, "mj_add_hamiltonian_coo: bad argument");
    int64_t n3 = 3 * c->n;
    if (op_l == MJ_OP_M && op_r == MJ_OP_M) {
        int rc = c->mm_generic ? -1 : try_compress(c, nnz, row1, col1, val, nullptr);
        if (rc == -2) MJ_FAIL(MJ_EINVAL, "mj_add_hamiltonian_coo: index out of range");
        if (rc != 0) {
            c->mm_terms_generic = 1;
            // generic path keeps the whole operator (previous stencil terms expanded into it)
            if (!c->mm_generic) {
                c->mm_generic = 1;
                for (size_t k = 0; k < c->st_blk.size() / 9; ++k) {
                    int o[3] = {c->st_ofs[3 * k], c->st_ofs[3 * k + 1], c->st_ofs[3 * k + 2]};
                    for (int64_t s = 0; s < c->n; ++s) {
                        int x, y, z;
                        site_xyz(c, s, x, y, z);
                        int g[3] = {x + o[0], y + o[1], z + o[2]};
                        bool ok = true;
                        for (int d = 0; d < 3; ++d) {
                            if (g[d] < 0 || g[d] >= c->dim[d]) {
                                if (!c->periodic[d]) ok = false;
                                g[d] = ((g[d] % c->dim[d]) + c->dim[d]) % c->dim[d];
                            }
                        }
                        if (!ok) continue;
                        int64_t t = g[0] + (int64_t)c->dim[0] * (g[1] + (int64_t)c->dim[1] * g[2]);
                        for (int a = 0; a < 3; ++a)
                            for (int b = 0; b < 3; ++b)
                                if (c->st_blk[9 * k + 3 * a + b] != 0.0) {
                                    c->g_rows.push_back(3 * s + a);
                                    c->g_cols.push_back((int)(3 * t + b));
                                    c->g_vals.push_back(c->st_blk[9 * k + 3 * a + b]);
                                }
                    }
                }
                c->st_ofs.clear();
                c->st_blk.clear();
            }
            for (int64_t k = 0; k < nnz; ++k) {
                c->g_rows.push_back((int64_t)row1[k] - 1);
                c->g_cols.push_back(col1[k] - 1);
                c->g_vals.push_back(val[k]);
            }
        }
        c->has_mm = 1;
    } else if ((op_l == MJ_OP_B && op_r == MJ_OP_M) || (op_l == MJ_OP_M && op_r == MJ_OP_B)) {
        // Zeeman "BM": the same 3x3 on-site block at every site
        double blk[9];
        int have[9] = {0};
        for (int q = 0; q < 9; ++q) blk[q] = 0.0;
        std::vector<int> seen(9, 0);
        for (int64_t k = 0; k < nnz; ++k) {
            int64_t r = (int64_t)row1[k] - 1, q = (int64_t)col1[k] - 1;
            if (r < 0 || r >= n3 || q < 0 || q >= n3) MJ_FAIL(MJ_EINVAL, "mj_add_hamiltonian_coo: index out of range");
            if (r / 3 != q / 3) MJ_FAIL(MJ_ENOTSUP, "mj_add_hamiltonian_coo: B-M term must be on-site");
            int a = (int)(r % 3), b = (int)(q % 3);
            if (op_l == MJ_OP_M) std::swap(a, b);   // store as Z(b-index, m-index)
            if (have[3 * a + b] && blk[3 * a + b] != val[k])
                MJ_FAIL(MJ_ENOTSUP, "mj_add_hamiltonian_coo: B-M term is not translation invariant");
            have[3 * a + b] = 1;
            blk[3 * a + b] = val[k];
            seen[3 * a + b]++;
        }
        for (int q = 0; q < 9; ++q) {
            if (seen[q] != 0 && seen[q] != c->n)
                MJ_FAIL(MJ_ENOTSUP, "mj_add_hamiltonian_coo: B-M term does not cover every site");
            c->zblk[q] += blk[q];
        }
        c->has_z = 1;
    } else {
        MJ_FAIL(MJ_ENOTSUP, "mj_add_hamiltonian_coo: only M-M and B-M rank-2 terms are supported");
    }
    c->dirty = 1;
    return MJ_OK;
}

extern "C" int mj_add_hamiltonian_stencil(mj_ctx *c, int nofs, const int *ofs3, const double *blocks9)
{
    if (!c || nofs < 0 || (nofs > 0 && (!ofs3 || !blocks9))) MJ_FAIL(MJ_EINVAL, "mj_add_hamiltonian_stencil: bad argument");
    if (c->mm_generic) MJ_FAIL(MJ_ENOTSUP, "mj_add_hamiltonian_stencil: context already uses the generic path");
    mj_ctx::MMTerm term;
    for (int k = 0; k < nofs; ++k) {
        int o[3];
        for (int d = 0; d < 3; ++d) o[d] = canon(ofs3[3 * k + d], c->dim[d], c->periodic[d]);
        add_block(c, o[0], o[1], o[2], blocks9 + 9 * k);
        term.ofs3.insert(term.ofs3.end(), o, o + 3);
        term.blk9.insert(term.blk9.end(), blocks9 + 9 * k, blocks9 + 9 * k + 9);
    }
    c->mm_terms.push_back(std::move(term));
    c->has_mm = 1;
    c->dirty = 1;
    return MJ_OK;
}

extern "C" int mj_set_bfield(mj_ctx *c, const double *b, int64_t n3)
{
    if (!c || !b || (n3 != 3 && n3 != 3 * c->n)) MJ_FAIL(MJ_EINVAL, "mj_set_bfield: size must be 3 or 3*Ncell");
    c->bvec.assign(b, b + n3);
    c->dirty = 1;
    return MJ_OK;
}

extern "C" int mj_set_tile(mj_ctx *c, int bx, int by, int bz)
{
    if (!c || bx < 0 || by < 0 || bz < 0) MJ_FAIL(MJ_EINVAL, "mj_set_tile: bad argument");
    if (bx && by && bz && (bx * by * bz > 1024 || (bx * by * bz) % 64 != 0))
        MJ_FAIL(MJ_EINVAL, "mj_set_tile: threads per tile must be a multiple of 64 and <= 1024");
    c->bx = bx;
    c->by = by;
    c->bz = bz;
    c->dirty = 1;
    return MJ_OK;
}

static int pow2ceil(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// effective tile for the current lattice
static void tile_shape(const mj_ctx *c, int &bx, int &by, int &bz)
{
    if (c->bx && c->by && c->bz) {
        bx = c->bx;
        by = c->by;
        bz = c->bz;
        return;
    }
    // 256-site tiles (32x8); on small lattices smaller ones (down to one
    // wavefront) while there would be fewer than 512 workgroups
    int nthr = 256;
    while (nthr > 64 && c->n / nthr < 512) nthr >>= 1;
    bx = std::min(pow2ceil(c->dim[0]), std::max(8, nthr / 8));
    by = std::min(pow2ceil(c->dim[1]), std::max(1, nthr / bx));
    bz = std::min(pow2ceil(c->dim[2]), std::max(1, nthr / (bx * by)));
    while (bx * by * bz < 64 && (bx * by * bz) < c->n) {
        // keep at least one full wavefront per tile when the lattice allows it
        if (by < pow2ceil(c->dim[1])) by <<= 1;
        else if (bz < pow2ceil(c->dim[2])) bz <<= 1;
        else if (bx < pow2ceil(c->dim[0])) bx <<= 1;
        else break;
    }
}

int mj_finalize(mj_ctx *c)
{
    if (!c->dirty) return MJ_OK;
    if (!c->has_mm && !c->has_z && !c->dip) MJ_FAIL(MJ_ESTATE, "no Hamiltonian term registered");
    drop_graph(c);
    MJ_TRY(set_dev(c));
    hipStreamSynchronize(c->stream);
    // Zeeman
    c->bmode = 0;
    if (c->has_z) {
        if (c->bvec.empty()) MJ_FAIL(MJ_ESTATE, "Zeeman term registered but no B order parameter set");
        bool uniform = true;
        if (c->bvec.size() != 3) {
            for (int64_t i = 1; i < c->n && uniform; ++i)
                for (int a = 0; a < 3; ++a)
                    if (c->bvec[3 * i + a] != c->bvec[a]) uniform = false;
        }
        if (uniform) {
            c->bmode = 1;
            for (int a = 0; a < 3; ++a) {
                double t = 0.0;
                for (int q = 0; q < 3; ++q) t = t + c->bvec[q] * c->zblk[q * 3 + a];
                c->zb[a] = t;
            }
        } else {
            c->bmode = 2;
            if (!c->d_bmap) MJ_HIP_CHECK(hipMalloc(&c->d_bmap, sizeof(double) * 3 * c->np));
            std::vector<double> soa(3 * c->np, 0.0);
            for (int64_t i = 0; i < c->n; ++i)
                for (int a = 0; a < 3; ++a) soa[i + a * c->np] = c->bvec[3 * i + a];
            MJ_HIP_CHECK(hipMemcpy(c->d_bmap, soa.data(), sizeof(double) * soa.size(), hipMemcpyHostToDevice));
        }
    }
    // M-M operator
    if (c->has_mm && c->mm_generic) {
        // CSR of H + H^T (3N rows), columns ascending
        int64_t n3 = 3 * c->n;
        std::vector<std::pair<int64_t, int>> ent;
        std::vector<double> vv;
        size_t m = c->g_vals.size();
        std::vector<int64_t> cnt(n3 + 1, 0);
        for (size_t k = 0; k < m; ++k) {
            cnt[c->g_rows[k] + 1]++;
            cnt[c->g_cols[k] + 1]++;
        }
        for (int64_t r = 0; r < n3; ++r) cnt[r + 1] += cnt[r];
        if (cnt[n3] > INT32_MAX) MJ_FAIL(MJ_ENOTSUP, "generic operator too large");
        std::vector<int> ci(cnt[n3]);
        std::vector<double> v(cnt[n3]);
        std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
        for (size_t k = 0; k < m; ++k) {
            int64_t r = c->g_rows[k], q = c->g_cols[k];
            ci[pos[r]] = (int)q;
            v[pos[r]++] = c->g_vals[k];
            ci[pos[q]] = (int)r;
            v[pos[q]++] = c->g_vals[k];
        }
        // sort each row by column
        for (int64_t r = 0; r < n3; ++r) {
            int64_t b = cnt[r], e = cnt[r + 1];
            std::vector<std::pair<int, double>> row;
            for (int64_t k = b; k < e; ++k) row.push_back({ci[k], v[k]});
            std::stable_sort(row.begin(), row.end(), [](auto &x, auto &y) { return x.first < y.first; });
            for (int64_t k = b; k < e; ++k) {
                ci[k] = row[k - b].first;
                v[k] = row[k - b].second;
            }
        }
        std::vector<int> rp(n3 + 1);
        for (int64_t r = 0; r <= n3; ++r) rp[r] = (int)cnt[r];
        if (c->d_rp) hipFree(c->d_rp);
        if (c->d_ci) hipFree(c->d_ci);
        if (c->d_v) hipFree(c->d_v);
        c->d_rp = nullptr;
        c->d_ci = nullptr;
        c->d_v = nullptr;
        MJ_HIP_CHECK(hipMalloc(&c->d_rp, sizeof(int) * rp.size()));
        MJ_HIP_CHECK(hipMalloc(&c->d_ci, sizeof(int) * std::max<size_t>(1, ci.size())));
        MJ_HIP_CHECK(hipMalloc(&c->d_v, sizeof(double) * std::max<size_t>(1, v.size())));
        MJ_HIP_CHECK(hipMemcpy(c->d_rp, rp.data(), sizeof(int) * rp.size(), hipMemcpyHostToDevice));
        if (!ci.empty()) {
            MJ_HIP_CHECK(hipMemcpy(c->d_ci, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice));
            MJ_HIP_CHECK(hipMemcpy(c->d_v, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
        }
        c->n_iso = c->n_dm = c->n_gen = 0;
        c->rx = c->ry = c->rz = 0;
        c->use_jit = 0;
    } else {
        // stencil: C_d = A_d + A_{-d}^T
        std::vector<std::array<int, 3>> offs;
        std::vector<std::array<double, 9>> C;
        auto slot = [&](int dx, int dy, int dz) -> std::array<double, 9> & {
            std::array<int, 3> o = {dx, dy, dz};
            for (size_t k = 0; k < offs.size(); ++k)
                if (offs[k] == o) return C[k];
            offs.push_back(o);
            C.push_back({0, 0, 0, 0, 0, 0, 0, 0, 0});
            return C.back();
        };
        for (size_t k = 0; k < c->st_blk.size() / 9; ++k) {
            int dx = c->st_ofs[3 * k], dy = c->st_ofs[3 * k + 1], dz = c->st_ofs[3 * k + 2];
            const double *A = &c->st_blk[9 * k];
            auto &f = slot(dx, dy, dz);
            for (int q = 0; q < 9; ++q) f[q] += A[q];
            auto &g = slot(canon(-dx, c->dim[0], c->periodic[0]), canon(-dy, c->dim[1], c->periodic[1]),
                           canon(-dz, c->dim[2], c->periodic[2]));
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) g[3 * a + b] += A[3 * b + a];
        }
        int bx, by, bz;
        tile_shape(c, bx, by, bz);
        c->rx = c->ry = c->rz = 0;
        for (auto &o : offs) {
            c->rx = std::max(c->rx, std::abs(o[0]));
            c->ry = std::max(c->ry, std::abs(o[1]));
            c->rz = std::max(c->rz, std::abs(o[2]));
        }
        int sx = bx + 2 * c->rx, sy = by + 2 * c->ry;
        std::vector<int> ofs_iso, ofs_dm, ofs_gen;
        std::vector<double> cf_iso, cf_dm, cf_gen;
        std::vector<int> o3_iso, o3_dm, o3_gen;
        for (size_t k = 0; k < offs.size(); ++k) {
            const auto &q = C[k];
            double mx = 0.0;
            for (int t = 0; t < 9; ++t) mx = std::max(mx, fabs(q[t]));
            if (mx == 0.0) continue;
            double tol = 1.0e-14 * mx;
            int lofs = offs[k][2] * sx * sy + offs[k][1] * sx + offs[k][0];
            bool diag_eq = fabs(q[0] - q[4]) <= tol && fabs(q[0] - q[8]) <= tol;
            bool offd0 = fabs(q[1]) <= tol && fabs(q[2]) <= tol && fabs(q[3]) <= tol && fabs(q[5]) <= tol &&
                         fabs(q[6]) <= tol && fabs(q[7]) <= tol;
            bool anti = fabs(q[1] + q[3]) <= tol && fabs(q[2] + q[6]) <= tol && fabs(q[5] + q[7]) <= tol;
            if (diag_eq && offd0 && q[0] == q[4] && q[0] == q[8]) {
                ofs_iso.push_back(lofs);
                o3_iso.insert(o3_iso.end(), offs[k].begin(), offs[k].end());
                cf_iso.push_back(q[0]);
            } else if (diag_eq && anti && q[0] == q[4] && q[0] == q[8]) {
                // C m = s m + w x m with w = (C_zy, C_xz, C_yx)
                ofs_dm.push_back(lofs);
                o3_dm.insert(o3_dm.end(), offs[k].begin(), offs[k].end());
                cf_dm.push_back(q[0]);
                cf_dm.push_back(0.5 * (q[7] - q[5]));
                cf_dm.push_back(0.5 * (q[2] - q[6]));
                cf_dm.push_back(0.5 * (q[3] - q[1]));
            } else {
                ofs_gen.push_back(lofs);
                o3_gen.insert(o3_gen.end(), offs[k].begin(), offs[k].end());
                cf_gen.insert(cf_gen.end(), q.begin(), q.end());
            }
        }
        c->n_iso = (int)ofs_iso.size();
        c->n_dm = (int)ofs_dm.size();
        c->n_gen = (int)ofs_gen.size();
        std::vector<int> ofs(ofs_iso);
        ofs.insert(ofs.end(), ofs_dm.begin(), ofs_dm.end());
        ofs.insert(ofs.end(), ofs_gen.begin(), ofs_gen.end());
        std::vector<double> cf(cf_iso);
        cf.insert(cf.end(), cf_dm.begin(), cf_dm.end());
        cf.insert(cf.end(), cf_gen.begin(), cf_gen.end());
        c->h_ofs3 = o3_iso;
        c->h_ofs3.insert(c->h_ofs3.end(), o3_dm.begin(), o3_dm.end());
        c->h_ofs3.insert(c->h_ofs3.end(), o3_gen.begin(), o3_gen.end());
        c->h_cf = cf;
        if (c->d_ofs) hipFree(c->d_ofs);
        if (c->d_cf) hipFree(c->d_cf);
        c->d_ofs = nullptr;
        c->d_cf = nullptr;
        MJ_HIP_CHECK(hipMalloc(&c->d_ofs, sizeof(int) * std::max<size_t>(1, ofs.size())));
        MJ_HIP_CHECK(hipMalloc(&c->d_cf, sizeof(double) * std::max<size_t>(1, cf.size())));
        if (!ofs.empty()) {
            MJ_HIP_CHECK(hipMemcpy(c->d_ofs, ofs.data(), sizeof(int) * ofs.size(), hipMemcpyHostToDevice));
            MJ_HIP_CHECK(hipMemcpy(c->d_cf, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice));
        }
        // per-Hamiltonian compiled kernels: MJ_JIT=0 off, 1 on, otherwise for lattices >= 64x64 sites
        const char *env = getenv("MJ_JIT");
        c->use_jit = env && env[0] == '0' ? 0 : (env && env[0] == '1') ? 1 : (c->n >= 4096 ? 1 : 0);
        if (c->use_jit) {
            int ntx = (c->dim[0] + bx - 1) / bx, nty = (c->dim[1] + by - 1) / by, ntz = (c->dim[2] + bz - 1) / bz;
            MJ_TRY(mj_jit_prepare(c, ofs.data(), cf.data(), bx, by, bz, ntx * nty * ntz, ntx, nty));
        }
    }
    // energy partial buffers
    int bx, by, bz;
    tile_shape(c, bx, by, bz);
    int nb = ((c->dim[0] + bx - 1) / bx) * ((c->dim[1] + by - 1) / by) * ((c->dim[2] + bz - 1) / bz);
    if (c->epart_cap < nb) {
        if (c->epart) hipFree(c->epart);
        c->epart = nullptr;
        // two energy-partial buffers (alternating steps) + one for the |B| maxima, + 1 scalar
        MJ_HIP_CHECK(hipMalloc(&c->epart, sizeof(double) * (3 * (size_t)std::max(nb, 1024) + 1)));
        c->epart_cap = std::max(nb, 1024);
    }
    c->dirty = 0;
    return MJ_OK;
}

static void base_args(const mj_ctx *c, KArgs &a)
{
    memset(&a, 0, sizeof(a));
    a.nx = c->dim[0];
    a.ny = c->dim[1];
    a.nz = c->dim[2];
    a.px = c->periodic[0];
    a.py = c->periodic[1];
    a.pz = c->periodic[2];
    tile_shape(c, a.bx, a.by, a.bz);
    a.rx = c->rx;
    a.ry = c->ry;
    a.rz = c->rz;
    a.sx = a.bx + 2 * a.rx;
    a.sy = a.by + 2 * a.ry;
    a.sz = a.bz + 2 * a.rz;
    a.ntx = (a.nx + a.bx - 1) / a.bx;
    a.nty = (a.ny + a.by - 1) / a.by;
    a.ntz = (a.nz + a.bz - 1) / a.bz;
    a.nblocks = a.ntx * a.nty * a.ntz;
    a.np = c->np;
    a.n_iso = c->n_iso;
    a.n_dm = c->n_dm;
    a.n_gen = c->n_gen;
    a.ofs = c->d_ofs;
    a.cf = c->d_cf;
    a.use_stencil = c->mm_generic ? 0 : 1;
    a.bmode = c->bmode;
    for (int q = 0; q < 3; ++q) a.zb[q] = c->zb[q];
    for (int q = 0; q < 9; ++q) a.z[q] = c->zblk[q];
    a.bmap = c->d_bmap;
}

int mj_base_args(const mj_ctx *c, KArgs &a)
{
    base_args(c, a);
    return MJ_OK;
}

// field sources for a stage evaluated on state `src`: generic MM product and dipolar grid
static int prep_field(mj_ctx *c, const double *src, KArgs &a)
{
    if (c->mm_generic && c->has_mm) {
        MJ_TRY(mj_launch_spmv(c, src, c->Fb, c->stream));
        a.fld_in = c->Fb;
    } else if (c->mm_generic) {
        a.use_stencil = 0;
        hipMemsetAsync(c->Fb, 0, sizeof(double) * 3 * c->np, c->stream);
        a.fld_in = c->Fb;
    }
    if (c->dip) MJ_TRY(dip_apply(c, src, &a));
    return MJ_OK;
}

static int ensure_evals(mj_ctx *c, int64_t n)
{
    if (c->evals_cap >= n) return MJ_OK;
    if (c->evals) hipFree(c->evals);
    c->evals = nullptr;
    MJ_HIP_CHECK(hipMalloc(&c->evals, sizeof(double) * (size_t)n));
    c->evals_cap = n;
    return MJ_OK;
}

extern "C" int mj_set_state(mj_ctx *c, const double *m, int64_t n3)
{
    if (!c || !m || n3 != 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_set_state: size must be 3*Ncell");
    MJ_TRY(set_dev(c));
    MJ_HIP_CHECK(hipMemcpyAsync(c->Nb, m, sizeof(double) * n3, hipMemcpyHostToDevice, c->stream));
    MJ_TRY(mj_launch_aos2soa(c->Nb, c->A, c->n, c->np, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->state_set = 1;
    c->field_valid = 0;
    return MJ_OK;
}

extern "C" int mj_set_state_device(mj_ctx *c, const double *m, int64_t n3)
{
    if (!c || !m || n3 != 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_set_state_device: size must be 3*Ncell");
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_launch_aos2soa(m, c->A, c->n, c->np, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    c->state_set = 1;
    c->field_valid = 0;
    return MJ_OK;
}

extern "C" int mj_get_state(mj_ctx *c, double *m, int64_t n3)
{
    if (!c || !m || n3 != 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_get_state: size must be 3*Ncell");
    if (!c->state_set) MJ_FAIL(MJ_ESTATE, "mj_get_state: no state set");
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_launch_soa2aos(c->A, c->Nb, c->n, c->np, 1.0e-50, c->stream));
    MJ_HIP_CHECK(hipMemcpyAsync(m, c->Nb, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

extern "C" int mj_get_state_device(mj_ctx *c, double *m, int64_t n3)
{
    if (!c || !m || n3 != 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_get_state_device: size must be 3*Ncell");
    if (!c->state_set) MJ_FAIL(MJ_ESTATE, "mj_get_state_device: no state set");
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_launch_soa2aos(c->A, m, c->n, c->np, 1.0e-50, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

extern "C" int mj_eff_field(mj_ctx *c, double *B, int64_t n3)
{
    if (!c || !B || n3 != 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_eff_field: size must be 3*Ncell");
    if (!c->state_set) MJ_FAIL(MJ_ESTATE, "mj_eff_field: no state set");
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_finalize(c));
    KArgs a;
    base_args(c, a);
    a.src = c->A;
    MJ_TRY(prep_field(c, c->A, a));
    a.fld_out = c->Bb;
    MJ_TRY(launch(c, MODE_FIELD, false, a));
    MJ_TRY(mj_launch_soa2aos(c->Bb, c->Nb, c->n, c->np, 0.0, c->stream));
    MJ_HIP_CHECK(hipMemcpyAsync(B, c->Nb, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

extern "C" int mj_keep_last_field(mj_ctx *c, int on)
{
    if (!c) MJ_FAIL(MJ_EINVAL, "mj_keep_last_field: null context");
    c->keep_field = on != 0;
    c->field_valid = 0;
    return MJ_OK;
}

extern "C" int mj_last_field(mj_ctx *c, double *B, int64_t n3)
{
    if (!c || !B || n3 != 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_last_field: size must be 3*Ncell");
    if (!c->field_valid)
        MJ_FAIL(MJ_ESTATE, "mj_last_field: no step has run since mj_keep_last_field(ctx, 1) or mj_set_state");
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_launch_soa2aos(c->Lf, c->Nb, c->n, c->np, 0.0, c->stream));
    MJ_HIP_CHECK(hipMemcpyAsync(B, c->Nb, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

static int energy_async(mj_ctx *c, double *dst_dev, double scale)
{
    KArgs a;
    base_args(c, a);
    a.src = c->A;
    MJ_TRY(prep_field(c, c->A, a));
    a.energy = 1;
    a.epart = c->epart;
    MJ_TRY(launch(c, MODE_FIELD, false, a));
    return mj_launch_reduce(c->epart, a.nblocks, dst_dev, scale, c->stream);
}

int mj_energy_into(mj_ctx *c, double *dst_dev) { return energy_async(c, dst_dev, 1.0); }

extern "C" int mj_energy(mj_ctx *c, double *E)
{
    if (!c || !E) MJ_FAIL(MJ_EINVAL, "mj_energy: null argument");
    if (!c->state_set) MJ_FAIL(MJ_ESTATE, "mj_energy: no state set");
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_finalize(c));
    MJ_TRY(ensure_evals(c, 1));
    MJ_TRY(energy_async(c, c->evals, 1.0));
    MJ_HIP_CHECK(hipMemcpyAsync(E, c->evals, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

int mj_launch_term_energy(const TermArgs &t, int nblocks, hipStream_t s);

// device copy of registered M-M term k (offsets, 3x3 blocks), made once: terms
// are only ever appended, so a copy stays valid for the context's lifetime
static int term_device(mj_ctx *c, size_t k, const int **ofs, const double **blk)
{
    if (c->term_ofs.size() < c->mm_terms.size()) {
        c->term_ofs.resize(c->mm_terms.size(), nullptr);
        c->term_blk.resize(c->mm_terms.size(), nullptr);
    }
    if (!c->term_ofs[k]) {
        const auto &tm = c->mm_terms[k];
        int *d_ofs = nullptr;
        double *d_blk = nullptr;
        MJ_HIP_CHECK(hipMalloc(&d_ofs, sizeof(int) * std::max<size_t>(1, tm.ofs3.size())));
        if (hipMalloc(&d_blk, sizeof(double) * std::max<size_t>(1, tm.blk9.size())) != hipSuccess) {
            hipFree(d_ofs);
            MJ_FAIL(MJ_ENOMEM, "term copy: device allocation failed");
        }
        if (!tm.ofs3.empty() &&
            (hipMemcpy(d_ofs, tm.ofs3.data(), sizeof(int) * tm.ofs3.size(), hipMemcpyHostToDevice) != hipSuccess ||
             hipMemcpy(d_blk, tm.blk9.data(), sizeof(double) * tm.blk9.size(), hipMemcpyHostToDevice) != hipSuccess)) {
            hipFree(d_ofs);
            hipFree(d_blk);
            MJ_FAIL(MJ_EHIP, "term copy: copy to the device failed");
        }
        c->term_ofs[k] = d_ofs;
        c->term_blk[k] = d_blk;
    }
    *ofs = c->term_ofs[k];
    *blk = c->term_blk[k];
    return MJ_OK;
}

// energy of every registered term of the current state, the reference's
// energy_resolved (hamiltonian_collection.f90:365-388): the M-M terms in the
// order they were registered, then the Zeeman term, then the FFT term
extern "C" int mj_energy_terms(mj_ctx *c, double *E, int *nterms)
{
    if (!c || !E || !nterms) MJ_FAIL(MJ_EINVAL, "mj_energy_terms: null argument");
    if (!c->state_set) MJ_FAIL(MJ_ESTATE, "mj_energy_terms: no state set");
    if (c->mm_terms_generic) MJ_FAIL(MJ_ENOTSUP, "mj_energy_terms: a term is not translation invariant");
    const int count = (int)c->mm_terms.size() + (c->has_z ? 1 : 0) + (c->dip ? 1 : 0);
    if (*nterms < count) {
        *nterms = count;
        MJ_FAIL(MJ_EINVAL, "mj_energy_terms: the array holds fewer entries than there are terms (count returned)");
    }
    *nterms = count;
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_finalize(c));
    MJ_TRY(ensure_evals(c, count));
    TermArgs t;
    memset(&t, 0, sizeof t);
    t.m = c->A;
    t.np = c->np;
    for (int d = 0; d < 3; ++d) {
        t.dim[d] = c->dim[d];
        t.per[d] = c->periodic[d];
    }
    t.part = c->epart;
    const int nb = (int)std::min<int64_t>((c->n + 255) / 256, c->epart_cap);   // grid-stride beyond
    int q = 0;
    for (size_t k = 0; k < c->mm_terms.size(); ++k) {
        const int nofs = (int)c->mm_terms[k].blk9.size() / 9;
        const int *d_ofs = nullptr;
        const double *d_blk = nullptr;
        MJ_TRY(term_device(c, k, &d_ofs, &d_blk));
        TermArgs a = t;
        a.mode = 0;
        a.nofs = nofs;
        a.ofs3 = d_ofs;
        a.blk9 = d_blk;
        MJ_TRY(mj_launch_term_energy(a, nb, c->stream));
        MJ_TRY(mj_launch_reduce(c->epart, nb, c->evals + q++, 1.0, c->stream));
    }
    if (c->has_z) {
        TermArgs a = t;
        a.mode = 1;
        for (int k = 0; k < 9; ++k) a.z[k] = c->zblk[k];
        a.bmode = c->bmode;
        for (int k = 0; k < 3; ++k) a.b[k] = c->bvec.size() >= 3 ? c->bvec[k] : 0.0;
        a.bmap = c->d_bmap;
        MJ_TRY(mj_launch_term_energy(a, nb, c->stream));
        MJ_TRY(mj_launch_reduce(c->epart, nb, c->evals + q++, 1.0, c->stream));
    }
    if (c->dip) {
        KArgs ka;
        base_args(c, ka);
        MJ_TRY(dip_apply(c, c->A, &ka));
        TermArgs a = t;
        a.mode = 2;
        a.hd = ka.hdip;
        a.dnx = ka.dnx;
        a.dny = ka.dny;
        a.hstride = ka.hstride;
        MJ_TRY(mj_launch_term_energy(a, nb, c->stream));
        MJ_TRY(mj_launch_reduce(c->epart, nb, c->evals + q++, 1.0, c->stream));
    }
    MJ_HIP_CHECK(hipMemcpyAsync(E, c->evals, sizeof(double) * count, hipMemcpyDeviceToHost, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

int mj_launch_term_mult(const TermArgs &t, int64_t n, hipStream_t s);

// product of one registered term with a vector, the backends' mult_r / mult_l
// (ham_base.f90:41,88-96; ham_dense.f90:30-87): res = alpha H vec + beta res
// (side 0) or alpha H^T vec + beta res (side 1); terms as mj_energy_terms
// numbers them (M-M terms in registration order, then the Zeeman term)
extern "C" int mj_term_mult(mj_ctx *c, int term, int side, const double *vec, int64_t nvec, double *res,
                            int64_t nres, double alpha, double beta)
{
    if (!c || !vec || !res) MJ_FAIL(MJ_EINVAL, "mj_term_mult: null argument");
    if (side != 0 && side != 1) MJ_FAIL(MJ_EINVAL, "mj_term_mult: side must be 0 (H v) or 1 (H^T v)");
    if (nvec != 3 * c->n || nres != 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_term_mult: vectors must hold 3*Ncell values");
    if (c->mm_terms_generic) MJ_FAIL(MJ_ENOTSUP, "mj_term_mult: a term is not translation invariant");
    const int n_mm = (int)c->mm_terms.size();
    if (term < 0 || term > n_mm || (term == n_mm && !c->has_z))
        MJ_FAIL(MJ_EINVAL, "mj_term_mult: no such term");
    MJ_TRY(set_dev(c));
    TermArgs t;
    memset(&t, 0, sizeof t);
    for (int d = 0; d < 3; ++d) {
        t.dim[d] = c->dim[d];
        t.per[d] = c->periodic[d];
    }
    t.alpha = alpha;
    t.beta = beta;
    t.side = side;
    const size_t bytes = sizeof(double) * 3 * (size_t)c->n;
    // staging vectors and the term's device copy are made once and kept
    if (!c->tm_in) MJ_HIP_CHECK(hipMalloc(&c->tm_in, bytes));
    if (!c->tm_out) MJ_HIP_CHECK(hipMalloc(&c->tm_out, bytes));
    if (term < n_mm) {
        t.mode = 3;
        t.nofs = (int)c->mm_terms[term].blk9.size() / 9;
        const int *d_ofs = nullptr;
        const double *d_blk = nullptr;
        MJ_TRY(term_device(c, (size_t)term, &d_ofs, &d_blk));
        t.ofs3 = d_ofs;
        t.blk9 = d_blk;
    } else {
        t.mode = 4;
        for (int k = 0; k < 9; ++k) t.z[k] = c->zblk[k];
    }
    MJ_HIP_CHECK(hipMemcpyAsync(c->tm_in, vec, bytes, hipMemcpyHostToDevice, c->stream));
    if (beta != 0.0) MJ_HIP_CHECK(hipMemcpyAsync(c->tm_out, res, bytes, hipMemcpyHostToDevice, c->stream));
    t.vin = c->tm_in;
    t.vout = c->tm_out;
    MJ_TRY(mj_launch_term_mult(t, c->n, c->stream));
    MJ_HIP_CHECK(hipMemcpyAsync(res, c->tm_out, bytes, hipMemcpyDeviceToHost, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

int mj_launch_math_probe(int fn, int64_t n, const double *x, double *y, hipStream_t s);

extern "C" int mj_math_probe(mj_ctx *c, int fn, int64_t n, const double *x, double *y)
{
    if (!c || !x || !y || n < 0 || fn < 0 || fn > 3) MJ_FAIL(MJ_EINVAL, "mj_math_probe: bad argument");
    if (n == 0) return MJ_OK;
    MJ_TRY(set_dev(c));
    double *d = nullptr;
    MJ_HIP_CHECK(hipMalloc(&d, sizeof(double) * 2 * (size_t)n));
    int rc = MJ_OK;
    if (hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess) {
        mj_set_error("mj_math_probe: copy in");
        rc = MJ_EHIP;
    }
    if (rc == MJ_OK) rc = mj_launch_math_probe(fn, n, d, d + n, c->stream);
    if (rc == MJ_OK && (hipMemcpyAsync(y, d + n, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                        hipStreamSynchronize(c->stream) != hipSuccess)) {
        mj_set_error("mj_math_probe: copy out");
        rc = MJ_EHIP;
    }
    hipFree(d);
    return rc;
}

extern "C" int mj_thermal_normals(mj_ctx *c, uint64_t seed, uint64_t step, double *out, int64_t n3)
{
    if (!c || !out || n3 != 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_thermal_normals: size must be 3*Ncell");
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_launch_normals(seed, step, c->n, c->Nb, c->stream));
    MJ_HIP_CHECK(hipMemcpyAsync(out, c->Nb, sizeof(double) * n3, hipMemcpyDeviceToHost, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

// ---------------------------------------------------------------------------
// the step loop
// ---------------------------------------------------------------------------
// Timing of the step loop's kernels (mj_profile_steps).  Steps of one kernel
// (the fused integration-2 kernel, Euler) are timed as one block of
// back-to-back launches (`block`): per-launch events would put the host's
// submission gaps into each launch.  Steps of two kernels get an event pair
// around every launch.
struct Prof {
    std::vector<hipEvent_t> ev;
    std::vector<int> kind;
    int used = 0;
    bool block = false;
    int block_kind = 1, block_count = 0;
    hipEvent_t b0 = nullptr, b1 = nullptr;   // around the steps' launches (block timing)
};

static int launch(mj_ctx *c, int mode, bool noise, const KArgs &a)
{
    // integration 4 runs the ahead-of-time stage kernel (no per-Hamiltonian form)
    if (c->use_jit && a.use_stencil && mode != MODE_RKS) return mj_jit_launch(c, mode, noise, a, c->stream);
    return mj_launch_stage(mode, noise, a, c->stream);
}

static int launch_timed(mj_ctx *c, int mode, bool noise, const KArgs &a, Prof *p, int kind)
{
    if (!p) return launch(c, mode, noise, a);
    if (p->block) {
        p->block_kind = kind;
        p->block_count += 1;
        return launch(c, mode, noise, a);
    }
    if ((size_t)p->used + 2 > p->ev.size()) {
        for (int k = 0; k < 64; ++k) {
            hipEvent_t e;
            MJ_HIP_CHECK(hipEventCreate(&e));
            p->ev.push_back(e);
        }
    }
    MJ_HIP_CHECK(hipEventRecord(p->ev[p->used], c->stream));
    int rc = launch(c, mode, noise, a);
    MJ_HIP_CHECK(hipEventRecord(p->ev[p->used + 1], c->stream));
    p->kind.push_back(kind);
    p->used += 2;
    return rc;
}

static const double kPi = 3.14159265358979323846264338327950288;   // constant.f90:32

// reduce the per-block |B| maxima, read it back and set the next timestep
static int next_timestep(mj_ctx *c, const double *bpart, int nblocks, double *bmax_dev, double disc, double &ts)
{
    MJ_TRY(mj_launch_reduce_max(bpart, nblocks, bmax_dev, c->stream));
    double bmax = 0.0;
    MJ_HIP_CHECK(hipMemcpyAsync(&bmax, bmax_dev, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    if (!(bmax > 1.0e-8)) MJ_FAIL(MJ_EINVAL, "update_time: max |B_eff| <= 1e-8 (the reference stops here)");
    ts = 2.0 * kPi * kHbar / bmax / disc;
    return MJ_OK;
}

// Adaptive timestep (update_time.f90:41-75), off unless disc > 0: after every
// step the timestep becomes 2 pi hbar / max_i |B_i| / disc, with B the field
// of the step's last stage (spindynamics.f90:438); ts_io carries the value
// between calls and ts_used (may be null) receives each step's timestep.
struct Adapt {
    double disc = 0.0;
    double *ts_io = nullptr;
    double *ts_used = nullptr;
};

// time-dependent B order parameter: b_i = sum_k w_k (x) F_k,i per stage
// (nmaps = 0: b uniform, w = b), w at [step][stage][3 * max(nmaps, 1)];
// w_end at [step][...]: the field of the energy after each step
struct Drive {
    int nmaps = 0;
    const double *w = nullptr;
    const double *w_end = nullptr;
};

static void apply_drive(const mj_ctx *c, KArgs &k, int nmaps, const double *w)
{
    if (nmaps == 0) {
        k.bmode = 1;
        for (int q = 0; q < 3; ++q)
            k.zb[q] = (w[0] * c->zblk[q] + w[1] * c->zblk[3 + q]) + w[2] * c->zblk[6 + q];
    } else {
        k.bmode = 3;
        k.nmaps = nmaps;
        for (int q = 0; q < 3 * nmaps; ++q) k.wmap[q] = w[q];
        k.maps = c->d_maps;
    }
}

// Butcher table of get_butcher_explicit (solver_order.f90:65-127) for
// integration 4, 1-based as there: T[row][col] for rows/cols 1..N+1; row 1 the
// factor of the stage timestep, rows 2..N+1 the weights of the stages' D.
static bool rk_butcher(int N, double T[9][9])
{
    for (int r = 0; r < 9; ++r)
        for (int q = 0; q < 9; ++q) T[r][q] = 0.0;
    auto col = [&](int q, std::initializer_list<double> v) {
        int r = 1;
        for (double x : v) T[r++][q] = x;
    };
    switch (N) {
    case 1:
        col(2, {1.0, 1.0});
        return true;
    case 2:
        col(2, {1.0, 1.0});
        col(3, {1.0, 0.5, 0.5});
        return true;
    case 3:
        col(2, {0.5, 0.5});
        col(3, {1.0, -1.0, 2.0, 0.0});
        col(4, {6.0 / 6.0, 1.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0});
        return true;
    case 4:
        col(2, {0.5, 0.5});
        col(3, {0.5, 0.0, 0.5});
        col(4, {1.0, 0.0, 0.0, 1.0});
        col(5, {6.0 / 6.0, 1.0 / 6.0, 2.0 / 6.0, 2.0 / 6.0, 1.0 / 6.0});
        return true;
    case 6:
        col(2, {0.25, 0.25});
        col(3, {12.0 / 32.0, 3.0 / 32.0, 9.0 / 32.0});
        col(4, {2028.0 / 2197.0, 1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0});
        col(5, {1.0, 439.0 / 216.0, -8.0, 3680.0 / 513.0, -845.0 / 4104.0});
        col(6, {0.5, -8.0 / 27.0, 2.0, -3544.0 / 2565.0, 1859.0 / 4104.0, -11.0 / 40.0});
        col(7, {1.0, 25.0 / 216.0, 0.0, 1408.0 / 2565.0, 2197.0 / 4104.0, -1.0 / 5.0, 0.0});
        return true;
    case 7:
        col(2, {1.0 / 5.0, 1.0 / 5.0});
        col(3, {12.0 / 40.0, 3.0 / 40.0, 9.0 / 40.0});
        col(4, {4.0 / 5.0, 44.0 / 45.0, -56.0 / 15.0, 32.0 / 9.0});
        col(5, {8.0 / 9.0, 19372.0 / 6561.0, -25360.0 / 2187.0, 64448.0 / 6561.0, -212.0 / 729.0});
        col(6, {1.0, 9017.0 / 3168.0, -355.0 / 33.0, 46732.0 / 5247.0, 49.0 / 176.0, -5103.0 / 18656.0});
        col(7, {1.0, 35.0 / 384.0, 0.0, 500.0 / 1113.0, 125.0 / 192.0, -2187.0 / 6784.0, 11.0 / 84.0});
        col(8, {1.0, 5179.0 / 57600.0, 0.0, 7571.0 / 16695.0, 393.0 / 640.0, -92097.0 / 339200.0, 187.0 / 2100.0,
                1.0 / 40.0});
        return true;
    default:
        return false;
    }
}

extern "C" int mj_set_rk_order(mj_ctx *c, int n_order)
{
    double T[9][9];
    if (!c) MJ_FAIL(MJ_EINVAL, "mj_set_rk_order: null context");
    if (!rk_butcher(n_order, T)) MJ_FAIL(MJ_EINVAL, "mj_set_rk_order: N_order must be 1, 2, 3, 4, 6 or 7");
    c->rk_order = n_order;
    return MJ_OK;
}

static int run_steps(mj_ctx *c, int nsteps, int integ, double timestep, double alpha, double kT,
                     uint64_t seed, uint64_t step0, const double *noise_host, double *energies, Prof *prof,
                     const Adapt *ad = nullptr, const Drive *dr = nullptr, const double *kT_steps = nullptr)
{
    // kT_steps: one kT per step (the thermal amplitude formed per step); kT is
    // then their maximum, which decides whether the steps carry a thermal term
    if (kT_steps) {
        kT = 0.0;
        for (int j = 0; j < nsteps; ++j) {
            if (!(kT_steps[j] >= 0.0)) MJ_FAIL(MJ_EINVAL, "kT per step must be >= 0");
            kT = std::max(kT, kT_steps[j]);
        }
    }
    if (nsteps < 0) MJ_FAIL(MJ_EINVAL, "nsteps must be >= 0");
    if (integ != MJ_EULER && integ != MJ_HEUN && integ != MJ_RK && integ != MJ_SIB && integ != MJ_DEPONDT)
        MJ_FAIL(MJ_EINVAL, "unknown integrator");
    if (integ == MJ_RK && kT * alpha > 1.0e-10)
        MJ_FAIL(MJ_ENOTSUP, "integration 4 has no thermal field (the reference selects none for it, "
                            "solver_commun.f90:112-121)");
    if (integ == MJ_RK && dr) MJ_FAIL(MJ_ENOTSUP, "integration 4 with a time-dependent field is not supported");
    if (!c->state_set) MJ_FAIL(MJ_ESTATE, "no state set");
    if (!(timestep > 0.0) || !(alpha >= 0.0) || !(kT >= 0.0)) MJ_FAIL(MJ_EINVAL, "timestep/damping/kT out of range");
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_finalize(c));
    if (nsteps == 0) return MJ_OK;
    const bool noisy = kT * alpha > 1.0e-10;
    const int nl = (integ == MJ_EULER) ? 1 : (integ == MJ_RK) ? c->rk_order : 2;
    double rkT[9][9];
    if (integ == MJ_RK) {
        rk_butcher(c->rk_order, rkT);
        for (int k = 0; k < c->rk_order - 1; ++k)   // before any capture
            if (!c->Drk[k]) MJ_HIP_CHECK(hipMalloc(&c->Drk[k], sizeof(double) * 3 * (size_t)c->np));
    }
    const bool adapt = ad && ad->disc > 0.0;
    double ts = timestep;
    auto stage_dt = [&](double t) { return (integ == MJ_EULER || integ == MJ_HEUN) ? t / (double)nl : t; };
    auto sigma_kt = [&](double dt_stage, double kt) {
        // reference: sqrt(damping*kT*hbar/dt) with dt = timestep/N_loop (spindynamics.f90:266-269),
        // no thermal term where kT*damping <= 1e-10 (:268);
        // SIB/Depondt (Stratonovich, field in both stages): 2*damping*kT*hbar/dt
        if (!(kt * alpha > 1.0e-10)) return 0.0;
        return (integ == MJ_EULER || integ == MJ_HEUN) ? sqrt(alpha * kt * kHbar / dt_stage)
                                                       : sqrt(2.0 * alpha * kt * kHbar / dt_stage);
    };
    auto sigma_of = [&](double dt_stage) { return noisy ? sigma_kt(dt_stage, kT) : 0.0; };
    if (energies) MJ_TRY(ensure_evals(c, nsteps));
    KArgs base;
    base_args(c, base);
    base.alpha = alpha;
    base.den = 1.0 + alpha * alpha;
    base.rden = 1.0 / base.den;
    base.dt = stage_dt(ts);
    base.sigma = sigma_of(base.dt);
    base.seed = seed;
    const double escale = 1.0 / (double)c->n;
    double *bpart = c->epart + 2 * (size_t)c->epart_cap;
    double *bmax_dev = bpart + c->epart_cap;

    // integration 2 as one kernel per step (both stages, halo recomputed in
    // LDS): per-Hamiltonian kernels on the stencil path with a static field,
    // the thermal field drawn in-kernel and no per-step host work (MJ_FUSE=0
    // runs the two stage kernels instead; both give the same bits)
    const char *fenv = getenv("MJ_FUSE");
    const bool fused = (integ == MJ_HEUN || integ == MJ_SIB || integ == MJ_DEPONDT) && c->use_jit && !c->mm_generic &&
                       !c->dip && !dr && !adapt && !noise_host && c->bmode != 3 && mj_jit_fused_ok(c) &&
                       !(fenv && fenv[0] == '0');
    const int fmode = integ == MJ_SIB ? MODE_SIBF : integ == MJ_DEPONDT ? MODE_DEPF : MODE_HEUNF;

    // the launches of one step; e_dst: where stage 2 (or the reduce) puts the
    // energy of the state entering the step, or null
    const int wlen = dr ? 3 * std::max(dr->nmaps, 1) : 0;
    // fused steps: the partials of step j (the state entering it) go to buffer
    // j & 1 and the launch of step j + 1 folds them; fuse_fold is where the
    // next launch puts that energy (null: nothing to fold)
    double *fuse_fold = nullptr;
    // the last stage of the call's last step writes its field when kept
    // (spindynamics.f90:298 evaluates it, :400 prints it)
    c->field_valid = 0;
    if (c->keep_field && !c->Lf) MJ_HIP_CHECK(hipMalloc(&c->Lf, sizeof(double) * 3 * (size_t)c->np));
    double *const lf = c->keep_field ? c->Lf : nullptr;
    bool last_step = false;
    auto enqueue_step = [&](KArgs a, int parity, double *e_dst) -> int {
        const bool e_here = e_dst != nullptr;
        double *const fo = last_step ? lf : nullptr;
        double *ep = c->epart + (size_t)(parity & 1) * c->epart_cap;
        const double *wj = dr ? dr->w + (size_t)parity * nl * wlen : nullptr;   // parity = step index here
        if (fused) {
            KArgs s = a;
            s.src = c->A;
            s.dst = c->Bb;   // other tiles read m0 around this one: out of place
            if (e_here) {
                s.energy = 1;
                s.epart = ep;
            }
            if (fuse_fold) {
                s.ered_in = c->epart + (size_t)((parity + 1) & 1) * c->epart_cap;
                s.n_ered = s.nblocks;
                s.e_out = fuse_fold;
                s.e_scale = escale;
            }
            s.fld_out = fo;
            MJ_TRY(launch_timed(c, fmode, noisy, s, prof, 1));
            fuse_fold = e_dst;
            std::swap(c->A, c->Bb);
            return MJ_OK;
        }
        if (integ == MJ_RK) {
            // spindynamics.f90:282-318: every stage from the previous stage's state,
            // dt_i = Butcher(1, i+1) dt_{i-1} from timestep / N (get_dt_mode)
            const int N = nl;
            double dt = a.dt / (double)N;   // a.dt: the step's timestep here (timestep_int / real(N_loop))
            for (int i = 1; i <= N; ++i) {
                dt = rkT[1][i + 1] * dt;
                KArgs s = a;
                s.dt = dt;
                s.src = c->A;
                MJ_TRY(prep_field(c, c->A, s));
                s.dst = c->Bb;
                s.nprev = i - 1;
                for (int k = 0; k < i - 1; ++k) s.dprev[k] = c->Drk[k];
                for (int k = 0; k < i; ++k) s.wrk[k] = rkT[k + 2][i + 1];
                s.dout = i < N ? c->Drk[i - 1] : nullptr;
                if (e_here && i == 1) {
                    s.energy = 1;
                    s.epart = ep;
                }
                if (e_here && i == N && N > 1) {
                    s.ered_in = ep;
                    s.n_ered = s.nblocks;
                    s.e_out = e_dst;
                    s.e_scale = escale;
                }
                if (adapt && i == N) s.bmax_part = bpart;
                if (i == N) s.fld_out = fo;
                MJ_TRY(launch_timed(c, MODE_RKS, false, s, prof, i == N ? 1 : 0));
                std::swap(c->A, c->Bb);
            }
            if (e_here && N == 1) MJ_TRY(mj_launch_reduce(ep, a.nblocks, e_dst, escale, c->stream));
            if (adapt) MJ_TRY(next_timestep(c, bpart, a.nblocks, bmax_dev, ad->disc, ts));
            return MJ_OK;
        }
        if (integ == MJ_EULER) {
            KArgs s1 = a;
            if (dr) apply_drive(c, s1, dr->nmaps, wj);
            s1.src = c->A;
            MJ_TRY(prep_field(c, c->A, s1));
            s1.dst = c->Bb;
            if (e_here) {
                s1.energy = 1;
                s1.epart = ep;
            }
            if (adapt) s1.bmax_part = bpart;
            s1.fld_out = fo;
            MJ_TRY(launch_timed(c, MODE_EULER, noisy, s1, prof, 0));
            if (e_here) MJ_TRY(mj_launch_reduce(ep, s1.nblocks, e_dst, escale, c->stream));
            std::swap(c->A, c->Bb);
            if (adapt) MJ_TRY(next_timestep(c, bpart, s1.nblocks, bmax_dev, ad->disc, ts));
            return MJ_OK;
        }
        int m1 = integ == MJ_HEUN ? MODE_HEUN1 : integ == MJ_SIB ? MODE_SIB1 : MODE_DEP1;
        int m2 = integ == MJ_HEUN ? MODE_HEUN2 : integ == MJ_SIB ? MODE_SIB2 : MODE_DEP2;
        KArgs s1 = a;
        if (dr) apply_drive(c, s1, dr->nmaps, wj);
        s1.src = c->A;
        MJ_TRY(prep_field(c, c->A, s1));
        s1.dst = c->Bb;
        s1.aux = c->Db;
        if (e_here) {
            s1.energy = 1;
            s1.epart = ep;
        }
        // integration 2 adds the thermal term in its last stage only
        MJ_TRY(launch_timed(c, m1, noisy && integ != MJ_HEUN, s1, prof, 0));
        KArgs s2 = a;
        if (dr) apply_drive(c, s2, dr->nmaps, wj + wlen);
        s2.src = c->Bb;
        MJ_TRY(prep_field(c, c->Bb, s2));
        s2.own = c->A;
        s2.aux = c->Db;
        s2.dst = c->A;   // own-site read then write: safe in place
        if (e_here) {
            s2.ered_in = ep;
            s2.n_ered = s1.nblocks;
            s2.e_out = e_dst;
            s2.e_scale = escale;
        }
        if (adapt) s2.bmax_part = bpart;
        s2.fld_out = fo;
        MJ_TRY(launch_timed(c, m2, noisy, s2, prof, 1));
        if (adapt) MJ_TRY(next_timestep(c, bpart, s2.nblocks, bmax_dev, ad->disc, ts));
        return MJ_OK;
    };

    int j = 0;
    // Chunks of kGraphChunk steps replay one captured graph: the thermal-field
    // step comes from a device counter, energies land in a staging array that
    // is copied into place after each replay.  Only for the stencil path
    // without FFT term (no memset or library call in the captured stream),
    // without per-step host work (adaptive timestep, injected noise, timing).
    // Not for the fused integration-2 kernel unless MJ_GRAPH=1: with one launch
    // per step the eager launches keep the GPU as busy (1024x1024 triangular:
    // 0.045 ms per call + 41.7 us per step eager, 0.177 ms + 42.2 us replayed;
    // 64x64 and 256x256 equal within 1%).  MJ_GRAPH=0 never replays.
    const char *genv = getenv("MJ_GRAPH");
    const bool graph_ok = !prof && !adapt && !noise_host && !kT_steps && !dr && !c->mm_generic && !c->dip &&
                          nsteps >= 2 * kGraphChunk &&
                          !(genv && genv[0] == '0') && (!fused || (genv && genv[0] == '1'));
    if (graph_ok) {
        const double key_arr[] = {(double)integ, energies ? 1.0 : 0.0, noisy ? 1.0 : 0.0, timestep, alpha, kT,
                                  (double)(seed >> 32), (double)(seed & 0xffffffffu), (double)(uintptr_t)c->A,
                                  (double)(uintptr_t)c->Bb, (double)(uintptr_t)c->epart, (double)c->use_jit,
                                  fused ? 1.0 : 0.0};
        std::vector<double> key(key_arr, key_arr + sizeof key_arr / sizeof key_arr[0]);
        if (key != c->gkey) drop_graph(c);
        if (!c->d_step) MJ_HIP_CHECK(hipMalloc(&c->d_step, sizeof(uint64_t)));
        if (!c->estage) MJ_HIP_CHECK(hipMalloc(&c->estage, sizeof(double) * (kGraphChunk + 1)));
        if (!c->gexec) {
            if (c->use_jit && fused) {   // compile before capturing
                MJ_TRY(mj_jit_resolve(c, fmode, noisy));
            } else if (c->use_jit) {
                const int modes[4][2] = {{MODE_EULER, -1}, {MODE_HEUN1, MODE_HEUN2}, {MODE_SIB1, MODE_SIB2},
                                         {MODE_DEP1, MODE_DEP2}};
                const int row = integ == MJ_EULER ? 0 : integ == MJ_HEUN ? 1 : integ == MJ_SIB ? 2 : 3;
                for (int q = 0; q < 2; ++q)
                    if (modes[row][q] >= 0)
                        MJ_TRY(mj_jit_resolve(c, modes[row][q], noisy && !(integ == MJ_HEUN && q == 0)));
            }
            hipGraph_t g = nullptr;
            MJ_HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            int rc = MJ_OK;
            // fused steps: launch 0 folds the partials of the step before the
            // replay (the previous replay's last) into the extra slot
            fuse_fold = (fused && energies) ? c->estage + kGraphChunk : nullptr;
            for (int l = 0; l < kGraphChunk && rc == MJ_OK; ++l) {
                KArgs a = base;
                a.step = (uint64_t)l;
                a.stepp = c->d_step;
                rc = enqueue_step(a, l, energies ? c->estage + l : nullptr);
            }
            fuse_fold = nullptr;
            if (rc == MJ_OK) rc = mj_launch_step_counter(c->d_step, (uint64_t)kGraphChunk, 1, c->stream);
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (rc != MJ_OK || e != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                if (rc != MJ_OK) return rc;
                MJ_FAIL(MJ_EHIP, std::string("step graph capture: ") + hipGetErrorString(e));
            }
            e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                c->gexec = nullptr;
                MJ_FAIL(MJ_EHIP, std::string("step graph: ") + hipGetErrorString(e));
            }
            c->gkey = key;
        }
        MJ_TRY(mj_launch_step_counter(c->d_step, step0, 0, c->stream));
        for (; nsteps - j >= kGraphChunk + (lf ? 1 : 0); j += kGraphChunk) {
            MJ_HIP_CHECK(hipGraphLaunch(c->gexec, c->stream));
            if (energies && !fused) {   // staging[l] = E after step j+l-1
                if (j == 0)
                    MJ_HIP_CHECK(hipMemcpyAsync(c->evals, c->estage + 1, sizeof(double) * (kGraphChunk - 1),
                                                hipMemcpyDeviceToDevice, c->stream));
                else
                    MJ_HIP_CHECK(hipMemcpyAsync(c->evals + (j - 1), c->estage, sizeof(double) * kGraphChunk,
                                                hipMemcpyDeviceToDevice, c->stream));
            } else if (energies) {
                // fused: staging[l] = E after step j+l-2 (l >= 1, launch l folds step l-1),
                // staging[chunk] = E after step j-2 (launch 0 folds the previous replay's last step)
                if (j == 0) {
                    MJ_HIP_CHECK(hipMemcpyAsync(c->evals, c->estage + 1, sizeof(double) * (kGraphChunk - 2),
                                                hipMemcpyDeviceToDevice, c->stream));
                } else {
                    MJ_HIP_CHECK(hipMemcpyAsync(c->evals + (j - 2), c->estage + kGraphChunk, sizeof(double),
                                                hipMemcpyDeviceToDevice, c->stream));
                    MJ_HIP_CHECK(hipMemcpyAsync(c->evals + (j - 1), c->estage, sizeof(double) * (kGraphChunk - 1),
                                                hipMemcpyDeviceToDevice, c->stream));
                }
                // the partials of the replay's last step wait in buffer 1 for the next launch
                fuse_fold = c->evals + (j + kGraphChunk - 2);
            }
        }
    }
    if (prof && prof->block && prof->b0) MJ_HIP_CHECK(hipEventRecord(prof->b0, c->stream));
    for (; j < nsteps; ++j) {
        KArgs a = base;
        a.step = step0 + (uint64_t)j;
        if (adapt) {
            a.dt = stage_dt(ts);
            a.sigma = sigma_of(a.dt);
            if (ad->ts_used) ad->ts_used[j] = ts;
        }
        if (kT_steps) a.sigma = sigma_kt(a.dt, kT_steps[j]);
        if (noisy && noise_host) {
            MJ_HIP_CHECK(hipMemcpyAsync(c->Fb, noise_host + (size_t)j * 3 * c->n, sizeof(double) * 3 * c->n,
                                        hipMemcpyHostToDevice, c->stream));
            MJ_TRY(mj_launch_aos2soa(c->Fb, c->Nb, c->n, c->np, c->stream));
            a.noise = c->Nb;
        }
        if (dr) {   // energies after each step with that step's end-of-step field
            last_step = j == nsteps - 1;
            MJ_TRY(enqueue_step(a, j, nullptr));
            if (energies) {
                KArgs e;
                base_args(c, e);
                apply_drive(c, e, dr->nmaps, dr->w_end ? dr->w_end + (size_t)j * wlen : dr->w + ((size_t)j * nl + nl - 1) * wlen);
                e.src = c->A;
                MJ_TRY(prep_field(c, c->A, e));
                e.energy = 1;
                e.epart = c->epart;
                MJ_TRY(launch(c, MODE_FIELD, false, e));
                MJ_TRY(mj_launch_reduce(c->epart, e.nblocks, c->evals + j, escale, c->stream));
            }
            continue;
        }
        // energy of the state entering this step = E after step j-1
        last_step = j == nsteps - 1;
        MJ_TRY(enqueue_step(a, j, (energies && j > 0) ? c->evals + (j - 1) : nullptr));
    }
    c->field_valid = lf && nsteps > 0;
    // fused steps: the last step's partials (the state entering it)
    if (fuse_fold)
        MJ_TRY(mj_launch_fold(c->epart + (size_t)((nsteps - 1) & 1) * c->epart_cap, base.nblocks, fuse_fold, escale,
                              base.bx * base.by * base.bz, c->stream));
    if (prof && prof->block && prof->b1) MJ_HIP_CHECK(hipEventRecord(prof->b1, c->stream));
    if (adapt) *ad->ts_io = ts;
    if (energies && dr) {
        MJ_HIP_CHECK(hipMemcpyAsync(energies, c->evals, sizeof(double) * nsteps, hipMemcpyDeviceToHost, c->stream));
    } else if (energies) {
        MJ_TRY(energy_async(c, c->evals + (nsteps - 1), escale));
        MJ_HIP_CHECK(hipMemcpyAsync(energies, c->evals, sizeof(double) * nsteps, hipMemcpyDeviceToHost, c->stream));
    }
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

extern "C" int mj_steps(mj_ctx *c, int nsteps, int integrator, double timestep, double damping, double kT,
                        uint64_t seed, uint64_t step0, double *energies)
{
    if (!c) MJ_FAIL(MJ_EINVAL, "mj_steps: null context");
    return run_steps(c, nsteps, integrator, timestep, damping, kT, seed, step0, nullptr, energies, nullptr);
}

extern "C" int mj_steps_temperature(mj_ctx *c, int nsteps, int integrator, double timestep, double damping,
                                    const double *kT_steps, uint64_t seed, uint64_t step0, double *energies)
{
    if (!c || (!kT_steps && nsteps > 0)) MJ_FAIL(MJ_EINVAL, "mj_steps_temperature: null argument");
    return run_steps(c, nsteps, integrator, timestep, damping, 0.0, seed, step0, nullptr, energies, nullptr, nullptr,
                     nullptr, kT_steps);
}

extern "C" int mj_steps_adaptive(mj_ctx *c, int nsteps, int integrator, double *timestep, double damping,
                                 double kT, uint64_t seed, uint64_t step0, double discretization, double *energies,
                                 double *timesteps)
{
    if (!c || !timestep) MJ_FAIL(MJ_EINVAL, "mj_steps_adaptive: null argument");
    if (!(discretization > 0.0)) MJ_FAIL(MJ_EINVAL, "mj_steps_adaptive: discretization must be > 0");
    Adapt ad;
    ad.disc = discretization;
    ad.ts_io = timestep;
    ad.ts_used = timesteps;
    return run_steps(c, nsteps, integrator, *timestep, damping, kT, seed, step0, nullptr, energies, nullptr, &ad);
}

// B order parameter maps of the time-dependent field (mj_steps_driven)
extern "C" int mj_set_bfield_maps(mj_ctx *c, int nmaps, const double *maps, int64_t n)
{
    if (!c || nmaps < 0 || nmaps > 4 || (nmaps > 0 && !maps)) MJ_FAIL(MJ_EINVAL, "mj_set_bfield_maps: 0..4 maps");
    if (n != (int64_t)nmaps * 3 * c->n) MJ_FAIL(MJ_EINVAL, "mj_set_bfield_maps: size must be nmaps*3*Ncell");
    MJ_TRY(set_dev(c));
    if (c->d_maps) hipFree(c->d_maps);
    c->d_maps = nullptr;
    c->nmaps = nmaps;
    if (nmaps == 0) return MJ_OK;
    MJ_HIP_CHECK(hipMalloc(&c->d_maps, sizeof(double) * 3 * (size_t)c->np * nmaps));
    for (int k = 0; k < nmaps; ++k) {
        MJ_HIP_CHECK(hipMemcpyAsync(c->Nb, maps + (size_t)k * 3 * c->n, sizeof(double) * 3 * c->n,
                                    hipMemcpyHostToDevice, c->stream));
        MJ_TRY(mj_launch_aos2soa(c->Nb, c->d_maps + (size_t)k * 3 * c->np, c->n, c->np, c->stream));
    }
    MJ_HIP_CHECK(hipStreamSynchronize(c->stream));
    return MJ_OK;
}

// steps with a time-dependent B order parameter (excitations, update_exc before each
// stage, spindynamics.f90:285-292): weights per stage, see Drive
extern "C" int mj_steps_driven(mj_ctx *c, int nsteps, int integrator, double timestep, double damping, double kT,
                               uint64_t seed, uint64_t step0, const double *weights, const double *weights_end,
                               double *energies)
{
    if (!c || !weights) MJ_FAIL(MJ_EINVAL, "mj_steps_driven: null argument");
    if (!c->has_z) MJ_FAIL(MJ_ESTATE, "mj_steps_driven: no Zeeman term registered");
    Drive dr;
    dr.nmaps = c->nmaps;
    dr.w = weights;
    dr.w_end = weights_end;
    return run_steps(c, nsteps, integrator, timestep, damping, kT, seed, step0, nullptr, energies, nullptr, nullptr,
                     &dr);
}

extern "C" int mj_steps_noise(mj_ctx *c, int nsteps, int integrator, double timestep, double damping, double kT,
                              const double *noise, double *energies)
{
    if (!c || !noise) MJ_FAIL(MJ_EINVAL, "mj_steps_noise: null argument");
    return run_steps(c, nsteps, integrator, timestep, damping, kT, 0, 0, noise, energies, nullptr);
}

extern "C" int mj_profile_steps(mj_ctx *c, int nsteps, int integrator, double timestep, double damping,
                                double kT, uint64_t seed, int energies, double *total_ms, double *kind_ms,
                                int *kind_count)
{
    if (!c || !total_ms || !kind_ms || !kind_count) MJ_FAIL(MJ_EINVAL, "mj_profile_steps: null argument");
    MJ_TRY(set_dev(c));
    MJ_TRY(mj_finalize(c));
    Prof p;
    // one kernel per step: the fused kernel of integration 2, SIB or Depondt
    // (as run_steps picks it) or Euler on the stencil path
    const char *fenv = getenv("MJ_FUSE");
    const bool fusable = integrator == MJ_HEUN || integrator == MJ_SIB || integrator == MJ_DEPONDT;
    p.block = !c->mm_generic && !c->dip &&
              (integrator == MJ_EULER ||
               (fusable && c->use_jit && c->bmode != 3 && mj_jit_fused_ok(c) && !(fenv && fenv[0] == '0')));
    hipEvent_t t0, t1;
    MJ_HIP_CHECK(hipEventCreate(&t0));
    MJ_HIP_CHECK(hipEventCreate(&t1));
    if (p.block) {
        MJ_HIP_CHECK(hipEventCreate(&p.b0));
        MJ_HIP_CHECK(hipEventCreate(&p.b1));
    }
    std::vector<double> eh(energies ? (size_t)std::max(nsteps, 1) : 0);
    MJ_HIP_CHECK(hipEventRecord(t0, c->stream));
    int rc = run_steps(c, nsteps, integrator, timestep, damping, kT, seed, 0, nullptr, energies ? eh.data() : nullptr,
                       &p);
    MJ_HIP_CHECK(hipEventRecord(t1, c->stream));
    MJ_HIP_CHECK(hipEventSynchronize(t1));
    float ms = 0.f;
    MJ_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
    *total_ms = ms;
    for (int k = 0; k < 3; ++k) {
        kind_ms[k] = 0.0;
        kind_count[k] = 0;
    }
    if (p.block) {
        float bm = 0.f;
        if (rc == MJ_OK && p.block_count > 0) MJ_HIP_CHECK(hipEventElapsedTime(&bm, p.b0, p.b1));
        kind_ms[p.block_kind] = bm;
        kind_count[p.block_kind] = p.block_count;
        hipEventDestroy(p.b0);
        hipEventDestroy(p.b1);
    }
    for (size_t i = 0; i < p.kind.size(); ++i) {
        float e = 0.f;
        hipEventElapsedTime(&e, p.ev[2 * i], p.ev[2 * i + 1]);
        kind_ms[p.kind[i]] += e;
        kind_count[p.kind[i]] += 1;
    }
    for (auto e : p.ev) hipEventDestroy(e);
    hipEventDestroy(t0);
    hipEventDestroy(t1);
    return rc;
}
