// Internal declarations shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/matjes_hip.h"

void mj_set_error(const std::string &msg);

#define MJ_HIP_CHECK(call)                                                                       \
    do {                                                                                         \
        hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            mj_set_error(std::string(#call) + ": " + hipGetErrorString(e_));                     \
            return MJ_EHIP;                                                                      \
        }                                                                                        \
    } while (0)

// kernel modes
enum {
    MODE_FIELD = 0,   // write field (SoA) and/or energy partials
    MODE_EULER = 1,   // integration 1: one stage
    MODE_HEUN1 = 2,   // integration 2, stage 1
    MODE_HEUN2 = 3,   // integration 2, stage 2
    MODE_SIB1 = 4,
    MODE_SIB2 = 5,
    MODE_DEP1 = 6,
    MODE_DEP2 = 7
};

// Everything a stage kernel needs, passed by value.
struct KArgs {
    int nx, ny, nz, px, py, pz;
    int bx, by, bz, rx, ry, rz, sx, sy, sz;
    int ntx, nty, ntz, nblocks;
    int64_t np;  // SoA component stride

    int n_iso, n_dm, n_gen;
    const int *ofs;     // LDS-linear offsets per stencil entry
    const double *cf;   // iso: s | dm: s,wx,wy,wz | gen: 9 (row-major)

    int use_stencil;      // 1: field from LDS stencil, 0: MM field read from fld_in
    const double *fld_in; // MM accumulation sum_j Hs_ij m_j (SoA), generic path

    const double *src;    // state the field is evaluated on (SoA)
    const double *own;    // second state read at the own site only (SoA) or null
    double *dst;          // output state (SoA)
    double *aux;          // stage-1 carried vector (SoA)
    double *fld_out;      // MODE_FIELD: B_eff (SoA)

    int bmode;            // 0 none, 1 uniform zb, 2 per-site map
    double zb[3];         // Z^T b (uniform)
    double z[9];          // Zeeman block (row-major, Z[a][c] = H_BM(3i+a,3i+c))
    const double *bmap;   // per-site b (SoA)

    const double *hdip;   // dipolar H_fft on the padded grid, component stride hstride
    int dnx, dny;         // padded grid dims (x, y)
    int64_t hstride;

    double alpha, den, dt, sigma;
    uint64_t seed, step;
    const double *noise;  // injected thermal field (SoA) or null

    int energy;           // 1: write per-block energy partials of src
    double *epart;
    const double *ered_in; // partials to reduce in block 0 (or null)
    int n_ered;
    double *e_out;
    double e_scale;
};

struct DipolarFFT;  // mj_fft.hip

struct mj_ctx {
    int device = 0;
    int dim[3] = {1, 1, 1};
    int periodic[3] = {0, 0, 0};
    int dim_mode = 3;
    int64_t n = 0;      // sites
    int64_t np = 0;     // padded SoA stride
    hipStream_t stream = nullptr;

    // state buffers (SoA, 3*np doubles each)
    double *A = nullptr, *Bb = nullptr, *Db = nullptr, *Fb = nullptr, *Nb = nullptr;
    double *epart = nullptr;   // 2 x max blocks
    double *evals = nullptr;   // energies device array
    int64_t evals_cap = 0;
    int epart_cap = 0;
    int state_set = 0;

    // Hamiltonian (host accumulation)
    std::vector<int> st_ofs;           // 3 per entry
    std::vector<double> st_blk;        // 9 per entry (H_MM(i,i+d) blocks, summed)
    int has_mm = 0;
    int mm_generic = 0;                // 1: generic CSR path
    std::vector<int64_t> g_rows;       // generic: 0-based COO of H_MM
    std::vector<int> g_cols;
    std::vector<double> g_vals;
    double zblk[9] = {0};
    int has_z = 0;
    std::vector<double> bvec;          // B order parameter (AoS 3N) or 3
    int bmode = 0;

    // device stencil
    int *d_ofs = nullptr;
    double *d_cf = nullptr;
    int n_iso = 0, n_dm = 0, n_gen = 0;
    int rx = 0, ry = 0, rz = 0;
    int bx = 0, by = 0, bz = 0;        // 0: auto
    int dirty = 1;
    double zb[3] = {0, 0, 0};
    double *d_bmap = nullptr;
    // generic CSR (H + H^T), device
    int *d_rp = nullptr, *d_ci = nullptr;
    double *d_v = nullptr;

    DipolarFFT *dip = nullptr;
};

int mj_finalize(mj_ctx *c);   // build device-side Hamiltonian if dirty

// dipolar (mj_fft.hip)
int dip_create(mj_ctx *c, const int N_rep[3], const int periodic[3], const double *K_n);
int dip_destroy(mj_ctx *c);
// H_fft of SoA state `m` into the padded real grid owned by the dipolar object;
// returns the grid pointer and geometry through KArgs fields
int dip_apply(mj_ctx *c, const double *m_soa, KArgs *ka);
