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

#include "mj_kargs.h"

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
    // each registered M-M term on its own (energy per term); mm_terms_generic
    // when a term is not translation invariant
    struct MMTerm {
        std::vector<int> ofs3;
        std::vector<double> blk9;
    };
    std::vector<MMTerm> mm_terms;
    int mm_terms_generic = 0;
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
    std::vector<int> h_ofs3;           // host copy of the classified stencil: offsets (3 each) and
    std::vector<double> h_cf;          // coefficients in the order iso | dm | gen
    int rx = 0, ry = 0, rz = 0;
    int bx = 0, by = 0, bz = 0;        // 0: auto
    int dirty = 1;
    double zb[3] = {0, 0, 0};
    double *d_bmap = nullptr;
    double *d_maps = nullptr;          // field maps of the time-dependent B (SoA, 3 np each)
    int nmaps = 0;
    // generic CSR (H + H^T), device
    int *d_rp = nullptr, *d_ci = nullptr;
    double *d_v = nullptr;

    DipolarFFT *dip = nullptr;
    void *jit = nullptr;               // per-Hamiltonian compiled stage kernels (mj_jit.hip)
    int use_jit = 0;
    void *adj = nullptr;               // differentiable SIB trajectory (mj_adjoint.hip)
    void *mc = nullptr;                // Metropolis Monte Carlo state (mj_mc.hip)
    // graph of a chunk of steps (small lattices: one replay instead of 2 launches per step)
    hipGraphExec_t gexec = nullptr;
    std::vector<double> gkey;          // the parameters it was captured for
    uint64_t *d_step = nullptr;        // device step counter read by the replayed kernels
    double *estage = nullptr;          // energies written by one replay
    int rk_order = 4;                  // integration 4: N_order (the reference's default)
    // per-term device copies of the registered M-M terms (offsets, blocks), made
    // on first use by mj_term_mult / mj_energy_terms and kept; staging vectors
    std::vector<int *> term_ofs;
    std::vector<double *> term_blk;
    double *tm_in = nullptr, *tm_out = nullptr;
    double *Drk[6] = {};               // integration 4: D of the stages before the last (SoA)
    // the effective field of the last stage of the last step (mj_keep_last_field)
    int keep_field = 0, field_valid = 0;
    double *Lf = nullptr;
};

int mj_finalize(mj_ctx *c);   // build device-side Hamiltonian if dirty

// per-Hamiltonian stage kernels (mj_jit.hip)
int mj_jit_prepare(mj_ctx *c, const int *ofs_lds, const double *cf, int bx, int by, int bz, int nblocks, int ntx,
                   int nty);
int mj_jit_launch(mj_ctx *c, int mode, bool noise, const KArgs &a, hipStream_t s);
int mj_jit_resolve(mj_ctx *c, int mode, bool noise);
int mj_jit_fused_ok(const mj_ctx *c);   // the fused integration-2 kernel was generated
void mj_jit_destroy(mj_ctx *c);

void mj_adj_destroy(mj_ctx *c);   // mj_adjoint.hip
void mj_mc_destroy(mj_ctx *c);    // mj_mc.hip
int mj_energy_into(mj_ctx *c, double *dst_dev);   // total energy of the state into device memory (mj_api.hip)

// dipolar (mj_fft.hip)
int dip_create(mj_ctx *c, const int N_rep[3], const int periodic[3], const double *K_n);
int dip_destroy(mj_ctx *c);
int dip_clone(mj_ctx *dst, const mj_ctx *src);   // dst->dip with src's operator, buffers and plans of its own
// the grid dip_apply writes H_fft into, with its row, plane and component strides
void dip_geometry(const mj_ctx *c, const double **grid, int *dnx, int *dny, int64_t *hstride);
// H_fft of SoA state `m` into the padded real grid owned by the dipolar object;
// returns the grid pointer and geometry through KArgs fields
int dip_apply(mj_ctx *c, const double *m_soa, KArgs *ka);
