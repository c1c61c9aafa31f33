// Kernel-argument block and stage modes shared by the ahead-of-time kernels
// (mj_kernels.hip) and the per-Hamiltonian kernels compiled at set-up
// (mj_jit.hip); plain C++ so both compilers see the same layout.
#pragma once
#include <stdint.h>

// kernel modes
enum {
    MODE_FIELD = 0,   // write field (SoA) and/or energy partials
    MODE_EULER = 1,   // integration 1: one stage
    MODE_HEUN1 = 2,   // integration 2, stage 1
    MODE_HEUN2 = 3,   // integration 2, stage 2
    MODE_SIB1 = 4,
    MODE_SIB2 = 5,
    MODE_DEP1 = 6,
    MODE_DEP2 = 7,
    MODE_HEUNF = 8,   // integration 2, both stages in one kernel (compile-time geometry only)
    MODE_RKS = 9,     // integration 4 (Runge-Kutta order N): one stage
    MODE_SIBF = 10,   // SIB, both stages in one kernel (compile-time geometry only)
    MODE_DEPF = 11    // Depondt, both stages in one kernel (compile-time geometry only)
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
    double *fld_out;      // B_eff (SoA) of the stage: MODE_FIELD, or a step's last stage when asked

    int bmode;            // 0 none, 1 uniform zb, 2 per-site map, 3 weighted sum of maps
    double zb[3];         // Z^T b (uniform)
    int nmaps;            // bmode 3: b_i = sum_k wmap[3k..3k+2] (x) maps[k] (SoA, 3 np each)
    double wmap[12];
    const double *maps;
    double z[9];          // Zeeman block (row-major, Z[a][c] = H_BM(3i+a,3i+c))
    const double *bmap;   // per-site b (SoA)

    const double *hdip;   // dipolar H_fft on the padded grid, component stride hstride
    int dnx, dny;         // padded grid dims (x, y)
    int64_t hstride;

    double alpha, den, rden, dt, sigma;   // den = 1 + alpha^2, rden = RN(1/den)
    uint64_t seed, step;
    const uint64_t *stepp;  // or null: the thermal-field step is *stepp + step (graph-replayed steps)
    const double *noise;  // injected thermal field (SoA) or null

    int energy;           // 1: write per-block energy partials of src
    double *epart;
    const double *ered_in; // partials to reduce in block 0 (or null)
    int n_ered;
    double *e_out;
    double e_scale;

    int vbid, vbase;      // batched kernels: tile within the system, first energy slot of the system

    double *bmax_part;    // or null: per-block max over sites of |B_eff| (adaptive timestep)

    // MODE_RKS: D of this stage -> dout (or null); Dint = sum_k wrk[k] D_k over the
    // nprev earlier stages (dprev, SoA) and this one (wrk[nprev])
    int nprev;
    const double *dprev[7];
    double wrk[8];
    double *dout;
};

// per-term energy pass (mj_energy_terms): mode 0 an M-M stencil term, 1 the
// Zeeman term, 2 the FFT term on its padded grid; modes 3 and 4 the term
// products of mj_term_mult
struct TermArgs {
    int mode;
    const double *m;      // state (SoA)
    int64_t np;
    int dim[3], per[3];
    int nofs;
    const int *ofs3;      // mode 0: offsets (3 each) and blocks H(i, i+d) (9 each, row-major)
    const double *blk9;
    double z[9];          // mode 1: Zeeman block, b uniform (bmode 1) or per site (bmode 2, SoA)
    int bmode;
    double b[3];
    const double *bmap;
    const double *hd;     // mode 2: H_fft grid
    int dnx, dny;
    int64_t hstride;
    double *part;         // per-block partial sums
    // mj_term_mult (mode 3 an M-M term, 4 the Zeeman term): out = alpha H v + beta out
    // (side 0) or alpha H^T v + beta out (side 1), AoS vectors
    const double *vin;
    double *vout;
    double alpha, beta;
    int side;
};
