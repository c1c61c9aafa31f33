/*
 * mjoracle.c -- CPU restatement of the reference's LLG hot path (test
 * infrastructure: the checker for the HIP path and the timed CPU baseline;
 * never linked into the product).
 *
 * Arithmetic follows the reference's operation order and is compiled with
 * -ffp-contract=off (the reference's x86-64 Fortran has no fused multiply-add):
 *   field      B = -((H.m + H^T.m) + Z^T.b)
 *              src/Hamiltonian_type/hamiltonian_collection.f90:433-454,
 *              src/Hamiltonian_derivative/deriv_rank2.f90:13-35,
 *              src/Hamiltonian_type/ham_dense.f90:30-87 (column-ordered sums)
 *   energy     E = m.(H_MM m) + b.(Z m)          hamiltonian_collection.f90:365-400, ham_base.f90:181-207
 *   LLG        D = -m^x(B + a m^xB)/(1+a^2)      src/solvers/propagator.f90:13-40
 *   euler      t = (m + Dint dt/hbar) + DT dt/hbar; m' = t |m|/|t|   src/solvers/solver.f90:157-178
 *   Heun/Euler stage loop                        src/dynamics/spindynamics.f90:260-323,
 *              src/solvers/solver_order.f90:385-429,436-453
 *   Runge-Kutta order N (integration 4)          src/solvers/solver_commun.f90:112-121,
 *              src/solvers/solver_order.f90:14-58,65-127
 *   langevin   DT = -m^x(BT + a m^xBT)/(1+a^2)   src/dynamics/BTeff_sd.f90:18-45
 *   truncate   |x| < eps -> 0                     src/vector_utils/precision.f90:39-51
 * Integrators without a reference implementation are restated from their
 * papers: SIB (Mentink et al., J. Phys.: Condens. Matter 22, 176001 (2010))
 * and Depondt (Depondt & Mertens, J. Phys.: Condens. Matter 21, 336005 (2009)).
 * The thermal generator is Philox4x32-10 (Salmon et al., SC'11) with
 * Box-Muller normals; the reference's own draw is gfortran random_number and
 * is replaced by an injected field in parity runs.
 *
 * Layout: m, B, b are AoS (3,N) doubles, x fastest, as the reference's
 * modes_v(3,Ncell).  CSR arrays are 0-based with ascending columns per row.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_HBAR 6.58211928e-1 /* eV fs, src/fixed-parameters/constant.f90:25 */
#define OR_PI 3.14159265358979323846264338327950288 /* src/fixed-parameters/constant.f90:32 */

typedef struct {
    int n; /* sites */
    const int *rp, *ci;
    const double *v; /* H_MM CSR (3N rows) */
    const int *rpt, *cit;
    const double *vt; /* H_MM^T CSR */
    const double *z;  /* Zeeman 3x3 block, z[a*3+b] = H_BM(3i+a,3i+b); NULL if absent */
    const double *b;  /* B order parameter (3N) */
    const double *hdip; /* optional extra field added as B -= 2*hdip (3N); NULL if absent */
} or_ham;

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                   */
/* ------------------------------------------------------------------ */
void or_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static const double OR_TWO_M32 = 2.3283064365386962890625e-10; /* 2^-32 */

/* Three standard normals for (site, step) under key seed: one Philox block,
 * counter (site, step lo, step hi, 0); Box-Muller on 32-bit uniforms, words
 * (0, 1) -> n0, n1 and words (2, 3) -> n2 (matjes_amd/csrc/mj_device.h). */
void or_normals3(uint64_t seed, uint64_t step, uint32_t site, double out[3])
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {site, (uint32_t)step, (uint32_t)(step >> 32), 0u};
    uint32_t x[4];
    or_philox4x32_10(ctr, key, x);
    double u1 = ((double)x[0] + 1.0) * OR_TWO_M32; /* (0,1] */
    double u2 = (double)x[1] * OR_TWO_M32;         /* [0,1) */
    double u3 = ((double)x[2] + 1.0) * OR_TWO_M32;
    double u4 = (double)x[3] * OR_TWO_M32;
    double r1 = sqrt(-2.0 * log(u1)), r2 = sqrt(-2.0 * log(u3));
    double t1 = 6.283185307179586476925286766559 * u2;
    double t2 = 6.283185307179586476925286766559 * u4;
    out[0] = r1 * cos(t1);
    out[1] = r1 * sin(t1);
    out[2] = r2 * cos(t2);
}

/* sigma * normals for all sites of one step, AoS (3,N) */
void or_thermal_field(uint64_t seed, uint64_t step, int n, double sigma, double *bt)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double g[3];
        or_normals3(seed, step, (uint32_t)i, g);
        bt[3 * i] = sigma * g[0];
        bt[3 * i + 1] = sigma * g[1];
        bt[3 * i + 2] = sigma * g[2];
    }
}

/* ------------------------------------------------------------------ */
/* field and energy                                                   */
/* ------------------------------------------------------------------ */
static void spmv(int n3, const int *rp, const int *ci, const double *v, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n3; ++r) {
        double acc = 0.0;
        for (int k = rp[r]; k < rp[r + 1]; ++k) acc = acc + v[k] * x[ci[k]];
        y[r] = acc;
    }
}

/* B_eff = -(H m + H^T m + Z^T b) - 2 hdip    (hamiltonian_collection.f90:442-453) */
void or_field(const or_ham *h, const double *m, double *B, double *work /* 3N */)
{
    int n3 = 3 * h->n;
    spmv(n3, h->rp, h->ci, h->v, m, B);            /* B = H m + 0           */
    spmv(n3, h->rpt, h->cit, h->vt, m, work);      /* H^T m                 */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < h->n; ++i) {
        for (int a = 0; a < 3; ++a) {
            double v = work[3 * i + a] + B[3 * i + a];
            if (h->z) {
                /* matmul(b, Z) restricted to the block: sum_c b_c Z(c,a), c ascending */
                double t = 0.0;
                for (int c = 0; c < 3; ++c) t = t + h->b[3 * i + c] * h->z[c * 3 + a];
                v = t + v;
            }
            v = -v;
            if (h->hdip) v = v - 2.0 * h->hdip[3 * i + a];
            B[3 * i + a] = v;
        }
    }
}

/* Total energy (serial sums in array order, as dot_product/sum). */
double or_energy(const or_ham *h, const double *m, double *work /* 3N */)
{
    int n3 = 3 * h->n;
    spmv(n3, h->rp, h->ci, h->v, m, work);
    double emm = 0.0;
    for (int r = 0; r < n3; ++r) emm = emm + m[r] * work[r];
    double ez = 0.0;
    if (h->z) {
        for (int i = 0; i < h->n; ++i)
            for (int a = 0; a < 3; ++a) {
                double t = 0.0;
                for (int c = 0; c < 3; ++c) t = t + h->z[a * 3 + c] * m[3 * i + c];
                ez = ez + h->b[3 * i + a] * t;
            }
    }
    double e = (0.0 + emm) + ez;
    if (h->hdip) {
        double ed = 0.0;
        for (int r = 0; r < n3; ++r) ed = ed + h->hdip[r] * m[r];
        e = e + ed;
    }
    return e;
}

/* ------------------------------------------------------------------ */
/* per-site kernels                                                    */
/* ------------------------------------------------------------------ */
static inline double nrm3(const double *x) { return sqrt((x[0] * x[0] + x[1] * x[1]) + x[2] * x[2]); }

static inline void cross3(const double *a, const double *b, double *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* propagator.f90:13-40 (also BTeff_sd.f90:18-45 with B -> BT) */
void or_llg(int n, double damping, const double *m, const double *B, double *D)
{
    double den = 1.0 + damping * damping;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double mn[3] = {m[3 * i], m[3 * i + 1], m[3 * i + 2]};
        double nr = nrm3(mn);
        if (nr > 1.0e-8) { mn[0] /= nr; mn[1] /= nr; mn[2] /= nr; }
        double c[3], l[3], o[3];
        cross3(mn, B + 3 * i, c);
        for (int a = 0; a < 3; ++a) l[a] = B[3 * i + a] + damping * c[a];
        cross3(mn, l, o);
        for (int a = 0; a < 3; ++a) D[3 * i + a] = -o[a] / den;
    }
}

/* solver.f90:157-178 */
void or_euler(int n, const double *m, const double *dint, const double *dt_field, double dt, double *out)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
        double t[3];
        double mn = nrm3(m + 3 * i);
        for (int a = 0; a < 3; ++a) {
            double v = m[3 * i + a] + dint[3 * i + a] * dt / OR_HBAR;
            if (dt_field) v = v + dt_field[3 * i + a] * dt / OR_HBAR;
            else v = v + 0.0 * dt / OR_HBAR;
            t[a] = v;
        }
        double tn = nrm3(t);
        for (int a = 0; a < 3; ++a) out[3 * i + a] = (tn > 1.0e-8) ? t[a] * mn / tn : m[3 * i + a];
    }
}

void or_truncate(int n3, double *m, double eps)
{
    for (int r = 0; r < n3; ++r)
        if (fabs(m[r]) < eps) m[r] = 0.0;
}

/* implicit-midpoint solve x + a x x... : (I + [a]x) x = y, y = m - a x m */
static inline void cayley(const double *m, const double *a, double *out)
{
    double axm[3], y[3], axy[3];
    cross3(a, m, axm);
    for (int k = 0; k < 3; ++k) y[k] = m[k] - axm[k];
    cross3(a, y, axy);
    double ay = (a[0] * y[0] + a[1] * y[1]) + a[2] * y[2];
    double aa = (a[0] * a[0] + a[1] * a[1]) + a[2] * a[2];
    double inv = 1.0 / (1.0 + aa);
    for (int k = 0; k < 3; ++k) out[k] = ((y[k] - axy[k]) + ay * a[k]) * inv;
}

/* a = (dt/2) * Omega,  Omega = -(B + alpha p x B)/((1+alpha^2) hbar) */
static inline void sib_vec(const double *p, const double *B, double alpha, double dt, double *a)
{
    double c[3];
    cross3(p, B, c);
    double f = -0.5 * dt / ((1.0 + alpha * alpha) * OR_HBAR);
    for (int k = 0; k < 3; ++k) a[k] = f * (B[k] + alpha * c[k]);
}

/* Btilde = (B + alpha m^ x B)/(1+alpha^2) */
static inline void dep_field(const double *m, const double *B, double alpha, double *bt)
{
    double mn[3] = {m[0], m[1], m[2]}, c[3];
    double nr = nrm3(mn);
    if (nr > 1.0e-8) { mn[0] /= nr; mn[1] /= nr; mn[2] /= nr; }
    cross3(mn, B, c);
    double den = 1.0 + alpha * alpha;
    for (int k = 0; k < 3; ++k) bt[k] = (B[k] + alpha * c[k]) / den;
}

/* rotate m about b by angle |b| dt / hbar (dm/dt = (1/hbar) b x m) */
static inline void dep_rotate(const double *m, const double *b, double dt, double *out)
{
    double bn = nrm3(b);
    if (!(bn > 0.0)) { out[0] = m[0]; out[1] = m[1]; out[2] = m[2]; return; }
    double k[3] = {b[0] / bn, b[1] / bn, b[2] / bn};
    double th = bn * dt / OR_HBAR;
    double ct = cos(th), st = sin(th);
    double kxm[3];
    cross3(k, m, kxm);
    double km = (k[0] * m[0] + k[1] * m[1]) + k[2] * m[2];
    for (int q = 0; q < 3; ++q) out[q] = (m[q] * ct + kxm[q] * st) + k[q] * (km * (1.0 - ct));
}

/* ------------------------------------------------------------------ */
/* the step loop                                                       */
/* ------------------------------------------------------------------ */
enum { OR_EULER = 1, OR_HEUN = 2, OR_RKN = 4, OR_SIB = 5, OR_DEPONDT = 6 };

typedef void (*or_field_cb)(void *user, const double *m, double *hdip_out);

typedef struct {
    int integrator;
    double timestep, damping, kT;
    uint64_t seed, step0;      /* Philox noise when noise == NULL and kT*damping > 1e-10 */
    const double *noise;       /* injected BT, (nsteps,3N), already sigma-scaled; or NULL */
    double trunc_eps;          /* 1e-50 in the reference */
    or_field_cb dip;           /* optional: fills hdip for a state (dipolar H_fft), NULL if none */
    void *dip_user;
    double discretization;     /* > 0: update_time after every step (update_time.f90:41-75) */
    double *timesteps;         /* optional (nsteps): timestep used by each step */
    const double *b_stage;     /* optional (nsteps, stages, 3N): B order parameter of each stage
                                  (update_exc before each stage, spindynamics.f90:285-292) */
    const double *b_end;       /* optional (nsteps, 3N): B of the energy after each step */
    int rk_order;              /* integration 4: N_order (1, 2, 3, 4, 6, 7; the reference's default 4) */
} or_dyn;

/* Butcher table of get_butcher_explicit (src/solvers/solver_order.f90:65-127),
 * 1-based as there: T[row][col], rows/cols 1..N+1; row 1 holds the factor of
 * the stage timestep (get_dt_mode), rows 2..N+1 the weights of the stages'
 * D (get_Dmode_int).  0 if the order has no table (the reference stops). */
static int rk_butcher(int N, double T[9][9])
{
    for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 9; ++c) T[r][c] = 0.0;
    switch (N) {
    case 1:
        T[1][2] = 1.0; T[2][2] = 1.0;
        return 1;
    case 2:
        T[1][2] = 1.0; T[2][2] = 1.0;
        T[1][3] = 1.0; T[2][3] = 0.5; T[3][3] = 0.5;
        return 1;
    case 3:
        T[1][2] = 0.5; T[2][2] = 0.5;
        T[1][3] = 1.0; T[2][3] = -1.0; T[3][3] = 2.0; T[4][3] = 0.0;
        T[1][4] = 6.0 / 6.0; T[2][4] = 1.0 / 6.0; T[3][4] = 2.0 / 6.0; T[4][4] = 1.0 / 6.0;
        return 1;
    case 4:
        T[1][2] = 0.5; T[2][2] = 0.5;
        T[1][3] = 0.5; T[2][3] = 0.0; T[3][3] = 0.5;
        T[1][4] = 1.0; T[2][4] = 0.0; T[3][4] = 0.0; T[4][4] = 1.0;
        T[1][5] = 6.0 / 6.0; T[2][5] = 1.0 / 6.0; T[3][5] = 2.0 / 6.0; T[4][5] = 2.0 / 6.0; T[5][5] = 1.0 / 6.0;
        return 1;
    case 6:
        T[1][2] = 0.25; T[2][2] = 0.25;
        T[1][3] = 12.0 / 32.0; T[2][3] = 3.0 / 32.0; T[3][3] = 9.0 / 32.0;
        T[1][4] = 2028.0 / 2197.0; T[2][4] = 1932.0 / 2197.0; T[3][4] = -7200.0 / 2197.0; T[4][4] = 7296.0 / 2197.0;
        T[1][5] = 1.0; T[2][5] = 439.0 / 216.0; T[3][5] = -8.0; T[4][5] = 3680.0 / 513.0; T[5][5] = -845.0 / 4104.0;
        T[1][6] = 0.5; T[2][6] = -8.0 / 27.0; T[3][6] = 2.0; T[4][6] = -3544.0 / 2565.0; T[5][6] = 1859.0 / 4104.0;
        T[6][6] = -11.0 / 40.0;
        T[1][7] = 1.0; T[2][7] = 25.0 / 216.0; T[3][7] = 0.0; T[4][7] = 1408.0 / 2565.0; T[5][7] = 2197.0 / 4104.0;
        T[6][7] = -1.0 / 5.0; T[7][7] = 0.0;
        return 1;
    case 7:
        T[1][2] = 1.0 / 5.0; T[2][2] = 1.0 / 5.0;
        T[1][3] = 12.0 / 40.0; T[2][3] = 3.0 / 40.0; T[3][3] = 9.0 / 40.0;
        T[1][4] = 4.0 / 5.0; T[2][4] = 44.0 / 45.0; T[3][4] = -56.0 / 15.0; T[4][4] = 32.0 / 9.0;
        T[1][5] = 8.0 / 9.0; T[2][5] = 19372.0 / 6561.0; T[3][5] = -25360.0 / 2187.0; T[4][5] = 64448.0 / 6561.0;
        T[5][5] = -212.0 / 729.0;
        T[1][6] = 1.0; T[2][6] = 9017.0 / 3168.0; T[3][6] = -355.0 / 33.0; T[4][6] = 46732.0 / 5247.0;
        T[5][6] = 49.0 / 176.0; T[6][6] = -5103.0 / 18656.0;
        T[1][7] = 1.0; T[2][7] = 35.0 / 384.0; T[3][7] = 0.0; T[4][7] = 500.0 / 1113.0; T[5][7] = 125.0 / 192.0;
        T[6][7] = -2187.0 / 6784.0; T[7][7] = 11.0 / 84.0;
        T[1][8] = 1.0; T[2][8] = 5179.0 / 57600.0; T[3][8] = 0.0; T[4][8] = 7571.0 / 16695.0; T[5][8] = 393.0 / 640.0;
        T[6][8] = -92097.0 / 339200.0; T[7][8] = 187.0 / 2100.0; T[8][8] = 1.0 / 40.0;
        return 1;
    default:
        return 0;
    }
}

/* max over sites of norm(B_i) = sqrt(sum(B_i**2)) (update_time.f90:58-68, vector.f90:353-357) */
static double max_norm(int n, const double *B)
{
    double mx = 0.0;
    for (int i = 0; i < n; ++i) {
        const double *b = B + 3 * i;
        double v = sqrt((b[0] * b[0] + b[1] * b[1]) + b[2] * b[2]);
        if (v > mx) mx = v;
    }
    return mx;
}

/* Advance m (3N) by nsteps. energies (nsteps) receives E/Ncell after each
 * step when non-NULL (spindynamics.f90:323). */
int or_steps(const or_ham *h_in, const or_dyn *d, double *m, int nsteps, double *energies)
{
    int n = h_in->n, n3 = 3 * n;
    or_ham h = *h_in;
    double *B = malloc(sizeof(double) * n3), *w = malloc(sizeof(double) * n3);
    double *m1 = malloc(sizeof(double) * n3), *m2 = malloc(sizeof(double) * n3);
    double *D1 = calloc((size_t)n3, sizeof(double)), *D2 = calloc((size_t)n3, sizeof(double));
    double *Di = malloc(sizeof(double) * n3), *DT = malloc(sizeof(double) * n3);
    double *bt = malloc(sizeof(double) * n3), *hd = d->dip ? malloc(sizeof(double) * n3) : NULL;
    double *lat = malloc(sizeof(double) * n3); /* state m1 carried between steps (lat_1) */
    if (!B || !w || !m1 || !m2 || !D1 || !D2 || !Di || !DT || !bt || !lat || (d->dip && !hd)) return -1;
    memcpy(lat, m, sizeof(double) * n3);
    h.hdip = hd;
    int noisy = d->kT * d->damping > 1.0e-10;
    /* integration 4: the reference selects no thermal field for it (solver_commun.f90:112-121) */
    double BT[9][9];
    double *Dk[8] = {0};
    if (d->integrator == OR_RKN) {
        if (noisy || !rk_butcher(d->rk_order, BT)) return -4;
        for (int k = 0; k < d->rk_order; ++k)
            if (!(Dk[k] = calloc((size_t)n3, sizeof(double)))) return -1;   /* Dmag, kept between steps */
    }

#define FIELD(state) do { if (d->dip) d->dip(d->dip_user, (state), hd); or_field(&h, (state), B, w); } while (0)
    const int NL = d->integrator == OR_EULER ? 1 : 2;
#define SETB(jj, l) do { if (d->b_stage) h.b = d->b_stage + ((size_t)(jj) * NL + (l)) * (size_t)n3; } while (0)

    double ts = d->timestep;                                          /* timestep_int */
    for (int j = 0; j < nsteps; ++j) {
        uint64_t step = d->step0 + (uint64_t)j;
        double bmax = 0.0;                                            /* max norm of the last stage's field */
        if (d->timesteps) d->timesteps[j] = ts;
        if (d->integrator == OR_EULER || d->integrator == OR_HEUN) {
            int nl = d->integrator == OR_EULER ? 1 : 2;
            double dt = ts / (double)nl;                              /* spindynamics.f90:266 */
            int have_t = 0;
            if (noisy) {
                if (d->noise) memcpy(bt, d->noise + (size_t)j * n3, sizeof(double) * n3);
                else or_thermal_field(d->seed, step, n, sqrt(d->damping * d->kT * OR_HBAR / dt), bt);
                or_llg(n, d->damping, lat, bt, DT);                   /* langevin_bath */
                have_t = 1;
            }
            /* stage 1: Butcher(1,2)=1 -> dt unchanged; Dint = 1*D1 + 0*D2 */
            SETB(j, 0);
            FIELD(lat);
            if (nl == 1) bmax = max_norm(n, B);
            or_llg(n, d->damping, lat, B, D1);
            if (nl == 1) {
                for (int r = 0; r < n3; ++r) Di[r] = (0.0 + 1.0 * D1[r]);
                or_euler(n, lat, Di, have_t ? DT : NULL, dt, m2);
            } else {
                for (int r = 0; r < n3; ++r) Di[r] = (0.0 + 1.0 * D1[r]) + 0.0 * D2[r];
                or_euler(n, lat, Di, NULL, dt, m2);
                memcpy(lat, m2, sizeof(double) * n3);
                /* stage 2: Butcher(1,3)=1; Dint = 0.5*D1 + 0.5*D2 */
                SETB(j, 1);
                FIELD(lat);
                bmax = max_norm(n, B);
                or_llg(n, d->damping, lat, B, D2);
                for (int r = 0; r < n3; ++r) Di[r] = (0.0 + 0.5 * D1[r]) + 0.5 * D2[r];
                or_euler(n, lat, Di, have_t ? DT : NULL, dt, m2);
            }
            memcpy(lat, m2, sizeof(double) * n3);
        } else if (d->integrator == OR_SIB || d->integrator == OR_DEPONDT) {
            double dt = ts;
            if (noisy) {
                if (d->noise) memcpy(bt, d->noise + (size_t)j * n3, sizeof(double) * n3);
                else or_thermal_field(d->seed, step, n, sqrt(2.0 * d->damping * d->kT * OR_HBAR / dt), bt);
            }
            SETB(j, 0);
            FIELD(lat);
            if (noisy) for (int r = 0; r < n3; ++r) B[r] = B[r] + bt[r];
            SETB(j, 1);
            if (d->integrator == OR_SIB) {
                for (int i = 0; i < n; ++i) {           /* predictor, midpoint p = (m + m~)/2 */
                    double a[3], mt[3];
                    sib_vec(lat + 3 * i, B + 3 * i, d->damping, dt, a);
                    cayley(lat + 3 * i, a, mt);
                    for (int k = 0; k < 3; ++k) m1[3 * i + k] = 0.5 * (lat[3 * i + k] + mt[k]);
                }
                FIELD(m1);
                bmax = max_norm(n, B);
                if (noisy) for (int r = 0; r < n3; ++r) B[r] = B[r] + bt[r];
                for (int i = 0; i < n; ++i) {           /* corrector */
                    double a[3];
                    sib_vec(m1 + 3 * i, B + 3 * i, d->damping, dt, a);
                    cayley(lat + 3 * i, a, m2 + 3 * i);
                }
            } else {
                for (int i = 0; i < n; ++i) {           /* predictor rotation */
                    dep_field(lat + 3 * i, B + 3 * i, d->damping, D1 + 3 * i);
                    dep_rotate(lat + 3 * i, D1 + 3 * i, dt, m1 + 3 * i);
                }
                FIELD(m1);
                bmax = max_norm(n, B);
                if (noisy) for (int r = 0; r < n3; ++r) B[r] = B[r] + bt[r];
                for (int i = 0; i < n; ++i) {           /* corrector with averaged field */
                    double b2[3], ba[3];
                    dep_field(m1 + 3 * i, B + 3 * i, d->damping, b2);
                    for (int k = 0; k < 3; ++k) ba[k] = 0.5 * (D1[3 * i + k] + b2[k]);
                    dep_rotate(lat + 3 * i, ba, dt, m2 + 3 * i);
                }
            }
            memcpy(lat, m2, sizeof(double) * n3);
        } else if (d->integrator == OR_RKN) {
            /* spindynamics.f90:260-318 with integration 4: every stage starts from
             * the previous stage's state, dt = Butcher(1, i+1) * dt cumulatively
             * from timestep / N (get_dt_mode), Dint = sum over all N columns of
             * Butcher(k+1, i+1) D_k (get_Dmode_int; D_k of later stages are last
             * step's, with weight 0) */
            const int N = d->rk_order;
            double dt = ts / (double)N;
            for (int i = 1; i <= N; ++i) {
                dt = BT[1][i + 1] * dt;
                FIELD(lat);
                if (i == N) bmax = max_norm(n, B);
                or_llg(n, d->damping, lat, B, Dk[i - 1]);
                for (int r = 0; r < n3; ++r) {
                    double v = 0.0;
                    for (int k = 2; k <= N + 1; ++k) v = v + BT[k][i + 1] * Dk[k - 2][r];
                    Di[r] = v;
                }
                or_euler(n, lat, Di, NULL, dt, m2);
                memcpy(lat, m2, sizeof(double) * n3);
            }
        } else {
            return -2;
        }
        memcpy(m, lat, sizeof(double) * n3);
        or_truncate(n3, m, d->trunc_eps);                             /* spindynamics.f90:319 */
        if (energies) {
            if (d->dip) d->dip(d->dip_user, m, hd);
            if (d->b_stage) h.b = d->b_end ? d->b_end + (size_t)j * n3 : d->b_stage + ((size_t)j * NL + NL - 1) * (size_t)n3;
            energies[j] = or_energy(&h, m, w) / (double)n;            /* spindynamics.f90:323 */
        }
        if (d->discretization > 0.0) {                                /* spindynamics.f90:438 */
            if (!(bmax > 1.0e-8)) return -3;                          /* 'error in update_time' */
            ts = 2.0 * OR_PI * OR_HBAR / bmax / d->discretization;
        }
    }
#undef FIELD
#undef SETB
    free(B); free(w); free(m1); free(m2); free(D1); free(D2); free(Di); free(DT); free(bt); free(lat);
    if (hd) free(hd);
    for (int k = 0; k < 8; ++k)
        if (Dk[k]) free(Dk[k]);
    return 0;
}

int or_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_num_threads(int t)
{
#ifdef _OPENMP
    omp_set_num_threads(t);
#else
    (void)t;
#endif
}

/* Mismatch count of the correctly rounded quotient a/b obtained from
 * r = RN(1/b): q0 = RN(a r), e = fma(-b, q0, a), q = fma(e, r, q0)
 * (Markstein 1990), against IEEE a/b, over n random (a, b) with exponents
 * in [-60, 60] and [-30, 30].  The device path divides this way. */
long or_div_rn_mismatches(long n, uint64_t seed)
{
    uint64_t s = seed | 1;
    long bad = 0;
    for (long i = 0; i < n; ++i) {
        double v[2];
        for (int k = 0; k < 2; ++k) {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            uint64_t u = s;
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            int ex = k == 0 ? (int)(s % 121) - 60 : (int)(s % 61) - 30;
            double x = ldexp(1.0 + (double)(u >> 12) * 0x1p-52, ex);
            v[k] = (k == 0 && (u & 1)) ? -x : x;
        }
        double a = v[0], b = v[1];
        double r = 1.0 / b, q0 = a * r, e = fma(-b, q0, a), q = fma(e, r, q0);
        if (q != a / b) ++bad;
    }
    return bad;
}

/* ------------------------------------------------------------------ */
/* Metropolis Monte Carlo                                              */
/* ------------------------------------------------------------------ */
/* energy_single (src/Hamiltonian_type/hamiltonian_collection.f90:402-431) with
 * the rank-2 single-site evaluation (src/Hamiltonian_type/eval_single.f90):
 * each term's row of the site times the state with the site set to s, times
 * mult_M_single (2 for pair terms, 1 for on-site and Zeeman terms):
 *   E = 2 s.(sum_{j != i} H_ij m_j) + s.(H_ii s) + b_i.(Z s) */
double or_energy_single(const or_ham *h, const double *m, int i, const double s[3])
{
    double pair = 0.0, onsite = 0.0;
    for (int a = 0; a < 3; ++a) {
        const int r = 3 * i + a;
        double pr = 0.0, os = 0.0;
        for (int k = h->rp[r]; k < h->rp[r + 1]; ++k) {
            const int c = h->ci[k];
            if (c / 3 == i)
                os = os + h->v[k] * s[c % 3];
            else
                pr = pr + h->v[k] * m[c];
        }
        pair = pair + s[a] * pr;
        onsite = onsite + s[a] * os;
    }
    double ez = 0.0;
    if (h->z) {
        for (int a = 0; a < 3; ++a) {
            double t = 0.0;
            for (int c = 0; c < 3; ++c) t = t + h->z[a * 3 + c] * s[c];
            ez = ez + h->b[3 * i + a] * t;
        }
    }
    return (2.0 * pair + onsite) + ez;
}

static int or_canon(int d, int n, int periodic)
{
    if (!periodic) return d;
    d %= n;
    if (d < 0) d += n;
    if (2 * d > n) d -= n;
    return d;
}

/* Colour classes of the Monte Carlo passes: class = (x + a y + b z) mod p,
 * the smallest p (then a, then b) with no stencil offset inside one class and
 * classes consistent across periodic wraps.  ofs3: the offsets d of the
 * blocks H(i, i+d) (either sign).  Returns 0 and (p, a, b), or -1. */
int or_mc_colouring(int nofs, const int *ofs3, const int dim[3], const int per[3], int out[3])
{
    int *o = (int *)malloc(sizeof(int) * 3 * (nofs > 0 ? nofs : 1));
    int n = 0, uses_y = 0, uses_z = 0;
    for (int k = 0; k < nofs; ++k) {
        int d[3];
        for (int q = 0; q < 3; ++q) d[q] = or_canon(ofs3[3 * k + q], dim[q], per[q]);
        if (d[0] == 0 && d[1] == 0 && d[2] == 0) continue;   /* on-site */
        for (int q = 0; q < 3; ++q) o[3 * n + q] = d[q];
        uses_y |= d[1] != 0;
        uses_z |= d[2] != 0;
        ++n;
    }
    int found = -1;
    for (int p = 1; p <= 64 && found < 0; ++p) {
        if (per[0] && dim[0] % p) continue;
        for (int a = 0; a < p && found < 0; ++a) {
            if (!uses_y && a) break;
            if (per[1] && ((long)a * dim[1]) % p) continue;
            for (int b = 0; b < p && found < 0; ++b) {
                if (!uses_z && b) break;
                if (per[2] && ((long)b * dim[2]) % p) continue;
                int ok = 1;
                for (int k = 0; k < n && ok; ++k) ok = ((long)o[3 * k] + (long)a * o[3 * k + 1] + (long)b * o[3 * k + 2]) % p != 0;
                if (ok) {
                    out[0] = p; out[1] = a; out[2] = b;
                    found = 0;
                }
            }
        }
    }
    free(o);
    return found;
}

typedef struct {
    double cone, cos_cone, acc, nb, rate;
} or_mc_state;

/* cone_update (src/Metropolis-MC/track_value.f90:149-160) applied n times at a
 * fixed rate, in closed form */
static void or_mc_adapt(or_mc_state *st, double n)
{
    const double pi = 3.14159265358979323846264338327950288;
    double cone = st->cone;
    if (st->rate > 0.5 && cone < pi) {
        double k = ceil((pi - cone) / 1.0e-4);
        if (k > n) k = n;
        cone = cone + k * 1.0e-4;
    } else if (st->rate < 0.5 && cone > 0.01) {
        double k = ceil((cone - 0.01) / 1.0e-4);
        if (k > n) k = n;
        cone = cone - k * 1.0e-4;
    }
    st->cone = cone;
    st->cos_cone = cos(cone);
}

/* sphereft (src/paratemp-MC/sampling.f90:9-62): uniform in the cap of
 * half-angle cone around s; u1 -> cos(theta), u2 -> phi */
static void or_sphereft(const double s[3], double cos_cone, double u1, double u2, double out[3])
{
    const double ct = u1 + cos_cone * (1.0 - u1);
    const double st = sqrt(1.0 - ct * ct);
    const double phi = 2.0 * acos(-1.0) * u2;
    const double cp = cos(phi), sp = sin(phi);
    const double cti = s[2];
    const double rr = s[0] * s[0] + s[1] * s[1];
    const double sti = sqrt(rr);
    double cpi = 1.0, spi = 0.0;
    if (!(rr < 0.00001)) {
        cpi = s[0] / sqrt(rr);
        spi = s[1] / sqrt(rr);
    }
    const double t = st * sp * cti + ct * sti;
    double n[3];
    n[1] = st * cp * cpi + t * spi;
    n[0] = -(st * cp * spi) + t * cpi;
    n[2] = -(st * sp * sti) + ct * cti;
    const double ss = sqrt((n[0] * n[0] + n[1] * n[1]) + n[2] * n[2]);
    for (int q = 0; q < 3; ++q) out[q] = n[q] / ss;
}

/* nsweeps sweeps of MCstep (src/Metropolis-MC/MCstep.f90:17-65) by colour
 * class: every site of class k proposes (sphereft or the Ising flip), the
 * energy change is energy_single(new) - energy_single(old), accept
 * (MCstep.f90:70-88) takes dE < 0 or u3 < exp(max(-dE/kT, -200)).  Random
 * numbers: Philox4x32-10, key seed, counter (site, sweep, stream), stream
 * 0x4D430000 + replica, u1, u2,
 * u3 = words 0..2 * 2^-32.  sampler 0 = sphere (cone adapted once per sweep
 * for its N proposals at the running rate), 1 = Ising.  m: AoS 3N, updated in
 * place. */
void or_mc_sweeps(const or_ham *h, double *m, const int dim[3], const int col[3], int nsweeps, int sampler, double kT,
                  uint64_t seed, uint64_t sweep0, or_mc_state *st, uint32_t stream)
{
    const int p = col[0], ca = col[1], cb = col[2];
    const int n = h->n;
    const double two_m32 = 2.3283064365386962890625e-10;
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    int *cls = (int *)malloc(sizeof(int) * (size_t)n);
    double *count = (double *)calloc((size_t)p, sizeof(double));
    for (int i = 0; i < n; ++i) {
        const int x = i % dim[0], y = (i / dim[0]) % dim[1], z = i / (dim[0] * dim[1]);
        int c = (x + ca * y + cb * z) % p;
        cls[i] = c;
        count[c] += 1.0;
    }
    const int adapt = sampler == 0;
    if (adapt && n > 0) or_mc_adapt(st, (double)n);
    for (int s = 0; s < nsweeps; ++s) {
        const uint64_t sweep = sweep0 + (uint64_t)s;
        long acc = 0;
        for (int k = 0; k < p; ++k) {
#pragma omp parallel for schedule(static) reduction(+ : acc)
            for (int i = 0; i < n; ++i) {
                if (cls[i] != k) continue;
                const uint32_t ctr[4] = {(uint32_t)i, (uint32_t)sweep, (uint32_t)(sweep >> 32), stream};
                uint32_t r[4];
                or_philox4x32_10(ctr, key, r);
                const double *so = m + 3 * (size_t)i;
                double sn[3];
                if (sampler == 1) {
                    sn[0] = -so[0]; sn[1] = -so[1]; sn[2] = -so[2];
                } else {
                    or_sphereft(so, st->cos_cone, (double)r[0] * two_m32, (double)r[1] * two_m32, sn);
                }
                const double e_old = or_energy_single(h, m, i, so);
                const double e_new = or_energy_single(h, m, i, sn);
                const double de = e_new - e_old;
                int ok = de < 0.0;
                if (!ok && kT > 0.0) {
                    double tmp = -de / kT;
                    if (tmp < -200.0) tmp = -200.0;
                    ok = (double)r[2] * two_m32 < exp(tmp);
                }
                if (ok) {
                    double *mi = m + 3 * (size_t)i;
                    mi[0] = sn[0]; mi[1] = sn[1]; mi[2] = sn[2];
                    ++acc;
                }
            }
        }
        st->acc = st->acc + (double)acc;
        st->nb = st->nb + (double)n;
        if (st->nb > 0.0) st->rate = st->acc / st->nb;
        if (adapt && s < nsweeps - 1) or_mc_adapt(st, (double)n);
    }
    free(cls);
    free(count);
}
