/*
 * matjes_hip.h -- C-ABI of the MI355X (gfx950) LLG hot path.
 *
 * Drop-in boundary for the reference's Fortran Hamiltonian/dynamics
 * interfaces (bind(C) from Fortran, ctypes from Python).  Plain pointers and
 * sizes only.  Every function returns 0 on success or a negative MJ_E* code;
 * mj_last_error() gives the message of the last failure on this thread.
 * Calls on one context must come from one host thread at a time.
 *
 * Data conventions (the reference's):
 *   order parameters are AoS (3, Ncell) doubles, x fastest
 *   (src/libraries_and_variables/order_parameter.f90:14-19, modes_v);
 *   COO matrices are 1-based (3N x 3N) with E = m_l^T H m_r
 *   (src/Hamiltonian_type/ham_coo.f90:103-118, pop_par);
 *   energies/fields in eV, B order parameter in T, times in fs.
 */
#ifndef MATJES_HIP_H
#define MATJES_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mj_ctx mj_ctx;
typedef struct mj_batch mj_batch;

enum {
    MJ_OK = 0,
    MJ_EINVAL = -1,   /* bad argument / shape */
    MJ_EHIP = -2,     /* HIP / rocFFT runtime failure */
    MJ_ESTATE = -3,   /* call order (e.g. no Hamiltonian set) */
    MJ_ENOMEM = -4,
    MJ_ENOTSUP = -5   /* input outside what the device path supports */
};

/* integrator codes: 1, 2, 4 as the reference's `integration` key
 * (src/solvers/solver_commun.f90:45-147); 5, 6 are new */
enum {
    MJ_EULER = 1,   /* explicit Euler (integration 1) */
    MJ_HEUN = 2,    /* the reference's two-half-step "Heun" (integration 2) */
    MJ_RK = 4,      /* the reference's Runge-Kutta of order N_order (integration 4), T = 0 */
    MJ_SIB = 5,     /* semi-implicit midpoint, Mentink et al. 2010 */
    MJ_DEPONDT = 6  /* rotation predictor-corrector, Depondt & Mertens 2009 */
};

/* order-parameter codes of op_l/op_r (src/libraries_and_variables/type_lattice.f90:15-21) */
enum { MJ_OP_M = 1, MJ_OP_B = 3 };
/* Monte Carlo proposals: the reference's `sampling sphere` and `ising` */
enum { MJ_MC_SPHERE = 0, MJ_MC_ISING = 1 };

int mj_last_error(char *buf, int len);

/* Create a context on `device` for a lattice dim_lat(3) with periodic(3)
 * flags (lattice%dim_lat, lattice%periodic; type_lattice.f90:26-60) and
 * dim_mode = 3*nmag.  Replaces the per-backend init of t_H
 * (ham_cusparse.f90:134-153 -> cuda_H.cu:45-133). */
int mj_create(mj_ctx **ctx, int device, const int dim_lat[3], const int periodic[3], int dim_mode);
int mj_destroy(mj_ctx **ctx);

/* Register one rank-2 Hamiltonian term from the reference's COO
 * (t_H_coo%pop_par, ham_coo.f90:103-118; same data cuda_H_init receives,
 * cuda_H.cu:45).  op_l/op_r: MJ_OP_M/MJ_OP_B.  Terms of the same space are
 * summed (combine_Hamiltonians, set_Hamiltonians.f90:182-231).
 * Translation-invariant M-M terms are compressed to a stencil of 3x3 blocks;
 * other M-M terms use the generic sparse path. */
int mj_add_hamiltonian_coo(mj_ctx *ctx, int op_l, int op_r, int64_t nnz, const int *row1,
                           const int *col1, const double *val);

/* Direct stencil form (extension): nofs offsets (dx,dy,dz) with 3x3 blocks
 * A(row-major) meaning H_MM(i, i+d) = A.  Equivalent to the COO of all
 * translated copies. */
int mj_add_hamiltonian_stencil(mj_ctx *ctx, int nofs, const int *ofs3, const double *blocks9);

/* The B order parameter (lattice%B%modes_v, AoS 3N); uniform fields may pass n3 == 3. */
int mj_set_bfield(mj_ctx *ctx, const double *b, int64_t n3);

/* FFT Hamiltonian (dipolar), same contract as fft_H%init_shape + init_op
 * (fft_ham_base.f90:191-224, fft_ham_fftw.f90:134-193, cuda_fft.cpp:96-143):
 * K_n(dim_mode, dim_mode, prod(N_rep)), column-major, real space, k index
 * x-fastest; periodic(3) as passed to init_shape.  The field contribution is
 * -2 * IFFT(K_F^T FFT(M)) (hamiltonian_collection.f90:450-453). */
int mj_set_fft_operator(mj_ctx *ctx, int dim_mode, const int N_rep[3], const int periodic[3],
                        const double *K_n);

/* State (lattice%M%modes_v, AoS 3N) host <-> device. */
int mj_set_state(mj_ctx *ctx, const double *m, int64_t n3);
int mj_get_state(mj_ctx *ctx, double *m, int64_t n3);
/* Same with device pointers (AoS 3N, on the context's device). */
int mj_set_state_device(mj_ctx *ctx, const double *m_dev, int64_t n3);
int mj_get_state_device(mj_ctx *ctx, double *m_dev, int64_t n3);

/* B_eff = -dH/dm of the current state into B (AoS 3N host array), exactly the
 * quantity of hamiltonian%get_eff_field(lat, B, 1)
 * (hamiltonian_collection.f90:433-454). */
int mj_eff_field(mj_ctx *ctx, double *B, int64_t n3);

/* Keep (on != 0) the effective field of the last stage of the last step of
 * every later mj_steps-family call: the Beff that spindynamics_run evaluates
 * on the last stage's state (spindynamics.f90:298, thermal term excluded) and
 * print_Beff writes as Beff_<tag>.dat (spindynamics.f90:400,
 * print_Beff.f90:10-29) under the input key gra_Beff.  The kept step runs
 * launch by launch; the buffer (24 B per site) is allocated on first use. */
int mj_keep_last_field(mj_ctx *ctx, int on);
/* That field (AoS 3N host array); MJ_ESTATE when no step ran since the field
 * was kept or since mj_set_state. */
int mj_last_field(mj_ctx *ctx, double *B, int64_t n3);

/* Total energy of the current state (hamiltonian%energy,
 * hamiltonian_collection.f90:390-400). */
int mj_energy(mj_ctx *ctx, double *E);

/* Total energy of the current state written to device memory (E_dev: one
 * double on the context's device), with no device allocation.  No counterpart
 * in the reference: for callers that cannot allocate (code under stream
 * capture, serving loops).  Its per-workgroup partials live in a static
 * device array of 131072 doubles per GPU (held until the process ends):
 * lattices with more tiles fail with MJ_EINVAL, and calls that overlap on one
 * GPU race on the array.  The Hamiltonian must already be on the device (an
 * earlier mj_energy / mj_steps after the last change, else MJ_ESTATE); not
 * with an FFT term or the generic sparse path (MJ_ENOTSUP).  The value lands
 * in the context's stream order: a later synchronising call, or a device
 * synchronisation, makes it visible. */
int mj_energy_noalloc(mj_ctx *ctx, double *E_dev);

/* Energy of each registered term of the current state (sum = mj_energy).
 * Replaces hamiltonian%energy_resolved (Hamiltonian_type/hamiltonian_collection.f90:365-388).
 * Order: the M-M terms in the order they were registered (one entry per
 * mj_add_hamiltonian_coo / mj_add_hamiltonian_stencil call), then the Zeeman
 * term, then the FFT term.  *nterms holds the capacity of E on entry and the
 * number of terms on return; MJ_EINVAL (count still returned) if E is too
 * short, MJ_ENOTSUP when an M-M term is not translation invariant. */
int mj_energy_terms(mj_ctx *ctx, double *E, int *nterms);

/* Product of one registered term with a vector, the backend interface
 * t_H_base%mult_r / mult_l (Hamiltonian_type/ham_base.f90:41,88-96, dense
 * form ham_dense.f90:30-87): side 0: res = alpha H vec + beta res; side 1:
 * res = alpha H^T vec + beta res (beta = 0: res is not read).  term numbers
 * as in mj_energy_terms (M-M terms in registration order, then the Zeeman
 * term, whose H maps the M space to the B space).  Host vectors, 3*Ncell
 * each (AoS).  The term's offsets and blocks are copied to the device on its
 * first call and kept there, with two staging vectors sized for the lattice,
 * so that later calls allocate nothing and only copy vec in and res out (the
 * reference's device backend, ham_cusparse.f90:174-236, also copies both
 * vectors per call). */
int mj_term_mult(mj_ctx *ctx, int term, int side, const double *vec, int64_t nvec, double *res, int64_t nres,
                 double alpha, double beta);

/* N_order of integration 4 (the input key `N_order`, read by
 * select_propagator, solver_commun.f90:112-121): 1, 2, 3, 4, 6 or 7, the orders
 * get_butcher_explicit has tables for (solver_order.f90:65-127); 4 until set. */
int mj_set_rk_order(mj_ctx *ctx, int n_order);

/* Advance the device state by nsteps of the LLG loop
 * (spindynamics.f90:260-323).  timestep in fs; kT in eV (T*k_B).  Thermal
 * field: if kT*damping > 1e-10, generated in-kernel from Philox4x32-10 with
 * key `seed` and step counter step0+j (integration 4 has none in the
 * reference, which selects no thermal field for it: MJ_ENOTSUP).  energies
 * (nsteps doubles, may be NULL) receives E/Ncell after each step
 * (spindynamics.f90:323). */
int mj_steps(mj_ctx *ctx, int nsteps, int integrator, double timestep, double damping, double kT,
             uint64_t seed, uint64_t step0, double *energies);

/* As mj_steps with one kT per step (kT_steps, nsteps values in eV): step j's
 * thermal amplitude is formed from kT_steps[j] as spindynamics.f90:266-271
 * forms it from kt (sqrt(damping*kT*hbar/dt), no thermal term where
 * kT*damping <= 1e-10).  For a temperature that varies in time such as the
 * pulse of update_TPulse (TPulse.f90:71-105), which the reference defines but
 * does not call: its loop keeps kt = ktini (spindynamics.f90:186-188).  The
 * steps run launch by launch. */
int mj_steps_temperature(mj_ctx *ctx, int nsteps, int integrator, double timestep, double damping,
                         const double *kT_steps, uint64_t seed, uint64_t step0, double *energies);

/* As mj_steps with the reference's adaptive timestep (update_time,
 * src/solvers/update_time.f90:41-75, called at spindynamics.f90:438; off in
 * the reference unless `update_time T`): step j runs with *timestep, then
 * *timestep = 2 pi hbar / max_i |B_i| / discretization, B the effective field
 * of the step's last stage.  On return *timestep holds the next step's value;
 * timesteps (nsteps doubles, may be NULL) receives the value each step used.
 * Fails with MJ_EINVAL where the reference stops (max |B| <= 1e-8). */
int mj_steps_adaptive(mj_ctx *ctx, int nsteps, int integrator, double *timestep, double damping, double kT,
                      uint64_t seed, uint64_t step0, double discretization, double *energies, double *timesteps);

/* Time-dependent B order parameter, the excitations of the reference
 * (update_exc before every stage, dynamics/spindynamics.f90:285-292,
 * excitations/excitations.f90:164-215: b_i(t) = sum_e T_e(t) (x) (A_i cos(phi_e(t)
 * + psi_i) + o_i), a sum of fixed site maps with time-dependent weights).
 * mj_set_bfield_maps: 0..4 maps F_k (nmaps x 3N AoS host doubles); with no map
 *   the field is uniform.
 * mj_steps_driven: as mj_steps with b_i = sum_k w_k (x) F_k,i in every stage
 *   (componentwise; nmaps = 0: b_i = w): weights at
 *   [step][stage][3 * max(nmaps, 1)], stages 1 (Euler) or 2; weights_end
 *   [step][3 * max(nmaps, 1)] the field of the energy after each step (NULL:
 *   the last stage's).  The static field of mj_set_bfield is not used by these
 *   steps and is kept for other calls.  Needs a Zeeman term (MJ_ESTATE). */
int mj_set_bfield_maps(mj_ctx *ctx, int nmaps, const double *maps, int64_t n);
int mj_steps_driven(mj_ctx *ctx, int nsteps, int integrator, double timestep, double damping, double kT,
                    uint64_t seed, uint64_t step0, const double *weights, const double *weights_end,
                    double *energies);

/* As mj_steps, with the thermal field supplied by the caller (the reference's
 * random_numbers%x stream already scaled by sigma; nsteps x 3N host doubles). */
int mj_steps_noise(mj_ctx *ctx, int nsteps, int integrator, double timestep, double damping,
                   double kT, const double *noise, double *energies);

/* Timing helper for measurements: runs nsteps like mj_steps (energies after
 * every step when `energies` is non-zero, as the reference's loop computes
 * them) with HIP events on the context's stream and returns total
 * milliseconds and, per kernel kind, the summed launch milliseconds and
 * launch counts (kinds: 0 stage-1 kernel, 1 stage-2 kernel or the kernel of a
 * one-kernel step, 2 other).  Steps of one kernel (the fused Heun, SIB and
 * Depondt kernels of a compiled Hamiltonian, Euler) are timed as one block of
 * back-to-back launches; steps of two kernels with an event pair around
 * every launch. */
int mj_profile_steps(mj_ctx *ctx, int nsteps, int integrator, double timestep, double damping,
                     double kT, uint64_t seed, int energies, double *total_ms, double *kind_ms, int *kind_count);

/* The device's correctly rounded sqrt and reciprocal against the library
 * operations, on n host values (tests of the arithmetic; no counterpart in
 * the reference): fn 0 the stage kernels' sqrt, 1 sqrt, 2 their reciprocal,
 * 3 1.0 / x. */
int mj_math_probe(mj_ctx *ctx, int fn, int64_t n, const double *x, double *y);

/* Thermal normals of one step (for inspection/tests): out (AoS 3N) = N(0,1) draws. */
int mj_thermal_normals(mj_ctx *ctx, uint64_t seed, uint64_t step, double *out, int64_t n3);

/* Kernel tile shape of the stencil path (0 = automatic). */
int mj_set_tile(mj_ctx *ctx, int bx, int by, int bz);

/* ---- Differentiable SIB trajectory (T = 0) --------------------------------
 * No counterpart in the reference (it has no gradients; its SIB-like update is
 * not implemented either).  The M-M Hamiltonian of the trajectory is
 * H_MM = sum_k theta_k H^(k) over `nparam` unit terms (e.g. exchange with J = 1,
 * DMI with D = 1, anisotropy with K = 1), each given as translation-invariant
 * 3x3 blocks H^(k)(i, i+d) in the reference's H_MM convention (the blocks that
 * mj_add_hamiltonian_stencil takes; c_H_J etc. already applied).  The Zeeman
 * block registered on the context (op_l = MJ_OP_B) is used with the per-site
 * order parameter b_i = b_ext + h_i (Tesla).
 *
 * mj_adjoint_setup: nofs[k] offsets for parameter k; ofs3 (3 ints per offset)
 *   and blocks9 (9 doubles per offset, row-major) concatenated over k.
 *   At most 13 parameters.
 * mj_adjoint_forward: device pointers, AoS (N,3) doubles: hmap (may be NULL),
 *   m0, mfinal.  Runs nsteps SIB steps, keeping a checkpoint every ckpt_every
 *   steps in device memory.
 * mj_adjoint_backward: given gfinal = dL/dm_final (device, AoS), returns
 *   dL/dtheta (host, nparam), dL/db_ext (host, 3), dL/dh (device AoS, may be
 *   NULL) and dL/dm0 (device AoS, may be NULL).  Segments are recomputed from
 *   their checkpoints; the reverse step is the exact transpose of the discrete
 *   forward step.  Both calls synchronise the library's stream before they
 *   return; device inputs must be ready when they are called. */
int mj_adjoint_setup(mj_ctx *ctx, int nparam, const int *nofs, const int *ofs3, const double *blocks9);
int mj_adjoint_forward(mj_ctx *ctx, const double *theta, const double *b_ext, const double *hmap_dev,
                       const double *m0_dev, int nsteps, double timestep, double damping, int ckpt_every,
                       double *mfinal_dev);
int mj_adjoint_backward(mj_ctx *ctx, const double *gfinal_dev, double *gtheta, double *gb_ext, double *ghmap_dev,
                        double *gm0_dev);

/* ---- Batched systems and parallel tempering --------------------------------
 * No counterpart in the reference (it runs one lattice per process and has no
 * replica exchange).  A batch advances many lattices with one launch per
 * integrator stage; the steps between exchanges are captured in a hipGraph.
 *
 * mj_batch_add: copies a context's finalized translation-invariant Hamiltonian
 *   (a uniform B or a per-site B map; an FFT dipolar term gets the batch's own
 *   copy of its transform, and a batch holding one runs every step eagerly,
 *   one launch per stage, the systems' H_fft before each) and current state; the system then runs the
 *   batch's integration with `damping`, kT in eV (T*k_B, as mj_steps) and its
 *   own Philox stream `seed`.  Its trajectory equals mj_steps on the context
 *   bit for bit.  *index receives the system number.
 * mj_batch_set_pt: systems first .. first+count-1 (same lattice, Hamiltonian
 *   and damping; no dipolar term) become replicas on the ladder kT_ladder[0..count-1] in eV (replica
 *   k starts at rung k); every `every` steps neighbouring rungs swap with the
 *   Metropolis probability of their replicas' energies (uniforms from Philox
 *   keyed by `seed`).
 * mj_batch_set_tiles: 2 ints per system (bx, by with bx*by = 256, or 0 0).
 * mj_batch_set_launch: how the captured segments of integration 2 run (the
 *   trajectories and exchanges are the same bit for bit in every form):
 *   MJ_BATCH_STAGES one launch per stage; MJ_BATCH_FUSED (default) one launch
 *   per step, both stages per tile (the tile plus twice the stencil reach in
 *   LDS, stage 1 repeated on the ring around the tile); MJ_BATCH_PERSISTENT /
 *   MJ_BATCH_PERSISTENT_FUSED one launch per segment (as many workgroups as
 *   are resident at once, grid barriers between the stages or steps).  A batch
 *   whose fused tile exceeds 64 KiB of LDS runs the per-stage counterpart.
 *   SIB and Depondt run the fused step as well unless MJ_BATCH_STAGES is
 *   asked for (no persistent form); Euler runs one launch per step.  A persistent barrier waits
 *   at most `spin` sleeps of about 0.25 us (negative: 2^22, about 1 s), then
 *   every workgroup leaves and the run returns MJ_EHIP with the states
 *   undefined.
 * mj_batch_steps / mj_batch_profile (HIP-event time of the steps):
 *   integrator MJ_EULER, MJ_HEUN, MJ_SIB or MJ_DEPONDT (not MJ_RK); for the
 *   integrators whose first stage carries the thermal field the exchange
 *   steps add a field pass for the entering energies.
 * mj_batch_energies: energy per site of every system's current state.
 * mj_batch_pt_state: rung of every replica, accepted/attempted swaps per rung
 *   pair (count entries each, the last unused). */
int mj_batch_create(mj_batch **batch, int device);
int mj_batch_destroy(mj_batch **batch);
int mj_batch_add(mj_batch *batch, mj_ctx *ctx, double damping, double kT, uint64_t seed, int *index);
int mj_batch_set_pt(mj_batch *batch, int first, int count, const double *kT_ladder, int every, uint64_t seed);
int mj_batch_set_tiles(mj_batch *batch, const int *tiles);
enum { MJ_BATCH_STAGES = 0, MJ_BATCH_FUSED = 1, MJ_BATCH_PERSISTENT = 2, MJ_BATCH_PERSISTENT_FUSED = 3 };
int mj_batch_set_launch(mj_batch *batch, int mode, int64_t spin);
int mj_batch_steps(mj_batch *batch, int nsteps, int integrator, double timestep);
int mj_batch_profile(mj_batch *batch, int nsteps, int integrator, double timestep, double *total_ms);
int mj_batch_get_state(mj_batch *batch, int sys, double *m, int64_t n3);
int mj_batch_energies(mj_batch *batch, double *energy_per_site);
int mj_batch_pt_state(mj_batch *batch, int *rung_of_replica, uint64_t *accepted, uint64_t *attempted);
int mj_batch_info(mj_batch *batch, int64_t *sites, int *workgroups, int *tiles);

/* ---- Metropolis Monte Carlo ------------------------------------------------
 * mj_energy_single: energy of single sites, hamiltonian%energy_single
 *   (Hamiltonian_type/hamiltonian_collection.f90:402-431): pair terms twice
 *   their row, on-site and Zeeman terms once.  sites1: n site indices, 1-based
 *   (the reference's i_spin); spins: 3 per site to evaluate the site with that
 *   spin instead of its current one, or NULL.  MJ_ENOTSUP with an FFT term or
 *   a generic (not translation invariant) M-M operator.
 * mj_mc_sweeps: nsweeps Metropolis sweeps of the state, the update of MCstep
 *   (Metropolis-MC/MCstep.f90:17-65) on every site once per sweep: the sites
 *   are split into colour classes that share no term of the Hamiltonian and
 *   each class is updated in one pass (replaces the random single-site order
 *   of choose_spin, choosespin.f90:14-22).  sampler MJ_MC_SPHERE (sphereft
 *   proposal in the adaptive cone, paratemp-MC/sampling.f90:9-62,
 *   track_value.f90:138-160) or MJ_MC_ISING (flip).  Random numbers from
 *   Philox4x32-10 keyed by seed, counter (site, sweep0 + s).  After every
 *   measure_every sweeps (0: never) the total energy and the sum of the spins
 *   are recorded: E_samples (nsweeps / measure_every), M_samples (3 each);
 *   either may be NULL.  kT in eV; kT = 0 accepts only energy decreases.
 * mj_mc_set_cone: the initial cone angle (io_MC%cone, track_value.f90:54-57),
 *   acceptance counters reset; mj_mc_stats: cone, accepted, proposed and
 *   acceptance rate (track_val cone, acc, nb, rate), and when colouring3 is
 *   not NULL the colouring (p, a, b): class = (x + a y + b z) mod p. */
int mj_energy_single(mj_ctx *ctx, int64_t n, const int64_t *sites1, const double *spins, double *E);
int mj_mc_sweeps(mj_ctx *ctx, int nsweeps, int sampler, double kT, uint64_t seed, uint64_t sweep0,
                 int measure_every, double *E_samples, double *M_samples);
int mj_mc_set_cone(mj_ctx *ctx, double cone);
int mj_mc_stats(mj_ctx *ctx, double *out4, int *colouring3);

/* Parallel tempering Monte Carlo (paratemp-MC/parallel_tempering.f90:24-217):
 * nrep replicas of the lattice, swept in the same launches, one temperature
 * each.
 * mj_mc_pt_setup: every replica starts from the context's state, replica r at
 *   rung r of kT_ladder (eV), cone angle `cone` and zero counters each.
 * mj_mc_pt_sweeps: nsweeps sweeps of every replica (as mj_mc_sweeps; replica r
 *   draws from Philox counter word 3 = 'MC' + r); every exchange_every sweeps
 *   (0: never) neighbouring rungs swap their replicas as reorder_temperature
 *   (paratemp.f90:47-121) decides: pairs (q, q+1) from q = (n + 1) mod 2 for the
 *   n-th exchange, accepted when (1/kT_{q+1} - 1/kT_q)(E_{q+1} - E_q) > 0 or
 *   u < exp(max(that, -200)).  Every measure_every sweeps (before that sweep's
 *   exchange) the total energy and summed spins of the replica at each rung
 *   are recorded: E_samples (nsamp x nrep), M_samples (nsamp x nrep x 3).
 * mj_mc_pt_state: replica at each rung, accepted / attempted swaps per rung
 *   pair (nrep - 1), and per replica cone, accepted, proposed, rate (4 each);
 *   any pointer may be NULL.  mj_mc_pt_get_state: the state at a rung (AoS). */
int mj_mc_pt_setup(mj_ctx *ctx, int nrep, const double *kT_ladder, double cone);
int mj_mc_pt_sweeps(mj_ctx *ctx, int nsweeps, int sampler, uint64_t seed, uint64_t sweep0, int exchange_every,
                    int measure_every, double *E_samples, double *M_samples);
int mj_mc_pt_state(mj_ctx *ctx, int *image_of_rung, double *swaps_accepted, double *swaps_attempted,
                   double *replica_stats);
int mj_mc_pt_get_state(mj_ctx *ctx, int rung, double *m, int64_t n3);

#ifdef __cplusplus
}
#endif
#endif
