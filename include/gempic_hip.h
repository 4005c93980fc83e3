/*
 * gempic_hip.h -- C ABI of the MI355X (gfx950) particle-field library.
 *
 * A Julia binding (`ccall`) or any C FFI drives GEMPIC.jl's particle-field
 * hot path through these entry points. They take plain pointers and sizes;
 * all device memory belongs to opaque handles. Nothing here falls back to
 * the CPU: without a usable GPU every compute call fails with
 * GEMPIC_ERR_DEVICE.
 *
 * Conventions (same as the reference, GEMPIC.jl):
 *   - particle arrays are the ParticleGroup matrix, column-major
 *     (D+V+n_weights) x N Float64: element (r, i) at aos[r + i*ld]
 *     (particle_group.jl:15-46);
 *   - grid arrays (e_dofs, b_dofs, j_dofs, rho) are nx Float64;
 *   - indices and counts are int64_t;
 *   - every entry point returns 0 (GEMPIC_OK) or an error code, with a
 *     message from gempic_last_error(). ArgumentError / AssertionError of
 *     the reference map to GEMPIC_ERR_ARGUMENT / GEMPIC_ERR_ASSERTION, and
 *     a non-finite or out-of-range particle position (InexactError from
 *     trunc(Int, x) in the reference) to GEMPIC_ERR_INEXACT.
 *   - A handle is not re-entrant: callers serialize calls on one handle.
 *     Each handle owns one HIP stream; calls return after their work is
 *     complete unless stated otherwise.
 */
#ifndef GEMPIC_HIP_H
#define GEMPIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GEMPIC_OK 0
#define GEMPIC_ERR_ARGUMENT 1
#define GEMPIC_ERR_ASSERTION 2
#define GEMPIC_ERR_DEVICE 3
#define GEMPIC_ERR_INEXACT 4
#define GEMPIC_ERR_STATE 5

/* smoothing_type of ParticleMeshCoupling1D (particle_mesh_coupling_1d.jl:56-65) */
#define GEMPIC_GALERKIN 0
#define GEMPIC_COLLOCATION 1

typedef struct gempic_splitting gempic_splitting;
typedef struct gempic_maxwell1d gempic_maxwell1d;
typedef struct gempic_maxwell2d gempic_maxwell2d;

/* Library version (major*10000 + minor*100 + patch) and last error message
 * of the calling thread (valid until the next call on that thread). */
int gempic_version(void);
const char *gempic_last_error(void);

/* Number of visible GPUs (0 when none). */
int gempic_device_count(void);
/* The current GPU's compute units and architecture name (as HIP reports it,
 * e.g. "gfx950:sramecc+:xnack-"); either may be NULL. No counterpart in the
 * reference: the launch autotuner keys its cache on it. */
int gempic_device_info(int *cus, char *arch, int arch_len);

/* ------------------------------------------------------------------------
 * HamiltonianSplitting{1,V} (hamiltonian_splitting.jl:20-86) together with
 * the Maxwell1DFEM (maxwell_1d_fem.jl:49-177), the two
 * ParticleMeshCoupling1D smoothers (particle_mesh_coupling_1d.jl:44-95) and
 * the ParticleGroup{1,V} (particle_group.jl:25-45) it holds.
 *
 * nx, xmin, xmax: the OneDGrid (mesh.jl:52-67); maxwell_degree: s_deg_0 of
 * Maxwell1DFEM (1..3); degree0/smoothing0: kernel_smoother_0, degree1/
 * smoothing1: kernel_smoother_1 (degrees 0..5); dim_v: 1 (1d1v) or 2 (1d2v);
 * n_weights must be 1; common_weight 0 means 1/N as in ParticleGroup.
 * ------------------------------------------------------------------------ */
int gempic_splitting_create(int64_t nx, double xmin, double xmax, int maxwell_degree,
                            int degree0, int smoothing0, int degree1, int smoothing1,
                            int dim_v, int64_t n_particles, int n_weights, double charge,
                            double mass, double common_weight, gempic_splitting **out);
int gempic_splitting_destroy(gempic_splitting *h);

/* Particle array in and out (the aliased ParticleGroup.array).
 * ld = 1 + dim_v + n_weights. */
int gempic_set_particles(gempic_splitting *h, const double *aos, int64_t ld, int64_t n);
int gempic_get_particles(gempic_splitting *h, double *aos, int64_t ld, int64_t n);

/* e_dofs[1], e_dofs[2], b_dofs (the aliased field vectors,
 * hamiltonian_splitting.jl:80-81); any pointer may be NULL to skip it. */
int gempic_set_fields(gempic_splitting *h, const double *e1, const double *e2, const double *b);
int gempic_get_fields(gempic_splitting *h, double *e1, double *e2, double *b);
/* j_dofs[1], j_dofs[2] as left by the last operator */
int gempic_get_j(gempic_splitting *h, double *j1, double *j2);

/* Sub-steps (drop-in for the Julia methods):
 *   operatorHB  -- hamiltonian_splitting_1d2v.jl:234-236 / _1d1v.jl:113-125
 *   operatorHE  -- hamiltonian_splitting_1d2v.jl:191-219 (1d2v only)
 *   operatorHp1 -- hamiltonian_splitting_1d2v.jl:41-112 / _1d1v.jl:63-98
 *   operatorHp2 -- hamiltonian_splitting_1d2v.jl:129-176 (1d2v only) */
int gempic_operator_hb(gempic_splitting *h, double dt);
int gempic_operator_he(gempic_splitting *h, double dt);
int gempic_operator_hp1(gempic_splitting *h, double dt);
int gempic_operator_hp2(gempic_splitting *h, double dt);

/* strang_splitting! (hamiltonian_splitting.jl:98-108; _1d1v.jl:11-21).
 * GEMPIC_STRANG_FUSED (default) runs one particle pass per step: the
 * step-boundary HE pair and Hp2-Hp1-Hp2 read only B between two field
 * solves, so they are fused; a second, small kernel then sums the
 * workgroups' deposits and runs the field solves. GEMPIC_STRANG_SUBSTEP
 * launches the seven sub-steps one by one.
 * Particle states and fields after each step are those of the sub-step
 * sequence. With GEMPIC_ASYNC the call returns once the work is queued. */
#define GEMPIC_STRANG_FUSED 0
#define GEMPIC_STRANG_SUBSTEP 1
#define GEMPIC_ASYNC 16
int gempic_strang_splitting(gempic_splitting *h, double dt, int64_t n_steps, int flags);

/* Wait for queued work and report any deferred error. */
int gempic_synchronize(gempic_splitting *h);

/* Ensembles (no counterpart in the reference; the batched form of
 * strang_splitting!, hamiltonian_splitting.jl:98-108): independent 1d2v
 * splittings, each with its own particles, grid size, domain and fields,
 * advanced together with one particle-pass launch and one field-part launch
 * per Strang step for all of them. Each splitting's results equal those of
 * its own gempic_strang_splitting call bit for bit. The handles stay owned
 * by the caller and must outlive the ensemble; they must share one GPU and
 * one smoother pair. The call is synchronous. One step of every splitting is
 * taken per launch pair, each with its own dt. */
typedef struct gempic_ensemble gempic_ensemble;
int gempic_ensemble_create(int n_instances, gempic_splitting *const *handles, gempic_ensemble **out);
int gempic_ensemble_destroy(gempic_ensemble *e);
/* dt: one time step per splitting (n_instances doubles). flags:
 * GEMPIC_GRAPH replays the steps between the first and the last as a HIP
 * graph of the step's launches (captured once per ensemble, again after a
 * change of the launch geometry); the due particle sorts run between
 * replays. Results equal those of the stream-queued steps bit for bit. */
#define GEMPIC_GRAPH 64
int gempic_ensemble_strang_splitting(gempic_ensemble *e, const double *dt, int64_t n_steps, int flags);
/* Launch autotuning of an ensemble (no counterpart in the reference; it
 * replaces the JULIA_NUM_THREADS knob of hamiltonian_splitting.jl:62 for a
 * batch): for each candidate number of particles per workgroup (>= 256),
 * every instance's pass takes ceil(N / ppw) workgroups; n_steps batched
 * steps are timed twice per candidate from the same state, which is
 * restored on the device after every trial; the fastest candidate stays set
 * on the member handles. ms_per_step (n_candidates doubles) may be NULL. */
int gempic_ensemble_autotune(gempic_ensemble *e, const double *dt, int64_t n_steps, const int *ppw,
                             int n_candidates, int *best_ppw, double *ms_per_step);
/* The reverse mode of every splitting of the ensemble (gempic_strang_splitting_adjoint
 * per splitting, batched: one launch per pass, re-run and field part for all
 * of them), with loss loss_weight[i] * sum_s 1/2 l2norm_squared(B_s, p-1)
 * and no cotangents of the final state. Each splitting's results equal those
 * of its own gempic_strang_splitting_adjoint bit for bit; gempic_get_adjoint
 * and gempic_get_adjoint_domain read them per splitting. */
int gempic_ensemble_strang_splitting_adjoint(gempic_ensemble *e, const double *dt, int64_t n_steps,
                                             int64_t checkpoint_every, const double *loss_weight,
                                             int flags);

/* Reverse mode of strang_splitting! (1d2v). The reference has no
 * derivatives; this is the pullback a ChainRulesCore.rrule of
 * strang_splitting!(h, dt, n_steps) (hamiltonian_splitting.jl:98-108) would
 * call. The call runs the n_steps forward (the handle ends in the same state
 * as after gempic_strang_splitting, without the periodic sort), recording the
 * fields of every step and the particles every checkpoint_every steps, then
 * sweeps back, re-running each checkpoint interval forward on the GPU.
 * The loss is
 *   loss_weight * sum_{s=1..n_steps} 1/2 l2norm_squared(B_s, p-1)
 *     + <ct_particles, final x, v1, v2> + <ct_e1, E1> + <ct_e2, E2> + <ct_b, B>
 * (l2norm_squared: maxwell_1d_fem.jl:299-313; with loss_weight = dt the sum
 * is the time-integrated magnetic energy of write_step!, diagnostics.jl:236).
 * ct_* may be NULL (zero); ct_particles is the caller's AoS layout (rows x,
 * v1, v2 read, ld = D+V+n_weights). 4 <= nx <= 128. Results: gempic_get_adjoint. */
int gempic_strang_splitting_adjoint(gempic_splitting *h, double dt, int64_t n_steps,
                                    int64_t checkpoint_every, double loss_weight,
                                    const double *ct_particles, int64_t ld, const double *ct_e1,
                                    const double *ct_e2, const double *ct_b, int flags);
/* The loss and the cotangents of the state the adjoint call started from:
 * g_particles in the caller's AoS layout (rows x, v1, v2; other rows zero),
 * g_e1, g_e2, g_b (nx each). Any output may be NULL. */
int gempic_get_adjoint(gempic_splitting *h, double *loss, double *g_particles, int64_t ld,
                       int64_t n, double *g_e1, double *g_e2, double *g_b);
/* The same call's explicit derivatives of the loss in the domain: g_length =
 * d loss / d Lx with the particles' positions, velocities and weights and the
 * fields held fixed (through dx = Lx/nx in every operator, and the periods the
 * mod adds in Hp1, hamiltonian_splitting_1d2v.jl:79); g_weight_sum = the sum
 * over particles of d loss / d w_i. With the cotangents of gempic_get_adjoint
 * they give the derivative in a parameter that moves Lx (a wavenumber k with
 * Lx = 2 pi / k) by the chain rule through the initial state. Needs a
 * smoother pair with p - 1 >= 1 (continuous splines along Hp1's path). */
int gempic_get_adjoint_domain(gempic_splitting *h, double *g_length, double *g_weight_sum);
/* Reverse mode of solve_poisson! (diagnostics.jl:15-31) in the positions:
 * g_x[i] = d <ct_e, E> / d x_i for E = compute_e_from_rho!(add_charge!(x))
 * of the handle's particles, in the caller's order (n doubles). */
int gempic_solve_poisson_adjoint(gempic_splitting *h, const double *ct_e, double *g_x);

/* Particles are kept sorted by position on the device: by cell, then by
 * sub-cell bin (their order in the caller's array is restored by
 * gempic_get_particles). interval = re-sort every `interval` Strang steps
 * (0 = never after the upload). */
int gempic_set_sort_interval(gempic_splitting *h, int64_t interval);
int gempic_sort_particles(gempic_splitting *h);

/* solve_poisson! (diagnostics.jl:15-31) on the handle's particles with
 * kernel_smoother_0; efield and rho are host arrays of nx (either may be
 * NULL). */
int gempic_solve_poisson(gempic_splitting *h, double *efield, double *rho);

/* write_step! (diagnostics.jl:186-250) for 1d2v: out[0..10] = Time,
 * KineticEnergy, Momentum1, Momentum2, PotentialEnergyE1,
 * PotentialEnergyE2, PotentialEnergyB3, Transfer, VVB, Poynting,
 * ErrorPoisson. e*_n are the "efield_dofs_n" arguments (NULL = current
 * fields), efield_poisson the host Poisson field (NULL = recompute with
 * gempic_solve_poisson). */
int gempic_write_step(gempic_splitting *h, double time, int degree, const double *e1_n,
                      const double *e2_n, const double *efield_poisson, double *out11);

/* Per-kernel timing with HIP events recorded on the handle's stream.
 * kernel ids: 0 fused pass (the second kernel of a two-kernel step),
 * 1 trailing pass, 2 HE, 3 Hp2, 4 Hp1, 5 HB/grid, 6 sort, 7 other,
 * 8 first kernel of a two-kernel step, 9 the deposit sum and field solves
 * that follow a deposit pass, 10 reverse pass, 11 forward re-run of a
 * checkpoint interval, 12 field part of a reverse step. total_ms / count accumulate since the
 * last reset. enable: 0 off, 1 every launch, n > 1 every n-th launch of each
 * kernel id (the event pair costs a few microseconds of stream time, so
 * sampling keeps it out of a timed run). */
int gempic_profile_enable(gempic_splitting *h, int enable);
int gempic_profile_read(gempic_splitting *h, int kernel_id, double *total_ms, int64_t *count);
int gempic_profile_reset(gempic_splitting *h);

/* Workgroups per particle pass (0 = automatic: CUs times the workgroups of
 * the pass kernel that fit on one CU; at most 8 per CU). */
int gempic_set_launch(gempic_splitting *h, int blocks);

/* Launch autotuner. No counterpart in the reference, whose parallelism is the
 * Julia thread count (hamiltonian_splitting.jl:62, nthreads()); it replaces
 * that knob. Times n_steps fused Strang steps of dt (gempic_strang_splitting
 * with flags 0) for every pair (blocks[i], sort_intervals[j]) on the handle's
 * own particles and fields, twice each, restoring particles, fields and the
 * sort phase after every trial, then keeps the fastest pair (as
 * gempic_set_launch and gempic_set_sort_interval would). blocks: workgroups
 * per particle pass, 0 = automatic; sort_intervals: Strang steps between
 * sorts, 0 = never. ms_per_step (may be NULL): n_blocks * n_intervals values,
 * row i for blocks[i]. On an error the state and both settings are left as
 * they were. Every pair is deterministic, but different pairs add the
 * deposits in different orders, so results agree to rounding, not bitwise. */
int gempic_autotune(gempic_splitting *h, double dt, int64_t n_steps, const int *blocks,
                    int n_blocks, const int64_t *sort_intervals, int n_intervals,
                    int *best_blocks, int64_t *best_interval, double *ms_per_step);

/* Fused Strang path of 1d2v in one particle kernel per step or in two
 * (HE pair + first Hp2, then Hp1 + second Hp2 + field solves). Both give
 * bitwise the same results; two kernels hold fewer registers. Default 1. */
int gempic_set_fusion(gempic_splitting *h, int kernels_per_step);

/* Arithmetic of the 1d2v particle passes run by gempic_strang_splitting and
 * the gempic_operator_* calls. GEMPIC_ARITH_REFERENCE (the default) rounds
 * every product and sum as the reference's Julia code does
 * (particle_mesh_coupling_1d.jl:261-453, low_level_bsplines.jl:63-80,
 * hamiltonian_splitting_1d2v.jl:41-219): per-particle results are bitwise
 * the reference's. GEMPIC_ARITH_FUSED has no counterpart in the reference: it
 * evaluates the same quantities with fewer FP64 operations (closed-form
 * spline bases, fused multiply-adds, each current segment's integral by its
 * midpoint value plus the exact curvature term instead of Gauss-Legendre),
 * which agree with the reference to rounding (a few ulp per operation), not
 * bitwise. It needs the one-kernel fused step. An ensemble's forward runs in
 * its members' arithmetic, which they must share; the reverse mode (of a
 * splitting or an ensemble), the Boris splitting and 1d1v always use the
 * reference's arithmetic. */
#define GEMPIC_ARITH_REFERENCE 0
#define GEMPIC_ARITH_FUSED 1
int gempic_set_arithmetic(gempic_splitting *h, int arith);

/* ------------------------------------------------------------------------
 * sample!(pg, alpha, k, sigma, mesh) (particle_sampling.jl:266-311) for
 * ParticleGroup{1,1} and {1,2}: Sobol.jl's SobolSeq(2) points, speeds
 * sigma sqrt(-2 log((i - 1/2)/n)), x = newton(r2, alpha, k) (:236-245),
 * weight = mesh.dimx. With V = 2, mesh.dimx must be 2 pi / k (:297).
 * ------------------------------------------------------------------------ */
/* into a host array: (2 + dim_v) x n column-major, leading dimension ld */
int gempic_sample_landau(int dim_v, int64_t n, double alpha, double k, double sigma, double dimx,
                         double *aos, int64_t ld);
/* directly into a splitting handle's GPU particles (dimx = the mesh's;
 * the particles are then sorted as after gempic_set_particles) */
int gempic_sample_landau_particles(gempic_splitting *h, double alpha, double k, double sigma);

/* ------------------------------------------------------------------------
 * HamiltonianSplittingBoris (hamiltonian_splitting_boris.jl:23-88) on a
 * splitting handle with V = 2 (the same particles, e_dofs, b_dofs and
 * j_dofs); its e_dofs_mid and b_dofs_mid live on the handle (zero at
 * creation, :50-60). One strang_splitting! step is one particle pass.
 * ------------------------------------------------------------------------ */
int gempic_boris_staggering(gempic_splitting *h, double dt);              /* staggering! :99-122 */
int gempic_boris_strang_splitting(gempic_splitting *h, double dt, int64_t n_steps,
                                  int flags);                             /* :132-177 */
int gempic_boris_push_v_epart(gempic_splitting *h, double dt);            /* :189-204 */
int gempic_boris_push_v_bpart(gempic_splitting *h, double dt);            /* :211-233 */
int gempic_boris_push_x_accumulate_j(gempic_splitting *h, double dt);     /* :250-288 */
/* e_dofs_mid[1], e_dofs_mid[2], b_dofs_mid (nx each; any may be NULL) */
int gempic_boris_set_mid(gempic_splitting *h, const double *e1_mid, const double *e2_mid,
                         const double *b_mid);
int gempic_boris_get_mid(gempic_splitting *h, double *e1_mid, double *e2_mid, double *b_mid);

/* ------------------------------------------------------------------------
 * ParticleMeshCoupling1D batched per-particle verbs over host arrays of n
 * particles, in the reference's particle order (one call = the reference's
 * loop over particles). nx/xmin/xmax/degree/smoothing as in the
 * constructor (particle_mesh_coupling_1d.jl:44-95).
 * ------------------------------------------------------------------------ */
/* add_charge! (:261-280): rho[nx] += sum_i deposit(x[i], q[i]) */
int gempic_pmc1d_add_charge(int64_t nx, double xmin, double xmax, int degree, int smoothing,
                            int64_t n, const double *x, const double *q, double *rho);
/* add_charge_pp! (:106-122) */
int gempic_pmc1d_add_charge_pp(int64_t nx, double xmin, double xmax, int degree, int smoothing,
                               int64_t n, const double *x, const double *q, double *rho);
/* evaluate (:438-453; batched form diagnostics.jl:260-268): out[i] = f(x[i]) */
int gempic_pmc1d_evaluate(int64_t nx, double xmin, double xmax, int degree, int smoothing,
                          int64_t n, const double *x, const double *field, double *out);
/* evaluate_pp (:242-250) on the (degree+1) x nx pp matrix from b_to_pp */
int gempic_pmc1d_evaluate_pp(int64_t nx, double xmin, double xmax, int degree, int smoothing,
                             int64_t n, const double *x, const double *field_pp, double *out);
/* add_current_update_v! with B (:296-376) when bfield != NULL, else the
 * 1d1v form without B (:471-529). j[nx] += current; v[i] updated in place
 * (unchanged without B). */
int gempic_pmc1d_add_current_update_v(int64_t nx, double xmin, double xmax, int degree,
                                      int smoothing, int64_t n, const double *x_old,
                                      const double *x_new, const double *q, double qm,
                                      const double *bfield, double *v, double *j);
/* add_current_update_v_pp! (:131-190) */
int gempic_pmc1d_add_current_update_v_pp(int64_t nx, double xmin, double xmax, int degree,
                                         int smoothing, int64_t n, const double *x_old,
                                         const double *x_new, const double *q, double qm,
                                         const double *bfield, double *v, double *j);

/* b_to_pp (splinepp.jl:241-261): pp[(degree+1) x ncells] from b[ncells] */
int gempic_b_to_pp(int degree, int64_t ncells, const double *b, double *pp);

/* ------------------------------------------------------------------------
 * Maxwell1DFEM (maxwell_1d_fem.jl) verbs on host arrays, run on the GPU.
 * ------------------------------------------------------------------------ */
int gempic_maxwell1d_create(int64_t nx, double xmin, double xmax, int degree,
                            gempic_maxwell1d **out);
int gempic_maxwell1d_destroy(gempic_maxwell1d *m);
/* eigenvalue tables in halfcomplex layout (:94-153); any may be NULL */
int gempic_maxwell1d_tables(gempic_maxwell1d *m, double *eig_mass0, double *eig_mass1,
                            double *eig_weak_ampere, double *eig_weak_poisson);
/* compute_e_from_rho! (:244-255) */
int gempic_maxwell1d_compute_e_from_rho(gempic_maxwell1d *m, double *e, const double *rho);
/* compute_e_from_j! (:263-289), component 1 or 2 */
int gempic_maxwell1d_compute_e_from_j(gempic_maxwell1d *m, double *e, const double *current,
                                      int component);
/* compute_e_from_b! (:384-396) */
int gempic_maxwell1d_compute_e_from_b(gempic_maxwell1d *m, double *e, double dt,
                                      const double *b);
/* compute_b_from_e! (:407-420) */
int gempic_maxwell1d_compute_b_from_e(gempic_maxwell1d *m, double *b, double dt,
                                      const double *e);
/* compute_rhs_from_function! (:188-220): coefs[i] = dx sum over the
 * support of spline i and the degree+1 Gauss-Legendre points of each cell of
 * w f(x) N_i(x); f is called on the host (a C callback, e.g. a Julia
 * @cfunction, with ctx passed through). 0 <= degree <= 3. */
typedef double (*gempic_fn1d)(double x, void *ctx);
int gempic_maxwell1d_compute_rhs_from_function(gempic_maxwell1d *m, gempic_fn1d f, void *ctx,
                                               int degree, double *coefs);
/* l2projection! (:347-374) for degree s_deg_0 or s_deg_0 - 1. As in the
 * reference, the result is the right-hand side over dx (the mass solve lands
 * in the reference's workspace and is not returned). */
int gempic_maxwell1d_l2projection(gempic_maxwell1d *m, gempic_fn1d f, void *ctx, int degree,
                                  double *coefs);

/* inner_product (:461-475) and l2norm_squared (:299-313) */
int gempic_maxwell1d_inner_product(gempic_maxwell1d *m, const double *c1, const double *c2,
                                   int degree, double *out);
int gempic_maxwell1d_l2norm_squared(gempic_maxwell1d *m, const double *c, int degree,
                                    double *out);

/* ------------------------------------------------------------------------
 * ParticleMeshCoupling2D batched verbs (particle_mesh_coupling_2d.jl) over
 * host arrays of n particles. nx/ny/bounds/degree/smoothing as in the
 * constructor (:22-45) on TwoDGrid (mesh.jl:31-43). Grid arrays hold nx*ny
 * doubles, x fastest (node (i, j) at i + j*nx). 1 <= degree <= 3,
 * degree+1 <= nx, ny <= 256.
 * ------------------------------------------------------------------------ */
/* add_charge! (:94-105): rho[nx*ny] += sum_i deposit(x[i], y[i], wp[i]);
 * the device deposit holds nx*ny <= 4096 */
int gempic_pmc2d_add_charge(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                            double ymax, int degree, int smoothing, int64_t n, const double *x,
                            const double *y, const double *wp, double *rho);
/* evaluate (:189-203): out[i] = field(x[i], y[i]) */
int gempic_pmc2d_evaluate(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                          double ymax, int degree, int smoothing, int64_t n, const double *x,
                          const double *y, const double *field, double *out);
/* evaluate_multiple (:214-231): two fields at once */
int gempic_pmc2d_evaluate_multiple(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                                   double ymax, int degree, int smoothing, int64_t n,
                                   const double *x, const double *y, const double *field1,
                                   const double *field2, double *out1, double *out2);

/* add_charge_pp! (:127-148): floor()+1 index, SplinePP (Horner) weights */
int gempic_pmc2d_add_charge_pp(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                               double ymax, int degree, int smoothing, int64_t n, const double *x,
                               const double *y, const double *wp, double *rho);
/* evaluate_pp (:162-178) on the ((degree+1)^2) x (nx*ny) column-major pp
 * matrix of b_to_pp_2d; a position outside the table is a BoundsError */
int gempic_pmc2d_evaluate_pp(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                             double ymax, int degree, int smoothing, int64_t n, const double *x,
                             const double *y, const double *pp, double *out);
/* b_to_pp_2d! (splinepp.jl:329-399): pp[(degree+1)^2 x nx*ny] from b[nx*ny] */
int gempic_b_to_pp_2d(int degree, int64_t nx, int64_t ny, const double *b, double *pp);

/* Device-resident forms of the three verbs above, for particles kept in GPU
 * memory (a ROCArray on the Julia side): x, y, wp, rho, field*, out* are
 * device pointers and `stream` a hipStream_t (NULL: the null stream).
 * flags: GEMPIC_DETERMINISTIC makes the deposit bit-reproducible from run
 * to run (per-lane patch caches added in lane order); without it the deposit
 * adds into a per-workgroup LDS grid with LDS atomics, and only the order of
 * those additions varies. GEMPIC_ASYNC returns once the work is queued; an
 * InexactError is then reported by the next synchronous call or by
 * gempic_pmc2d_sync. The calls share one workspace per device: calls that
 * overlap on one device race. */
#define GEMPIC_DETERMINISTIC 32
int gempic_pmc2d_add_charge_dev(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                                double ymax, int degree, int smoothing, int64_t n,
                                const double *x, const double *y, const double *wp, double *rho,
                                int flags, void *stream);
int gempic_pmc2d_evaluate_dev(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                              double ymax, int degree, int smoothing, int64_t n, const double *x,
                              const double *y, const double *field, double *out, int flags,
                              void *stream);
int gempic_pmc2d_evaluate_multiple_dev(int64_t nx, int64_t ny, double xmin, double xmax,
                                       double ymin, double ymax, int degree, int smoothing,
                                       int64_t n, const double *x, const double *y,
                                       const double *field1, const double *field2, double *out1,
                                       double *out2, int flags, void *stream);
/* wait for `stream` and report errors of earlier GEMPIC_ASYNC 2D calls */
int gempic_pmc2d_sync(void *stream);

/* Allocation-free add_charge! on device arrays. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Arguments, flags and results are those of
 * gempic_pmc2d_add_charge_dev; the call makes no hipMalloc/hipFree. Its
 * error flag, workgroup grids and split sums live in a fixed 32 MiB array of
 * device memory that every GPU the library runs on holds until the process
 * ends: the deposit uses at most (32 MiB - 256 B) / (8 nx ny) grids, so its
 * workgroup count is capped by the grid size (nx*ny <= 4096, as for the
 * other deposit). All calls on one GPU share that array and keep the last
 * call's contents: calls that overlap on one GPU race. With GEMPIC_ASYNC an
 * InexactError is reported by the next synchronous call of this entry point
 * on the same GPU. */
int gempic_pmc2d_add_charge_dev_noalloc(int64_t nx, int64_t ny, double xmin, double xmax,
                                        double ymin, double ymax, int degree, int smoothing,
                                        int64_t n, const double *x, const double *y,
                                        const double *wp, double *rho, int flags, void *stream);

/* ------------------------------------------------------------------------
 * TwoDMaxwell (maxwell_2d_fem.jl:11-88) with its TwoDPoisson
 * (poisson_2d_fem.jl:3-85) and TwoDLinearSolverSplineMass
 * (linear_solver_spline_mass_2d.jl) on host arrays of nx*ny doubles, run on
 * the GPU. 1 <= degree <= 3, 2*degree+1 <= nx, ny <= 256.
 * ------------------------------------------------------------------------ */
int gempic_maxwell2d_create(int64_t nx, double xmin, double xmax, int64_t ny, double ymin,
                            double ymax, int degree, gempic_maxwell2d **out);
int gempic_maxwell2d_destroy(gempic_maxwell2d *m);
/* eig_values_mass_{0,1}_{1,2} (:51-54); any may be NULL */
int gempic_maxwell2d_tables(gempic_maxwell2d *m, double *eig0x, double *eig0y, double *eig1x,
                            double *eig1y);
/* solve(inv_mass_1[component] or inv_mass_2[component], rhs) (:56-66,
 * linear_solver_spline_mass_2d.jl:16-32); form 1 or 2 */
int gempic_maxwell2d_solve(gempic_maxwell2d *m, int form, int component, const double *rhs,
                           double *out);
/* multiply_mass_2dkron! (:343-363) with the mass lines of inner_product's
 * (form, component) (:511-575) */
int gempic_maxwell2d_multiply_mass(gempic_maxwell2d *m, int form, int component, const double *in,
                                   double *out);
/* inner_product (:511-575) */
int gempic_maxwell2d_inner_product(gempic_maxwell2d *m, const double *c1, const double *c2,
                                   int component, int form, double *out);
/* compute_e_from_b! (:370-412): e1..e3 updated in place */
int gempic_maxwell2d_compute_e_from_b(gempic_maxwell2d *m, double dt, const double *b1,
                                      const double *b2, const double *b3, double *e1, double *e2,
                                      double *e3);
/* compute_b_from_e! (:423-444): b1..b3 updated in place */
int gempic_maxwell2d_compute_b_from_e(gempic_maxwell2d *m, double dt, const double *e1,
                                      const double *e2, const double *e3, double *b1, double *b2,
                                      double *b3);
/* compute_e_from_j! (:455-459), component 1..3 */
int gempic_maxwell2d_compute_e_from_j(gempic_maxwell2d *m, const double *current, int component,
                                      double *e);
/* compute_rho_from_e! (:468-499); e3 is not read and may be NULL */
int gempic_maxwell2d_compute_rho_from_e(gempic_maxwell2d *m, const double *e1, const double *e2,
                                        const double *e3, double *rho);
/* compute_e_from_rho! (:199-201 -> poisson_2d_fem.jl:237-262) */
int gempic_maxwell2d_compute_e_from_rho(gempic_maxwell2d *m, const double *rho, double *e1,
                                        double *e2);

/* ------------------------------------------------------------------------
 * The 2D3V Hamiltonian splitting. The reference includes
 * hamiltonian_splitting_2d3v.jl (src/GEMPIC.jl:33) but the file is empty;
 * this is the GEMPIC paper's splitting H = Hp1 + Hp2 + Hp3 + HE + HB on the
 * reference's 2D pieces, in the form of its 1D2V splitting
 * (hamiltonian_splitting.jl:20-108, hamiltonian_splitting_1d2v.jl):
 * ParticleGroup{2,3} (particle_group.jl:15-46; rows x1, x2, v1, v2, v3, w,
 * ld = 6), the 2D coupling's shape factor (particle_mesh_coupling_2d.jl:54-82)
 * with one degree per direction, and TwoDMaxwell(mesh, degree)
 * (maxwell_2d_fem.jl:26-87) with E1 in splines of degrees (p-1, p), E2
 * (p, p-1), E3 (p, p), B1 (p, p-1), B2 (p-1, p), B3 (p-1, p-1):
 *   operatorHp1/Hp2: the position moves along x1 / x2 with the exact line
 *     integrals of the degree p-1 splines (those of update_jv!,
 *     particle_mesh_coupling_1d.jl:385-414, here from the splines'
 *     antiderivative), the current j1 / j2 and the magnetic kicks of the
 *     other velocities, then E -= M1^{-1} j (compute_e_from_j!,
 *     maxwell_2d_fem.jl:455-459);
 *   operatorHp3: the kicks of v1, v2 by v3 B and the current j3 = dt q v3 S;
 *   operatorHE: v += dt qm E(x), then compute_b_from_e! (:423-444);
 *   operatorHB: compute_e_from_b! (:370-412);
 *   strang_splitting: HB HE Hp3 Hp2 Hp1 Hp2 Hp3 HE HB (half steps around Hp1).
 * Particles are kept sorted by cell on the device (every `interval` steps,
 * default 20); get_particles restores the caller's order. Degrees 1..3,
 * 2*degree+1 <= nx, ny and nx*ny <= 4096. The current deposits add in an
 * order that depends on timing between wave flushes: results agree with the
 * oracle to rounding, not bitwise.
 * ------------------------------------------------------------------------ */
typedef struct gempic_splitting2d gempic_splitting2d;
int gempic_splitting2d_create(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                              double ymax, int degree, int64_t n_particles, double charge,
                              double mass, double common_weight, gempic_splitting2d **out);
int gempic_splitting2d_destroy(gempic_splitting2d *h);
int gempic_splitting2d_set_particles(gempic_splitting2d *h, const double *aos, int64_t ld, int64_t n);
int gempic_splitting2d_get_particles(gempic_splitting2d *h, double *aos, int64_t ld, int64_t n);
/* e_dofs and b_dofs, nx*ny each (x fastest); any pointer may be NULL */
int gempic_splitting2d_set_fields(gempic_splitting2d *h, const double *e1, const double *e2,
                                  const double *e3, const double *b1, const double *b2,
                                  const double *b3);
int gempic_splitting2d_get_fields(gempic_splitting2d *h, double *e1, double *e2, double *e3,
                                  double *b1, double *b2, double *b3);
/* j_dofs as left by the last Hp1 / Hp2 / Hp3 (j3 includes its dt) */
int gempic_splitting2d_get_j(gempic_splitting2d *h, double *j1, double *j2, double *j3);
int gempic_splitting2d_operator_hb(gempic_splitting2d *h, double dt);
int gempic_splitting2d_operator_he(gempic_splitting2d *h, double dt);
int gempic_splitting2d_operator_hp1(gempic_splitting2d *h, double dt);
int gempic_splitting2d_operator_hp2(gempic_splitting2d *h, double dt);
int gempic_splitting2d_operator_hp3(gempic_splitting2d *h, double dt);
/* flags: GEMPIC_ASYNC returns once the work is queued */
int gempic_splitting2d_strang_splitting(gempic_splitting2d *h, double dt, int64_t n_steps, int flags);
int gempic_splitting2d_synchronize(gempic_splitting2d *h);
/* sort the particles by cell every `interval` Strang steps (0, the default:
   never; the passes run faster on unordered particles) */
int gempic_splitting2d_set_sort_interval(gempic_splitting2d *h, int64_t interval);
/* add_charge! (particle_mesh_coupling_2d.jl:94-105) of the handle's particles
 * with degree (p, p) and q = charge w common_weight, into rho (nx*ny, not
 * accumulated), and compute_rho_from_e! (maxwell_2d_fem.jl:468-499) of its
 * fields: their difference is Gauss's law, constant over the splitting. */
int gempic_splitting2d_rho(gempic_splitting2d *h, double *rho);
int gempic_splitting2d_rho_from_e(gempic_splitting2d *h, double *rho);

#ifdef __cplusplus
}
#endif
#endif
