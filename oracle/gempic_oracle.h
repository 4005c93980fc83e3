/*
 * gempic_oracle.h -- CPU restatement of GEMPIC.jl's particle-field path.
 *
 * TEST INFRASTRUCTURE. This library is the checker, not the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (gempic.jl_amd/csrc, include/gempic_hip.h) never links it.
 *
 * Every function follows one function of the reference (Julia, src/...) line
 * by line, including its rounding order; the file is compiled with
 * -ffp-contract=off so that no product/sum pair is fused into an FMA (Julia
 * does not contract by default). Citations are file:line in the reference.
 *
 * How each function is pinned (tests/test_oracle_kat.py, KAT = the
 * reference's own known-answer tests, copied into tests/golden/;
 * tests/test_oracle_invariants.py for the rest):
 *   - bspline basis, add_charge(_pp), evaluate(_pp), add_current_update_v(_pp):
 *     KATs of test/test_particle_mesh_coupling_spline_1d.jl, partition of
 *     unity, charge continuity of the line integral;
 *   - add_current_update_v_nob (1d1v): charge continuity for every move;
 *   - gausslegendre: polynomial exactness; splinepp tables, b_to_pp: de Boor;
 *   - Maxwell1DFEM tables and solves: eigenvalues = DFT of the mass lines,
 *     dense circulant solve, numpy rfft, the analytic tests of
 *     test/test_maxwell_1d_fem.jl, the splitting KAT's fields;
 *     l2norm_squared / inner_product: Gauss quadrature of the spline;
 *   - 1d2v operators: KAT of test/test_hamiltonian_splitting.jl (after each
 *     sub-step); 1d1v operators: Gauss's law kept by Hp1, HB's kick;
 *   - strang_splitting: second-order energy conservation;
 *   - write_step: energy conservation and the diagnostics as rates of change
 *     of the energies and momenta under HE, Hp2 and HB; ErrorPoisson by
 *     Gauss's law;
 *   - reverse mode (strang_adjoint, solve_poisson_adjoint, bspline_deriv,
 *     strang_adjoint_domain): central differences of the forward functions
 *     above along random directions of every input, and in the wavenumber k
 *     that moves the domain (tests/test_oracle_adjoint.py).
 * No function here is left unpinned.
 *
 * Layout conventions (same as the reference):
 *   - particle array: column-major (D+V+n_weights) x N, i.e. element (r, i)
 *     (both 0-based here) lives at a[r + i*ld] with ld = D+V+n_weights;
 *   - grid arrays: nx doubles, node k (0-based) = Julia index k+1.
 */
#ifndef GEMPIC_ORACLE_H
#define GEMPIC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_GALERKIN 0
#define OR_COLLOCATION 1
#define OR_MAXDEG 5

/* ---------- Julia Base semantics used on the path ---------- */
double or_jl_mod_f64(double x, double y);
int64_t or_jl_mod_i64(int64_t a, int64_t n);

/* ---------- low_level_bsplines.jl ---------- */
void or_bspline_basis(int degree, double offset, double *bspl);
void or_eval_uniform_periodic_spline_curve(int degree, int64_t n, const double *scoef, double *sval);

/* ---------- FastGaussQuadrature.gausslegendre (third-party, restated) ---------- */
int or_gausslegendre(int n, double *x, double *w);

/* ---------- splinepp.jl ---------- */
int or_splinepp_tables(int degree, double *poly_coeffs, double *poly_coeffs_fp);
void or_b_to_pp(int degree, int64_t ncells, const double *b_coeffs, double *pp_coeffs);

/* ---------- particle_mesh_coupling_1d.jl ---------- */
typedef struct {
    int64_t n_grid;
    double xmin, Lx, delta_x, scaling;
    int degree, n_span, n_quad;
    double quad_x[8], quad_w[8];
    double pp[(OR_MAXDEG + 1) * (OR_MAXDEG + 1)];    /* [power + piece*(deg+1)] */
    double pp_fp[(OR_MAXDEG + 1) * (OR_MAXDEG + 1)];
} or_pmc1d;

int or_pmc1d_init(or_pmc1d *p, double xmin, double xmax, int64_t nx, int degree, int smoothing);
void or_add_charge(double *rho, const or_pmc1d *p, double x, double q);
void or_add_charge_pp(double *rho, const or_pmc1d *p, double x, double q);
double or_evaluate(const or_pmc1d *p, double x, const double *field);
double or_evaluate_pp(const or_pmc1d *p, double x, const double *field_pp);
double or_add_current_update_v(double *j, const or_pmc1d *p, double x_old, double x_new,
                               double q, double qm, const double *b, double v);
double or_add_current_update_v_nob(double *j, const or_pmc1d *p, double x_old, double x_new,
                                   double q, double qm, double v);
double or_add_current_update_v_pp(double *j, const or_pmc1d *p, double x_old, double x_new,
                                  double q, double qm, const double *b, double v);

/* batched helpers (loops over particles in index order, as the reference's callers do) */
void or_add_charge_n(double *rho, const or_pmc1d *p, int64_t n, const double *x, const double *q);
void or_evaluate_n(const or_pmc1d *p, int64_t n, const double *x, const double *field, double *out);

/* ---------- maxwell_1d_fem.jl ---------- */
typedef struct {
    double xmin, Lx, delta_x;
    int64_t n;
    int s_deg_0, s_deg_1;
    double mass_0[4], mass_1[4];
    double *eig_mass0, *eig_mass1, *eig_weak_ampere, *eig_weak_poisson;
    double *work, *wsave, *eigvals;
    double *cos_tab, *sin_tab;   /* DFT twiddles, index (j*k mod n) */
} or_maxwell1d;

or_maxwell1d *or_maxwell1d_create(double xmin, double xmax, int64_t n, int degree);
void or_maxwell1d_destroy(or_maxwell1d *m);
void or_maxwell1d_tables(const or_maxwell1d *m, double *eig_mass0, double *eig_mass1,
                         double *eig_weak_ampere, double *eig_weak_poisson);
void or_r2hc(const or_maxwell1d *m, const double *in, double *out);
void or_hc2r(const or_maxwell1d *m, const double *in, double *out);
void or_solve_circulant(or_maxwell1d *m, const double *eigvals, const double *rhs);
void or_compute_e_from_rho(double *e, or_maxwell1d *m, const double *rho);
void or_compute_e_from_j(double *e, or_maxwell1d *m, const double *current, int component);
void or_compute_e_from_b(double *e, or_maxwell1d *m, double dt, const double *b);
void or_compute_b_from_e(double *b, or_maxwell1d *m, double dt, const double *e);
/* func(x) = a_cos*cos(k*x) + a_sin*sin(k*x) + c */
void or_compute_rhs_from_function(double *coefs, or_maxwell1d *m, double a_cos, double a_sin,
                                  double k, double c, int degree);
int or_l2projection(double *coefs, or_maxwell1d *m, double a_cos, double a_sin, double k,
                    double c, int degree);
double or_l2norm_squared(or_maxwell1d *m, const double *coefs, int degree);
double or_inner_product(or_maxwell1d *m, const double *c1, const double *c2, int degree);
const double *or_maxwell1d_work(const or_maxwell1d *m);

/* ---------- particle_mesh_coupling_2d.jl (gempic_oracle2d.c) ---------- */
typedef struct {
    int64_t nx, ny;
    double xmin, ymin, dx, dy, scaling;
    int degree, n_span;
    double pp[(OR_MAXDEG + 1) * (OR_MAXDEG + 1)];
    double pp_fp[(OR_MAXDEG + 1) * (OR_MAXDEG + 1)];
} or_pmc2d;

int or_pmc2d_init(or_pmc2d *p, double xmin, double xmax, int64_t nx, double ymin, double ymax,
                  int64_t ny, int degree, int smoothing);
void or_pmc2d_add_charge(double *rho, const or_pmc2d *p, double x, double y, double wp);
void or_pmc2d_add_charge_pp(double *rho, const or_pmc2d *p, double x, double y, double wp);
double or_pmc2d_evaluate(const or_pmc2d *p, double x, double y, const double *field);
void or_pmc2d_evaluate_multiple(const or_pmc2d *p, double x, double y, const double *f1,
                                const double *f2, double *out2);
void or_pmc2d_add_charge_n(double *rho, const or_pmc2d *p, int64_t n, const double *x,
                           const double *y, const double *wp);
void or_pmc2d_evaluate_n(const or_pmc2d *p, int64_t n, const double *x, const double *y,
                         const double *field, double *out);
void or_pmc2d_evaluate_multiple_n(const or_pmc2d *p, int64_t n, const double *x,
                                  const double *y, const double *f1, const double *f2,
                                  double *out1, double *out2);
void or_pmc2d_b_to_pp(const or_pmc2d *p, const double *b, double *pp);
double or_pmc2d_evaluate_pp(const or_pmc2d *p, double x, double y, const double *pp,
                            int *bounds_error);

/* ---------- maxwell_2d_fem.jl, poisson_2d_fem.jl, linear_solver_spline_mass_2d.jl ---------- */
typedef struct or_maxwell2d or_maxwell2d;
void or_spline_fem_mass_line(int degree, double *mass_line);
void or_spline_fem_compute_mass_eig(int64_t n_cells, int degree, const double *mass_line,
                                    double *eig);
or_maxwell2d *or_maxwell2d_create(double xmin, double xmax, int64_t nx, double ymin, double ymax,
                                  int64_t ny, int degree);
void or_maxwell2d_destroy(or_maxwell2d *m);
const double *or_maxwell2d_table(const or_maxwell2d *m, int which);
void or_maxwell2d_solve(or_maxwell2d *m, int form, int component, const double *rhs, double *out);
void or_maxwell2d_multiply_mass(or_maxwell2d *m, const double *line1, int deg1,
                                const double *line2, int deg2, const double *in, double *out);
double or_maxwell2d_inner_product(or_maxwell2d *m, const double *c1, const double *c2,
                                  int component, int form);
void or_maxwell2d_compute_e_from_b(or_maxwell2d *m, double dt, double *const *b, double *const *e);
void or_maxwell2d_compute_b_from_e(or_maxwell2d *m, double dt, double *const *e, double *const *b);
void or_maxwell2d_compute_e_from_j(or_maxwell2d *m, const double *current, int component,
                                   double *e);
void or_maxwell2d_compute_rho_from_e(or_maxwell2d *m, double *const *e, double *rho);
void or_maxwell2d_compute_e_from_rho(or_maxwell2d *m, const double *rho, double *const *e);
void or_maxwell2d_fem_rhs(or_maxwell2d *m, double (*f)(double, double, void *), void *ctx,
                          int component, int form, double *coefs);
void or_maxwell2d_l2projection(or_maxwell2d *m, double (*f)(double, double, void *), void *ctx,
                               int component, int form, double *out);

/* ---------- 2D3V Hamiltonian splitting (gempic_oracle2d3v.c; the reference's
 * hamiltonian_splitting_2d3v.jl is empty) ---------- */
typedef struct or_split2d or_split2d;
/* particles: 6 x n column-major (x1, x2, v1, v2, v3, w), e and b: 3 arrays of
 * nx*ny each; all ALIASED */
or_split2d *or_split2d_create(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                              double ymax, int degree, int64_t n, double charge, double mass,
                              double common_weight, double *particles, double *const *e,
                              double *const *b);
void or_split2d_destroy(or_split2d *h);
const double *or_split2d_j(const or_split2d *h, int component);
void or_split2d_operator_hb(or_split2d *h, double dt);
void or_split2d_operator_he(or_split2d *h, double dt);
void or_split2d_operator_hp1(or_split2d *h, double dt);
void or_split2d_operator_hp2(or_split2d *h, double dt);
void or_split2d_operator_hp3(or_split2d *h, double dt);
void or_split2d_strang(or_split2d *h, double dt, int64_t nsteps);
void or_split2d_rho(or_split2d *h, double *rho);
void or_split2d_rho_from_e(or_split2d *h, double *rho);
void or_split2d_energy(or_split2d *h, double *out7);

/* ---------- hamiltonian_splitting*.jl ---------- */
typedef struct or_split or_split;

/* particles/e1/e2/b are ALIASED (caller-owned), as HamiltonianSplitting aliases
 * the caller's e_dofs/b_dofs and the ParticleGroup array. */
or_split *or_split_create(int64_t nx, double xmin, double xmax, int maxwell_degree,
                          int degree0, int smoothing0, int degree1, int smoothing1,
                          int dim_v, int64_t n_particles, int n_weights, double charge,
                          double mass, double common_weight, double *particles, double *e1,
                          double *e2, double *b);
void or_split_destroy(or_split *h);

/* sample!(pg, alpha, k, sigma, mesh) (particle_sampling.jl:236-311) and the
 * Sobol.jl SobolSeq(2) points it draws */
void or_sobol2(uint64_t k, double *r1, double *r2);
int or_landau_newton(double r, double alpha, double k, double *out);
int or_sample_landau(int dim_v, int64_t n, double alpha, double k, double sigma, double dimx,
                     double *aos, int64_t ld);

/* HamiltonianSplittingBoris (hamiltonian_splitting_boris.jl) on an or_split
 * state with dim_v = 2. which: 0 = e_dofs_mid[1], 1 = e_dofs_mid[2], 2 =
 * b_dofs_mid. */
const double *or_boris_mid(const or_split *h, int which);
void or_boris_set_mid(or_split *h, const double *e1m, const double *e2m, const double *bm);
void or_boris_push_v_epart(or_split *h, double dt);
void or_boris_push_v_bpart(or_split *h, double dt);
void or_boris_push_x_accumulate_j(or_split *h, double dt);
void or_boris_staggering(or_split *h, double dt);
void or_boris_strang(or_split *h, double dt, int64_t number_steps);
const double *or_split_j(const or_split *h, int component);
void or_operator_hb(or_split *h, double dt);
void or_operator_he(or_split *h, double dt);
void or_operator_hp1(or_split *h, double dt);
void or_operator_hp2(or_split *h, double dt);
void or_strang_splitting(or_split *h, double dt, int64_t nsteps);
int or_split_set_particles(or_split *h, double *particles, int64_t n_particles);
/* n_jobs = nthreads() of the reference (hamiltonian_splitting.jl:61-66): the
 * 1d2v Hp1/Hp2 deposits per chunk of n_particles / ntasks particles, on
 * ntasks threads, summed in chunk order; HE on ntasks threads. Default 1.
 * Returns -1 when n_particles % ntasks != 0 (the reference's @assert). */
int or_split_set_tasks(or_split *h, int ntasks);

/* ---------- diagnostics.jl ---------- */
void or_solve_poisson(double *efield, or_split *h, double *rho);
/* out[0..10] = Time, KineticEnergy, Momentum1, Momentum2, PotentialEnergyE1,
 * PotentialEnergyE2, PotentialEnergyB3, Transfer, VVB, Poynting, ErrorPoisson */
void or_write_step(or_split *h, double time, int degree, const double *e1, const double *e2,
                   const double *b, const double *e1_n, const double *e2_n,
                   const double *efield_poisson, double *out);

/* ---------- reverse mode of the 1D2V Strang splitting (no reference
 * counterpart; pinned by central finite differences of the functions above) ---------- */
void or_bspline_deriv(int degree, double offset, double *d);
double or_strang_loss(or_split *h, double dt, int64_t nsteps, double wloss);
int or_strang_adjoint(or_split *h, double dt, int64_t nsteps, double wloss,
                      const double *ct_part, const double *ct_e1, const double *ct_e2,
                      const double *ct_b, double *loss, double *g_part, double *g_e1,
                      double *g_e2, double *g_b);
void or_solve_poisson_adjoint(or_split *h, const double *ct_e, double *g_x);
void or_strang_adjoint_domain(const or_split *h, double *g_length, double *g_wsum);

#ifdef __cplusplus
}
#endif
#endif
