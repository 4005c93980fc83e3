// lib_internal.hpp -- declarations shared by the kernels (pass1d.hip), the
// C ABI (capi.hip) and the 2D coupling and field solvers (pic2d.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gl {

constexpr int kBlock = 256;               // threads per workgroup of the particle passes
constexpr int kWaves = kBlock / 64;
// iterations ahead that a 1D particle pass loads its particles (1 or 2)
#ifndef GEMPIC_PASS_PREFETCH
#define GEMPIC_PASS_PREFETCH 1
#endif
constexpr int kPrefetch = GEMPIC_PASS_PREFETCH;
static_assert(kPrefetch == 1 || kPrefetch == 2, "prefetch depth 1 or 2");
constexpr int kNxMax = 256;               // largest 1D grid handled by the device path
constexpr int kGroup = 16;                // workgroups per first-level reduction group
constexpr int kEpiGroupsMax = 64;         // workgroups of the deposit-sum kernel
constexpr int kAdjNxMax = 128;            // largest 1D grid of the reverse-mode kernels

// grid slots (nx doubles each) of a splitting's device field buffer
// (S_E1M, S_E2M, S_BM: the Boris splitting's e_dofs_mid and b_dofs_mid)
enum GridSlot { S_E1 = 0, S_E2, S_B, S_E2P, S_J1, S_J2, S_RHO, S_EPOIS, S_E1N, S_E2N, S_E1M, S_E2M,
                S_BM, NSLOT };
// constant tables (nx doubles each)
// T_CM0, T_CM1, T_CAMP: first columns of the circulant matrices of the M0 and
// M1 inverse-mass solves and of the weak Ampere solve (the fused-arithmetic
// field part applies them as matrix-vector products)
enum TabSlot { T_COS = 0, T_SIN, T_INVM0, T_INVM1, T_AMP, T_POIS, T_M0, T_M1, T_CM0, T_CM1, T_CAMP,
               T_CMASS1, T_CPOIS, NTAB };
// deposit slots of one pass
enum DepSlot { D_J2A = 0, D_J1, D_RHO, D_J2B, NDEP };

// Sub-step content of one particle pass.
enum ModeBits : int {
    M_HE_PREV = 1,   // HE (or 1d1v HB) of the previous Strang step, with E2 saved before it
    M_HE = 2,        // HE (1d2v) / HB (1d1v) with the current fields
    M_HP2A = 4,      // Hp2 before Hp1
    M_HP1 = 8,       // Hp1 (+ rho deposit in 1d2v when no Hp2 follows in the pass)
    M_HP2B = 16,     // Hp2 after Hp1
    M_V1 = 32,       // 1d1v variant
    M_RHO = 64,      // charge deposit only (solve_poisson!)
    M_DIAG = 128,    // write_step! particle sums
    M_NOEPI = 256,   // write the deposit rows and stop (the next pass reduces them)
    M_WITH_J2A = 512,  // also reduce the Hp2 rows written by the previous pass
    M_FAST = 1024,   // fused arithmetic: closed-form spline bases, fma, the line integral
                     // by its midpoint rule with the curvature term (agrees with the
                     // reference to rounding, not bitwise)
};

constexpr int kMaxGen = 8;        // generations of a particle pass partition (PassArgs)

struct PassArgs {
    int64_t n;
    int64_t chunk;       // particles per workgroup (multiple of kBlock)
    int iters;           // chunk / kBlock
    int nx;
    int ngroups;         // number of reduction groups
    int more;            // fused pass: another Strang step follows in this call
    int degree;          // write_step! degree argument (diagnostics)
    double xmin, dx, Lx;
    double scaling0, scaling1;
    double qx[3], qw[3];          // Gauss-Legendre rule of kernel_smoother_1
    double charge, mass, cw, qm;
    double dt_he, dt_hp2, dt_hp1;
    double dtqm_he, dtqm_hp2;     // dt_he*qm and dt_hp2*qm, rounded as in `dt * qm * ...`
    double *x, *v1, *v2, *w;
    double *grid;                 // NSLOT x nx
    const double *tabs;           // NTAB x nx
    double *part, *gpart;         // reduction partials
    unsigned *cnt;                // ngroups + 1 counters (zero between launches)
    int *err;                     // error flags
    double *diag;                 // diagnostics output (11 doubles)
    double *hist;                 // fused Strang step with a record: k_epi1d writes B at the
                                  // step's end to hist[4][nx] and the next pass's fields
                                  // E1, E2, B, E2P to hist[5..8][nx] (nullptr: no record)
    double *cx, *cv1, *cv2;       // fused Strang pass: also store x, v1, v2 here (nullptr: no)
    // Particle pass partition in generations (gen_blocks > 0): workgroups
    // g * gen_blocks .. (g + 1) * gen_blocks - 1 form generation g and take
    // gen_chunk[g] particles each (a multiple of 64), from gen_base[g] on;
    // gen_blocks = 0 is the uniform partition by chunk.
    int gen_blocks;
    int64_t gen_chunk[kMaxGen];
    int64_t gen_base[kMaxGen];
};

// error flag bits set by kernels
enum ErrBits { E_INEXACT = 1, E_BOUNDS = 2, E_FAR = 4 };
// largest number of cells a particle may cross in one Hp1 push; beyond it
// the move is reported as an error instead of being integrated
constexpr int kMaxCellsPerStep = 1 << 16;

hipError_t launch_pass(int p0, int p1, int mode, const PassArgs &a, int blocks, size_t shmem,
                       hipStream_t s);

// An ensemble's instances in one launch (device arrays): PassArgs per
// instance, the first particle-pass workgroup of each instance (ninst + 1
// entries) and the first field-part workgroup of each (ninst + 1 entries).
struct BatchArgs {
    const PassArgs *args;
    const int *block_start;
    const int *epi_start;
    int ninst;
};
hipError_t launch_pass_batch(int p0, int p1, int mode, const BatchArgs &b, int total_blocks,
                             size_t shmem, hipStream_t s);
hipError_t launch_epi_batch(int p0, int p1, int mode, const BatchArgs &b, int total_groups,
                            size_t shmem, hipStream_t s);
size_t epi_shmem_bytes(int nx);
int epi_groups(int nblocks);      // workgroups of the field part after a pass of nblocks
// The deposit sum and field solves that follow a deposit pass (k_epi1d);
// pass_blocks = workgroups of that pass. Modes with M_NOEPI have none.
bool mode_has_epi(int mode);
hipError_t launch_epi(int p0, int p1, int mode, const PassArgs &a, int pass_blocks, hipStream_t s);
// workgroups of the (p0, p1, mode) pass kernel resident on one CU
int pass_blocks_per_cu(int p0, int p1, int mode, size_t shmem);
// LDS of a particle pass of this mode: padded fields, deposit rows (reused as
// reduction scratch and tables) and, for deposit modes, the cache-B slots.
size_t pass_shmem_bytes(int nx, int mode);
constexpr int kSlotsPerLane = 12;         // cache-B doubles per lane: three caches of <= 4 nodes

// Single-workgroup field program on the splitting's grid buffer.
enum GridOp : int {
    G_HB = 1,            // E2 += dt/dx * W(B)       (compute_e_from_b!)
    G_HE_B,              // B  += dt/dx * (E2[i-1] - E2[i])  (compute_b_from_e!)
    G_E_FROM_J1,         // E1 -= M1^{-1} J1 / dx
    G_E_FROM_J2,         // E2 -= M0^{-1} J2 / dx
    G_E_FROM_RHO,        // EPOIS from RHO           (compute_e_from_rho!)
    G_INNER_M0,          // scalar: sum(c1 .* M0 c2) dx
    G_INNER_M1,
    G_COPY,              // out = A
    G_AVG,               // out = (out + A) * 0.5
};
struct GridProg {
    int nops;
    int op[8];
    double dt[8];
    int slot_a[8], slot_b[8], slot_out[8];   // generic operands for the Maxwell1DFEM verbs
};
hipError_t launch_grid(const GridProg &g, int nx, double dx, double *grid, const double *tabs,
                       double *scalar_out, hipStream_t s);

// HamiltonianSplittingBoris particle pass: which pushers it runs, in this
// order (E kick with dtqm_he, B rotation with dtqm_hp2 = qm*0.5*dt, E kick,
// x push with dt_hp1 and the midpoint deposits of j1 and j2)
enum BorisOps : int { BO_EPART = 1, BO_BPART = 2, BO_EPART2 = 4, BO_PUSHX = 8 };
// field part after a BO_PUSHX pass: j_dofs only (push_x_accumulate_j!), the
// staggering!, or a strang_splitting! step (a.more: the next step's B
// update follows)
enum BorisEpi : int { BE_J = 0, BE_STAGGER, BE_STEP };
hipError_t launch_boris(int p0, int p1, int ops, const PassArgs &a, int blocks, size_t shmem,
                        hipStream_t s);
hipError_t launch_boris_epi(int kind, const PassArgs &a, int pass_blocks, hipStream_t s);
size_t boris_shmem_bytes(int nx);
int boris_blocks_per_cu(int p0, int p1, size_t shmem);

// ---------------------------------------------------------------------------
// Reverse mode of the fused 1D2V Strang pass (k_pass1d<P0,P1,HE_PREV|HE|HP2A|HP1|HP2B>).
// ---------------------------------------------------------------------------
// What a reverse-mode particle kernel does:
//   R_FWD     re-run the pass forward (no deposits) from x, v1, v2 into xo, v1o, v2o
//   R_REV     cotangents through the pass: (lx, lv1, lv2) after it -> before it,
//             deposit rows A_E1 (S1 basis at x), A_E2 (S0 basis at x), A_B (S1)
//   R_HEONLY  cotangents through the trailing HE of a call
//   R_POIS    lx = d <ct, compute_e_from_rho!(rho(x))> / dx with nu = C_pois D^T ct
//   R_PREV    (with R_FWD / R_REV) the pass starts with the previous step's HE
enum AdjMode : int { R_PREV = 1, R_FWD = 2, R_REV = 4, R_HEONLY = 8, R_POIS = 16 };
// adjoint grid slots (nx doubles each) and deposit rows of a reverse pass
enum AdjSlot { A_E1 = 0, A_E2, A_B, A_MU1, A_MU2, A_NU, A_LOSS, NASLOT };
enum AdjRow { AR_E1 = 0, AR_E2, AR_B, AR_SCAL };   // AR_SCAL: [0] d/d dx, [1] d/d Lx (mod), [2] sum d/d w
// scalars in the A_LOSS slot: the loss, then its explicit derivatives
enum AdjScalar { AS_LOSS = 0, AS_DX, AS_MOD, AS_W };
// field part of a reverse step: the reverse of the grid work between the
// passes (k_epi1d and the prologue), see k_adj_epi
enum AdjEpi : int { AE_TAIL = 0, AE_STEP, AE_HEAD, AE_POIS };
// field record of one step (nx doubles each): the fields the step's pass
// reads, then B at the step's end
enum HistSlot { H_E1 = 0, H_E2, H_B, H_E2P, H_BEND, NHIST };

struct AdjArgs {
    PassArgs p;                   // geometry, constants, w, reduction scratch, error flags
    const double *fields;         // this pass's record: [NHIST][nx]
    const double *bend;           // the field part's B at a step's end (nx), or nullptr
    double *agrid;                // [NASLOT][nx]
    const double *x, *v1, *v2;    // primal state before the pass
    double *xo, *v1o, *v2o;       // R_FWD: state after the pass
    double *lx, *lv1, *lv2;       // cotangents, updated in place
    double wloss;                 // weight of 1/2 l2norm_squared(B) per step
};
hipError_t launch_adj(int p0, int p1, int rm, const AdjArgs &a, int blocks, hipStream_t s);
// an ensemble's instances in one launch (device arrays as BatchArgs'); the
// field part runs as two launches (slice sums, then one workgroup per instance)
struct AdjBatch {
    const AdjArgs *args;
    const int *block_start;
    const int *epi_start;
    int ninst;
    int max_nx;
};
hipError_t launch_adj_batch(int p0, int p1, int rm, const AdjBatch &b, int total_blocks,
                            hipStream_t s);
hipError_t launch_adj_epi_batch(int kind, const AdjBatch &b, int total_groups, hipStream_t s);
hipError_t launch_adj_epi(int kind, const AdjArgs &a, int pass_blocks, hipStream_t s);
int adj_blocks_per_cu(int p0, int p1, int rm);
// AoS rows 0..2 of the caller's array (particle id[i] of the caller) into SoA, and back
hipError_t launch_adj_rows_in(const double *aos, int64_t ld, int64_t n, const int64_t *id,
                              double *r0, double *r1, double *r2, hipStream_t s);
hipError_t launch_adj_rows_out(double *aos, int64_t ld, int64_t n, const int64_t *id,
                               const double *r0, const double *r1, const double *r2, hipStream_t s);
// dst[perm[i]] = src[i] for the 3 rows of n cotangents: back across a sort
// whose perm[new index] = old index
hipError_t launch_adj_unsort(const int *perm, int64_t n, const double *src, double *dst,
                             hipStream_t s);

// Stable counting sort by (cell, sub-cell bin): bins = nx * sub, tile-major
// counts. Sorting finer than the cell keeps each wave's particles within a
// fraction of a cell, so the drift between sorts only disturbs the waves
// next to a cell boundary.
struct SortArgs {
    int64_t n;
    int nx, sub, nbins;
    int ntiles, tiles_per_group, ngroups;
    double xmin, dx;
    const double *x, *v1, *v2, *w;
    const int64_t *id;
    double *x2, *v12, *v22, *w2;
    int64_t *id2;
    int *counts;      // [ntiles][nbins]: bin counts, then (in place) scatter offsets
    int *gsum;        // [ngroups][nbins]: group sums, then (in place) group bases
    int *perm;        // if not null: perm[new index] = old index
};
constexpr int kSortTile = 64 * 16;     // particles per wave tile
constexpr int kSortBinsMax = 1024;
struct SortGeometry {
    int sub, nbins, ntiles, tiles_per_group, ngroups;
};
SortGeometry sort_geometry(int64_t n, int nx);
hipError_t launch_sort(const SortArgs &a, int has_v2, hipStream_t s);
// the sorts of an ensemble's instances (1d2v), one launch per stage (device
// arrays: SortArgs per instance, first tile workgroup and first group
// column of each instance, ninst + 1 entries each)
struct SortBatch {
    const SortArgs *args;
    const int *tile_start;
    const int *group_start;
    int ninst;
};
hipError_t launch_sort_batch(const SortBatch &b, int total_tiles, int total_groups, int max_bins,
                             hipStream_t s);

// sample!(pg, alpha, k, sigma, mesh) into SoA arrays (id = identity); a
// newton that does not converge sets E_FAR in err
hipError_t launch_sample_landau(int64_t n, int dim_v, double alpha, double k, double sigma,
                                double dimx, double *x, double *v1, double *v2, double *w,
                                int64_t *id, int *err, hipStream_t s);

// AoS (reference ParticleGroup layout) <-> SoA with the permutation id.
hipError_t launch_aos_to_soa(const double *aos, int64_t ld, int64_t n, int dim_v, double *x,
                             double *v1, double *v2, double *w, int64_t *id, hipStream_t s);
hipError_t launch_soa_to_aos(double *aos, int64_t ld, int64_t n, int dim_v, const double *x,
                             const double *v1, const double *v2, const double *w,
                             const int64_t *id, hipStream_t s);

// Batched ParticleMeshCoupling1D verbs.
struct PmcArgs {
    int64_t n;
    int64_t chunk;
    int iters;
    int nx;
    int ngroups;
    int degree;
    double xmin, dx, scaling;
    double qx[3], qw[3];
    double pc[36], pf[36];        // SplinePP tables
    double qm;
    const double *x, *x_new, *q;
    const double *field;          // field dofs / pp matrix / bfield
    double *out;                  // evaluate output / v in-out
    double *acc;                  // rho or j (accumulated into)
    double *part, *gpart;
    unsigned *cnt;
    int *err;
};
enum PmcOp { PMC_EVAL = 0, PMC_EVAL_PP, PMC_CHARGE, PMC_CHARGE_PP, PMC_CURRENT_B, PMC_CURRENT_NOB,
             PMC_CURRENT_PP };
hipError_t launch_pmc(int op, const PmcArgs &a, int blocks, hipStream_t s);

// shared by the C ABI files: record the thread's last error message and
// return code; check that a GPU is visible; compute units of the device
int abi_fail(int code, const char *msg);
int abi_check_device();
int abi_device_cus();

}  // namespace gl

struct gempic_maxwell2d;

namespace gl {
// TwoDMaxwell operations on device arrays (pic2d.hip), queued on the
// solver's own stream: compute_e_from_b!, compute_b_from_e!, E -= M1^{-1} j
// (compute_e_from_j!) and compute_rho_from_e! (maxwell_2d_fem.jl:370-499)
hipStream_t m2d_stream(gempic_maxwell2d *m);
void m2d_e_from_b(gempic_maxwell2d *m, double dt, double *const *b, double *const *e);
void m2d_b_from_e(gempic_maxwell2d *m, double dt, double *const *e, double *const *b);
void m2d_e_from_j(gempic_maxwell2d *m, const double *j, int component, double *e);
// j = (sum of the partial grids part2[0 .. splits - 1], in order) * scale,
// then e -= the inverse 1-form mass of j's component; false (nothing
// queued) where the grid takes the general path (m2d_e_from_j)
bool m2d_e_from_parts(gempic_maxwell2d *m, const double *part2, int splits, double scale, double *j,
                      int component, double *e);
void m2d_rho_from_e(gempic_maxwell2d *m, double *const *e, double *rho);
}  // namespace gl
