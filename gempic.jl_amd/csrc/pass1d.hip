// pass1d.hip -- CDNA4 kernels of the 1D particle-field path.
//
// One kernel template (k_pass1d) runs a particle pass: it streams the SoA
// particle arrays once, with the fields of the pass in LDS, applies the
// sub-steps selected by MODE to each particle in registers, and accumulates
// the charge/current deposits with the deterministic per-lane cell caches of
// device_common.hpp, writing one partial row per workgroup. The field part
// (k_epi1d, a second launch) sums the rows in a fixed order (workgroup slices,
// then the last of its workgroups to finish sums the group rows) and runs the
// field solves of the sub-steps, so a whole Strang step is two launches.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_common.hpp"
#include "lib_internal.hpp"

using namespace gk;

namespace gl {

// ---------------------------------------------------------------------------
// Field-side helpers run by one workgroup on LDS arrays
// ---------------------------------------------------------------------------

// solve_circulant! (maxwell_1d_fem.jl:222-240) with the FFTW halfcomplex
// transform pair written out as direct sums over a twiddle table:
// out = HC2R(eig (.) R2HC(rhs)) / n. Threads [t0, t0 + nt) of the
// workgroup take part (several independent solves can run side by side);
// `sync` is a barrier over the whole workgroup, so every thread calls this
// with the same sequence of solves. Each sum runs left to right over j (or
// k), with the twiddle index advanced incrementally instead of (j*k) % n.
// Output of a circulant solve: out[j] = value (the default), or a fused
// update of a field by the thread that owns node j.
struct StoreOut {
    double *out;
    __device__ void operator()(int j, double v) const { out[j] = v; }
};

template <class Emit>
__device__ void circ_solve_part(const double *eig, const double *rhs, double rhs_scale, Emit emit,
                                double *hc, const double *cos_t, const double *sin_t, int n, int t0,
                                int nt)
{
    // first pass: R2HC of rhs * rhs_scale (1, or the sub-step's dt that the
    // reference applies to j_dofs before the solve), then the halfcomplex
    // product with eig by the thread that holds both halves of mode k
    const int tid = (int)threadIdx.x - t0;
    const bool mine = tid >= 0 && tid < nt;
    if (mine)
        for (int k = tid; k <= n / 2; k += nt) {
            double re = 0.0, im = 0.0;
            int t = 0;
#pragma unroll 8
            for (int j = 0; j < n; ++j) {
                const double r = rhs[j] * rhs_scale;
                re += r * cos_t[t];
                im -= r * sin_t[t];
                t += k;
                if (t >= n) t -= n;
            }
            if (k == 0 || k == n / 2) {
                hc[k] = re * eig[k];
            } else {
                double re_p = re * eig[k] - im * eig[n - k];
                double im_p = re * eig[n - k] + im * eig[k];
                hc[k] = re_p;
                hc[n - k] = im_p;
            }
        }
    __syncthreads();
    // second pass: HC2R / n
    if (mine)
        for (int j = tid; j < n; j += nt) {
            double s = hc[0];
            int t = j;
#pragma unroll 8
            for (int k = 1; k < n - k; ++k) {
                s += 2.0 * (hc[k] * cos_t[t] - hc[n - k] * sin_t[t]);
                t += j;
                if (t >= n) t -= n;
            }
            if ((n & 1) == 0) s += hc[n / 2] * ((j & 1) ? -1.0 : 1.0);
            emit(j, s / (double)n);
        }
    __syncthreads();
}

// out = solve_circulant(eig, rhs) (maxwell_1d_fem.jl:222-240) by the whole
// workgroup
__device__ void circ_solve(const double *eig, const double *rhs, double *out, double *hc,
                           const double *cos_t, const double *sin_t, int n)
{
    circ_solve_part(eig, rhs, 1.0, StoreOut{out}, hc, cos_t, sin_t, n, 0, blockDim.x);
}

// One circulant solve per third of the workgroup: solve g (0, 1, 2) reads
// rhs_g (times scale_g), multiplies by eig_g in halfcomplex form and writes
// out_g, with hc_g as its halfcomplex workspace. The arguments are picked
// with selects so that they stay in registers.
struct SolveArgs {
    const double *eig, *rhs;
    double scale;
    double *out, *hc;
};

__device__ void circ_solve3(SolveArgs s0, SolveArgs s1, SolveArgs s2, const double *cos_t,
                            const double *sin_t, int n)
{
    const int G = blockDim.x / 3;
    const int g = (int)threadIdx.x / G;
    const double *eig = g == 1 ? s1.eig : (g == 2 ? s2.eig : s0.eig);
    const double *rhs = g == 1 ? s1.rhs : (g == 2 ? s2.rhs : s0.rhs);
    const double scale = g == 1 ? s1.scale : (g == 2 ? s2.scale : s0.scale);
    double *out = g == 1 ? s1.out : (g == 2 ? s2.out : s0.out);
    double *hc = g == 1 ? s1.hc : (g == 2 ? s2.hc : s0.hc);
    circ_solve_part(eig, rhs, scale, StoreOut{out}, hc, cos_t, sin_t, n, (g < 3 ? g : 3) * G, G);
}

__device__ void ep_e_from_j(double *E, const double *j, const double *inv_mass, double dx,
                            double *work, double *hc, const double *tabs, int n)
{
    (void)work;
    auto emit = [E, dx](int i, double v) {
        double wv = v / dx;
        E[i] -= wv;
    };
    circ_solve_part(inv_mass, j, 1.0, emit, hc, tabs + T_COS * n, tabs + T_SIN * n, n, 0,
                    blockDim.x);
}

// compute_b_from_e! (maxwell_1d_fem.jl:407-420): B[i] += dt/dx (E[i-1] - E[i])
__device__ void ep_b_from_e(double *B, const double *E, double dt, double dx, int n)
{
    const double coef = dt / dx;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int im = i == 0 ? n - 1 : i - 1;
        B[i] = B[i] + coef * (E[im] - E[i]);
    }
    __syncthreads();
}

// compute_e_from_b! (maxwell_1d_fem.jl:384-396): E += dt/dx W(B)
__device__ void ep_e_from_b(double *E, const double *B, double dt, double dx, double *work,
                            double *hc, const double *tabs, int n)
{
    const double coef = dt / dx;
    (void)work;
    auto emit = [E, coef](int i, double v) { E[i] += coef * v; };
    circ_solve_part(tabs + T_AMP * n, B, 1.0, emit, hc, tabs + T_COS * n, tabs + T_SIN * n, n, 0,
                    blockDim.x);
}

// compute_e_from_rho! (maxwell_1d_fem.jl:244-255)
__device__ void ep_e_from_rho(double *E, const double *rho, double *work, double *hc,
                              const double *tabs, int n)
{
    circ_solve(tabs + T_POIS * n, rho, work, hc, tabs + T_COS * n, tabs + T_SIN * n, n);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int im = i == 0 ? n - 1 : i - 1;
        E[i] = work[im] - work[i];
    }
    __syncthreads();
}

// sum(c1 .* solve_circulant(eig, c2)) * dx (inner_product, maxwell_1d_fem.jl:461-475),
// summed left to right by one thread.
__device__ double ep_inner(const double *c1, const double *c2, const double *eig, double dx,
                           double *work, double *hc, const double *tabs, int n)
{
    circ_solve(eig, c2, work, hc, tabs + T_COS * n, tabs + T_SIN * n, n);
    __shared__ double s_r;
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int i = 0; i < n; ++i) r += c1[i] * work[i];
        s_r = r * dx;
    }
    __syncthreads();
    double r = s_r;
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// Deterministic grid-wide reduction of the per-workgroup deposit rows.
// Workgroups are grouped by kGroup; the last one to finish in a group sums
// the group's rows, and the last group-finisher sums the group rows into
// `red` (LDS). Every sum runs in a fixed order that depends only on the
// launch geometry. Returns true only in that final workgroup. Counters are
// reset for the next launch.
// ---------------------------------------------------------------------------

// out[idx] = sum over rows r in [0, nrows) of base[r*len + idx], for the
// entries idx whose deposit slot (idx / nx) is set in depmask. The rows are
// split into `k` contiguous chunks summed by k threads (loads issued eight at
// a time), then the k chunk sums are added in chunk order. `scratch` is LDS
// of at least blockDim.x doubles. Called by all threads of the workgroup.
__device__ void sum_rows(const double *base, int nrows, int len, unsigned depmask, int nx,
                         double *out, double *scratch)
{
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int k = len >= nt ? 1 : nt / len;
    for (int idx0 = 0; idx0 < len; idx0 += nt / k) {
        const int idx = idx0 + tid % (nt / k);
        const int h = tid / (nt / k);
        double s = 0.0;
        const bool live = idx < len && h < k && ((depmask >> (idx / nx)) & 1u);
        if (live) {
            const int r0 = (int)(((long)nrows * h) / k);
            const int r1 = (int)(((long)nrows * (h + 1)) / k);
            int r = r0;
            for (; r + 8 <= r1; r += 8) {
                double t[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) t[u] = base[(size_t)(r + u) * len + idx];
#pragma unroll
                for (int u = 0; u < 8; ++u) s += t[u];
            }
            for (; r < r1; ++r) s += base[(size_t)r * len + idx];
        }
        scratch[tid] = s;
        __syncthreads();
        if (h == 0 && idx < len) {
            double tot = 0.0;
            if ((depmask >> (idx / nx)) & 1u)
                for (int hh = 0; hh < k; ++hh) tot += scratch[hh * (nt / k) + tid];
            out[idx] = tot;
        }
        __syncthreads();
    }
}

__device__ bool grid_reduce(double *part, double *gpart, unsigned *cnt, int ngroups, int len,
                            unsigned depmask, int nx, double *red, double *scratch)
{
    __shared__ int s_last;
    const int tid = threadIdx.x;
    __threadfence();
    __syncthreads();
    const int g = blockIdx.x / kGroup;
    const int g0 = g * kGroup;
    const int gn = min(kGroup, (int)gridDim.x - g0);
    if (tid == 0) {
        unsigned prev = atomicAdd(&cnt[g], 1u);
        s_last = (prev == (unsigned)(gn - 1));
    }
    __syncthreads();
    if (!s_last) return false;
    __threadfence();
    sum_rows(part + (size_t)g0 * len, gn, len, depmask, nx, red, scratch);
    for (int idx = tid; idx < len; idx += blockDim.x) gpart[(size_t)g * len + idx] = red[idx];
    __threadfence();
    __syncthreads();
    if (tid == 0) {
        cnt[g] = 0u;
        unsigned prev = atomicAdd(&cnt[ngroups], 1u);
        s_last = (prev == (unsigned)(ngroups - 1));
    }
    __syncthreads();
    if (!s_last) return false;
    __threadfence();
    sum_rows(gpart, ngroups, len, depmask, nx, red, scratch);
    for (int idx = tid; idx < len; idx += blockDim.x)
        if (!((depmask >> (idx / nx)) & 1u)) red[idx] = 0.0;
    if (tid == 0) cnt[ngroups] = 0u;
    __syncthreads();
    return true;
}

// Copy the constant tables into LDS (all threads).
__device__ void load_tabs(const double *tabs, double *lds, int nx)
{
    for (int i = threadIdx.x; i < NTAB * nx; i += blockDim.x) lds[i] = tabs[i];
    __syncthreads();
}

// ---------------------------------------------------------------------------
// Per-particle pieces
// ---------------------------------------------------------------------------

// field value from precomputed basis values at cell `cell` (evaluate,
// particle_mesh_coupling_1d.jl:438-453)
template <int P>
__device__ __forceinline__ double gather(const double (&b)[P + 1], const double *f, int cell, int nx)
{
    int k = first_node(cell, P, nx);
    double v = 0.0;
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        v += f[k] * b[i];
        k = next_node(k, nx);
    }
    return v;
}

// Fields copied to LDS with kPad wrapped nodes on each side: node k of the
// periodic grid, k in [-kPad, nx + kPad), sits at fp[k] with fp = base + kPad.
constexpr int kPad = 4;


__device__ __forceinline__ void load_padded(double *dst, const double *src, int nx)
{
    for (int i = threadIdx.x; i < nx + 2 * kPad; i += blockDim.x) {
        int k = i - kPad;
        while (k < 0) k += nx;
        while (k >= nx) k -= nx;
        dst[i] = src[k];
    }
}

// evaluate with basis values b at cell `cell` in [0, nx): nodes cell-P..cell
template <int P>
__device__ __forceinline__ double gather_p(const double (&b)[P + 1], const double *fp, int cell)
{
    const double *f = fp + (cell - P);
    double v = 0.0;
#pragma unroll
    for (int i = 0; i <= P; ++i) v += f[i] * b[i];
    return v;
}

// one update_jv! segment (particle_mesh_coupling_1d.jl:385-414): the
// Gauss-Legendre line integral of the degree-P1 splines over [lower, upper],
// scaled by sign*dx.
template <int P1, int NQ>
__device__ __forceinline__ void segment(double lower, double upper, double sign, const double *qx,
                                        const double *qw, double dx, double (&sv)[P1 + 1])
{
    double c1 = 0.5 * (upper - lower);
    double c2 = 0.5 * (upper + lower);
    bspline<P1>(c1 * qx[0] + c2, sv);
    const double s0 = qw[0] * c1;
#pragma unroll
    for (int k = 0; k <= P1; ++k) sv[k] = sv[k] * s0;
#pragma unroll
    for (int jq = 1; jq < NQ; ++jq) {
        double svm[P1 + 1];
        bspline<P1>(c1 * qx[jq] + c2, svm);
#pragma unroll
        for (int k = 0; k <= P1; ++k) sv[k] = sv[k] + svm[k] * qw[jq] * c1;
    }
    const double s1 = sign * dx;
#pragma unroll
    for (int k = 0; k <= P1; ++k) sv[k] = sv[k] * s1;
}

// The same segment in fused arithmetic: for splines of degree P1 <= 3 the
// integral over [lower, upper] is exactly d (S(m) + d^2/24 S''(m)), d = upper -
// lower, m the midpoint (the Gauss-Legendre rule of the reference is exact for
// these polynomials too, so both give the same value up to rounding).
template <int P1>
__device__ __forceinline__ void segment_fast(double lower, double upper, double sign, double dx,
                                             double (&sv)[P1 + 1])
{
    const double d = upper - lower;
    const double m = __builtin_fma(0.5, d, lower);
    bspline_fast<P1>(m, sv);
    const double f = (sign * dx) * d;
    if constexpr (P1 >= 2) {
        double c[P1 + 1];
        bspline_d2_fast<P1>(m, c);
        const double h = (d * d) * (1.0 / 24.0);
#pragma unroll
        for (int k = 0; k <= P1; ++k) sv[k] = __builtin_fma(h, c[k], sv[k]) * f;
    } else {
#pragma unroll
        for (int k = 0; k <= P1; ++k) sv[k] = sv[k] * f;
    }
}

// gather_p in fused arithmetic
template <int P>
__device__ __forceinline__ double gather_pf(const double (&b)[P + 1], const double *fp, int cell)
{
    const double *f = fp + (cell - P);
    double v = f[0] * b[0];
#pragma unroll
    for (int i = 1; i <= P; ++i) v = __builtin_fma(f[i], b[i], v);
    return v;
}

// Segment s of add_current_update_v! (:296-376 / :471-529): the old cell,
// the new cell, then the cells in between, each with its integration bounds.
__device__ __forceinline__ void segment_bounds(int s, int io, double ro, int inw, double rn,
                                               double &lo, double &up, double &sg, int &cell)
{
    if (io == inw) {
        cell = io;
        if (ro < rn) { lo = ro; up = rn; sg = 1.0; }
        else { lo = rn; up = ro; sg = -1.0; }
    } else if (io < inw) {
        sg = 1.0;
        if (s == 0) { lo = ro; up = 1.0; cell = io; }
        else if (s == 1) { lo = 0.0; up = rn; cell = inw; }
        else { lo = 0.0; up = 1.0; cell = io + s - 1; }
    } else {
        sg = -1.0;
        if (s == 0) { lo = rn; up = 1.0; cell = inw; }
        else if (s == 1) { lo = 0.0; up = ro; cell = io; }
        else { lo = 0.0; up = 1.0; cell = inw + s - 1; }
    }
}

__device__ __forceinline__ bool pos_ok(double xi) { return fabs(xi) < 1.0e9; }

// Element i (0 <= i < 2^28) of a workgroup's slice of a particle array, with
// the byte offset as a 32-bit unsigned value: the access is then the slice's
// (uniform) base address plus a 32-bit lane offset, with no 64-bit address
// arithmetic per lane.
template <class T>
__device__ __forceinline__ T &elem(T *p, int i)
{
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + (uint32_t)i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ const T &elem(const T *p, int i)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + (uint32_t)i * (uint32_t)sizeof(T));
}

// The particles of workgroup bid of a particle pass: base .. base + cnt - 1.
// With generations (PassArgs::gen_blocks), workgroup bid = g * gen_blocks + r
// takes gen_chunk[g] particles from gen_base[g] + r * gen_chunk[g]; else bid
// takes `chunk` from bid * chunk. Iteration it gives lane l of wave wv the
// particle it * kBlock + wv * 64 + l, and `iters` counts the calling wave's
// iterations up to its last particle (the loops have no barrier inside, so
// the waves of a workgroup may run different counts).
struct PassChunk {
    int64_t base;
    int cnt, iters;
};

__device__ __forceinline__ PassChunk pass_chunk(const PassArgs &a, int bid)
{
    PassChunk c;
    int64_t chunk;
    if (a.gen_blocks > 0) {
        const int g = bid / a.gen_blocks;
        chunk = a.gen_chunk[g];
        c.base = a.gen_base[g] + (int64_t)(bid - g * a.gen_blocks) * chunk;
    } else {
        chunk = a.chunk;
        c.base = (int64_t)bid * a.chunk;
    }
    const int64_t left = a.n - c.base;
    c.cnt = left <= 0 ? 0 : left < chunk ? (int)left : (int)chunk;
    const int wave_first = (int)(threadIdx.x / kWave) * kWave;
    // the same in every lane of the wave: kept in a scalar register, so the
    // loops over it branch on the scalar unit
    c.iters = __builtin_amdgcn_readfirstlane(c.cnt > wave_first ? (c.cnt - wave_first + kBlock - 1) / kBlock : 0);
    return c;
}

// ---------------------------------------------------------------------------
// The particle pass
// ---------------------------------------------------------------------------
template <int P0, int P1, int MODE>
__device__ __forceinline__ void pass1d_body(const PassArgs &a, int bid, double *smem)
{
    constexpr bool HE_PREV = (MODE & M_HE_PREV) != 0;
    constexpr bool HE = (MODE & M_HE) != 0;
    constexpr bool HP2A = (MODE & M_HP2A) != 0;
    constexpr bool HP1 = (MODE & M_HP1) != 0;
    constexpr bool HP2B = (MODE & M_HP2B) != 0;
    constexpr bool V1 = (MODE & M_V1) != 0;
    constexpr bool RHO_ONLY = (MODE & M_RHO) != 0;
    constexpr bool FAST = (MODE & M_FAST) != 0;
    static_assert(!FAST || (!V1 && !RHO_ONLY), "fused arithmetic covers the 1d2v sub-steps");
    // operatorHp1 (1d2v) also deposits rho into j_dofs[2]
    // (hamiltonian_splitting_1d2v.jl:90-108); when Hp2 follows in the same pass
    // it clears j_dofs[2] first (:132-133) and nothing reads the rho between, so
    // that deposit is dead and skipped.
    constexpr bool DEP_RHO = (HP1 && !V1 && !HP2B) || RHO_ONLY;
    constexpr bool NEED_NEW = DEP_RHO || HP2B;           // work at the post-Hp1 position
    constexpr int NEW_ROW = HP2B ? D_J2B : D_RHO;        // the new-position cache's row
    constexpr unsigned DEPMASK = (HP2A ? 1u << D_J2A : 0u) | (HP1 ? 1u << D_J1 : 0u) |
                                 (DEP_RHO ? 1u << D_RHO : 0u) | (HP2B ? 1u << D_J2B : 0u);
    constexpr int NQ = (P1 + 2) / 2;
    constexpr int W0 = P0 + 1, W1 = P1 + 1;
    static_assert(P0 <= kPad && P1 <= kPad, "LDS field padding must cover the spline support");

    const int nx = a.nx;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int nxp = nx + 2 * kPad;
    double *pE1 = smem;                                   // padded LDS copies of the fields
    double *pE2p = pE1 + nxp;
    double *pE2 = pE2p + nxp;
    double *pB = pE2 + nxp;
    double *rows = pB + nxp;                              // [kWaves][NDEP][nx] deposit rows
    // this wave's cache-B slots (deposit modes only)
    double *slots = rows + kWaves * NDEP * nx + wave * (kSlotsPerLane * kWave);
    // unpadded views (node k at s*[k])
    double *sE1 = pE1 + kPad, *sE2p = pE2p + kPad, *sE2 = pE2 + kPad, *sB = pB + kPad;

    // this workgroup's chunk: particles base .. base + cnt_here - 1, addressed
    // through per-chunk base pointers and a 32-bit index
    const PassChunk pc = pass_chunk(a, bid);
    const int64_t base = pc.base;
    const int cnt_here = pc.cnt;
    const int iters = pc.iters;
    double *__restrict__ px = a.x + base;
    double *__restrict__ pv1 = a.v1 + base;
    double *__restrict__ pv2 = a.v2 + base;
    const double *__restrict__ pw = a.w + base;

    // software pipeline: the loads of iteration it + kPrefetch are issued
    // before the work of iteration it; the first ones before the fields are
    // staged, so that their latencies overlap
    // (a lane past the chunk's end loads its last particle again: every lane
    // loads, so the registers need no copies for the lanes that do not)
    auto load_particle = [&](int j, double &lx, double &lv1, double &lv2, double &lw) {
        const int jj = j < cnt_here ? j : cnt_here - 1;
        lx = elem(px, jj);
        lw = elem(pw, jj);
        if (!RHO_ONLY) lv1 = elem(pv1, jj);
        if (!V1 && !RHO_ONLY) lv2 = elem(pv2, jj);
    };
    double nx_x = 0.0, nx_v1 = 0.0, nx_v2 = 0.0, nx_w = 0.0;     // iteration it + 1
    double n2_x = 0.0, n2_v1 = 0.0, n2_v2 = 0.0, n2_w = 0.0;     // it + 2 (kPrefetch 2)
    if (iters > 0) load_particle(tid, nx_x, nx_v1, nx_v2, nx_w);
    if (kPrefetch > 1 && iters > 1) load_particle(tid + kBlock, n2_x, n2_v1, n2_v2, n2_w);

    load_padded(pE1, a.grid + S_E1 * nx, nx);
    load_padded(pE2p, a.grid + S_E2P * nx, nx);
    load_padded(pE2, a.grid + S_E2 * nx, nx);
    load_padded(pB, a.grid + S_B * nx, nx);
    for (int i = tid; i < kWaves * NDEP * nx; i += kBlock) rows[i] = 0.0;
    __syncthreads();
    double *myrow = rows + wave * NDEP * nx;

    static_assert(W0 + W1 + W0 <= kSlotsPerLane, "cache-B slots");
    Cache2<W0> c_j2a;
    Cache2<W1> c_j1;
    Cache2<W0> c_new;
    if constexpr (HP2A) c_j2a.init(slots);
    if constexpr (HP1) c_j1.init(slots + W0 * kWave);
    if constexpr (NEED_NEW) c_new.init(slots + (W0 + W1) * kWave);

    const double xmin = a.xmin, dx = a.dx, qm = a.qm;
    const double rdx = 1.0 / dx;     // div_by: (x - xmin) / dx, correctly rounded
    bool e_inexact = false, e_far = false;   // error flags of this lane's particles

    for (int it = 0; it < iters; ++it) {
        const int i = it * kBlock + tid;
        const bool valid = i < cnt_here;
        double x = nx_x, v1 = nx_v1, v2 = nx_v2, w = nx_w;
        if constexpr (kPrefetch > 1) {
            nx_x = n2_x;
            nx_v1 = n2_v1;
            nx_v2 = n2_v2;
            nx_w = n2_w;
            if (it + 2 < iters) load_particle(i + 2 * kBlock, n2_x, n2_v1, n2_v2, n2_w);
        } else {
            (void)n2_x;
            (void)n2_v1;
            (void)n2_v2;
            (void)n2_w;
            if (it + 1 < iters) load_particle(i + kBlock, nx_x, nx_v1, nx_v2, nx_w);
        }
        // marker charge, as (w*charge)*common_weight (hamiltonian_splitting_1d2v.jl:62-64)
        const double q = (w * a.charge) * a.cw;

        double xi = div_by(x - xmin, dx, rdx);
        e_inexact |= valid && !pos_ok(xi);
        const int ic = (int)xi;                          // trunc(Int, xi)
        const double off = xi - (double)ic;
        const int key_old = imod(ic, nx);

        double b1[W1];
        double b0[W0];
        if constexpr (FAST && (HE_PREV || HE || HP2A)) bspline_pair_fast<P0, P1>(off, b0, b1);
        else if constexpr ((HE_PREV || HE || HP2A) && !V1) bspline_pair<P0, P1>(off, b0, b1);
        else if constexpr (HE_PREV || HE || HP2A) bspline<P1>(off, b1);

        if constexpr (V1 && (HE_PREV || HE)) {
            // operatorHB 1d1v (hamiltonian_splitting_1d1v.jl:113-125): v += dt * E1(x)
            const double e1 = gather_p<P1>(b1, sE1, key_old);
            if (HE_PREV) v1 = v1 + a.dt_he * e1;
            if (HE) v1 = v1 + a.dt_he * e1;
        }
        if constexpr (FAST && (HE_PREV || HE)) {
            // operatorHE 1d2v in fused arithmetic
            const double e1 = gather_pf<P1>(b1, sE1, key_old);
            const double dtqm = a.dtqm_he;
            if (HE_PREV) {
                const double e2 = gather_pf<P0>(b0, sE2p, key_old);
                v1 = __builtin_fma(dtqm, e1, v1);
                v2 = __builtin_fma(dtqm, e2, v2);
            }
            if (HE) {
                const double e2 = gather_pf<P0>(b0, sE2, key_old);
                v1 = __builtin_fma(dtqm, e1, v1);
                v2 = __builtin_fma(dtqm, e2, v2);
            }
        } else if constexpr (!V1 && (HE_PREV || HE)) {
            // operatorHE 1d2v (hamiltonian_splitting_1d2v.jl:198-215)
            const double e1 = gather_p<P1>(b1, sE1, key_old);
            const double dtqm = a.dtqm_he;
            if (HE_PREV) {
                const double e2 = gather_p<P0>(b0, sE2p, key_old);
                v1 = v1 + dtqm * e1;
                v2 = v2 + dtqm * e2;
            }
            if (HE) {
                const double e2 = gather_p<P0>(b0, sE2, key_old);
                v1 = v1 + dtqm * e1;
                v2 = v2 + dtqm * e2;
            }
        }
        // fused arithmetic: B at the old cell's nodes, for Hp2 and Hp1's first piece
        double bold[W1];
        if constexpr (FAST && (HP2A || HP1)) {
#pragma unroll
            for (int k = 0; k < W1; ++k) bold[k] = sB[key_old - P1 + k];
        }
        if constexpr (HP2A) {
            // operatorHp2 (hamiltonian_splitting_1d2v.jl:146-163)
            double vals[W0];
            if constexpr (FAST) {
                const double bf = gather_pf<P1>(b1, bold, P1);
                v1 = __builtin_fma(a.dtqm_hp2 * v2, bf, v1);
                const double ws = (q * v2) * a.scaling0;
#pragma unroll
                for (int k = 0; k < W0; ++k) vals[k] = ws * b0[k];
            } else {
                const double bf = gather_p<P1>(b1, sB, key_old);
                v1 = v1 + a.dtqm_hp2 * v2 * bf;
                const double wq = q * v2;
#pragma unroll
                for (int k = 0; k < W0; ++k) vals[k] = wq * b0[k] * a.scaling0;
            }
            cache_add<W0>(c_j2a, valid, key_old, vals, myrow + D_J2A * nx, nx);
        }
        double xi_new = 0.0;       // (x - xmin) / dx after Hp1, when the mod kept x_new
        bool xi_new_ok = false;
        int inw_new = 0;           // its cell (trunc) and offset
        double rn_new = 0.0;
        if constexpr (HP1) {
            // operatorHp1 (hamiltonian_splitting_1d2v.jl:53-82 / _1d1v.jl:308-329)
            const double xn = FAST ? __builtin_fma(a.dt_hp1, v1, x) : x + a.dt_hp1 * v1;
            const double xin = div_by(xn - xmin, dx, rdx);
            e_inexact |= valid && !pos_ok(xin);
            const int inw = V1 ? (int)floor(xin) : (int)xin;
            const double rn = xin - (double)inw;
            const int dseg = inw - ic;
            // a move across more than kMaxCellsPerStep cells (or a non-finite
            // position) is reported instead of integrated segment by segment
            const bool far = !pos_ok(xin) || !pos_ok(xi) || dseg > kMaxCellsPerStep ||
                             dseg < -kMaxCellsPerStep;
            e_far |= valid && far;
            if constexpr (FAST) {
                // The path from xi to xin as signed integrals over its piece in
                // each cell (the midpoint rule of segment_fast is odd in the
                // length, so a piece run backwards is its negative): first the
                // old cell, from the old offset to the new one when the cell is
                // kept, else to the cell's edge in the direction of motion; then,
                // for the lanes that cross, the new cell from its edge to the new
                // offset and the whole cells in between.
                const bool cross = dseg != 0 && !far;
                const bool right = dseg > 0;
                const double qs = q * a.scaling1;
                auto piece = [&](bool act, int key, double from, double to, const double *bk) {
                    double sv[W1];
                    segment_fast<P1>(from, to, 1.0, dx, sv);
                    double vals[W1];
                    double sb = sv[0] * bk[0];
#pragma unroll
                    for (int m = 0; m < W1; ++m) {
                        vals[m] = qs * sv[m];
                        if (m > 0) sb = __builtin_fma(sv[m], bk[m], sb);
                    }
                    const double v2s = __builtin_fma(-qm, sb, v2);
                    if (act) v2 = v2s;
                    cache_add<W1>(c_j1, act, key, vals, myrow + D_J1 * nx, nx);
                };
                piece(valid, key_old, off, cross ? (right ? 1.0 : 0.0) : rn, bold);
                if (__ballot(valid && cross)) {
                    const int extra = cross ? (right ? dseg : -dseg) : 0;
                    const double edge = right ? 0.0 : 1.0;
                    for (int s = 1;; ++s) {
                        const bool act = valid && s <= extra;
                        if (!__ballot(act)) break;
                        const int cell = s == 1 ? inw : (right ? ic + s - 1 : ic - s + 1);
                        const int key = imod(cell, nx);
                        piece(act, key, edge, s == 1 ? rn : 1.0 - edge, sB + (key - P1));
                    }
                }
            } else {
            const int nseg = far ? 1 : 1 + (dseg < 0 ? -dseg : dseg);
            for (int s = 0;; ++s) {
                const bool act = valid && s < nseg;
                if (!__ballot(act)) break;
                double lo, up, sg;
                int cell;
                segment_bounds(s, ic, off, inw, rn, lo, up, sg, cell);
                double sv[W1];
                segment<P1, NQ>(lo, up, sg, a.qx, a.qw, dx, sv);
                const int key = imod(cell, nx);
                double vals[W1];
                const double *bk = sB + (key - P1);
                double v2s = v2;
#pragma unroll
                for (int m = 0; m < W1; ++m) {
                    vals[m] = q * sv[m] * a.scaling1;
                    if (!V1) v2s = v2s - qm * sv[m] * bk[m];
                }
                if (act) v2 = v2s;
                cache_add<W1>(c_j1, act, key, vals, myrow + D_J1 * nx, nx);
            }
            }
            x = jl_mod(xn, a.Lx);
            xi_new = xin;
            xi_new_ok = __double_as_longlong(x) == __double_as_longlong(xn);
            inw_new = inw;
            rn_new = rn;
        }
        if constexpr (NEED_NEW) {
            // rho deposit at the new position (hamiltonian_splitting_1d2v.jl:93-102,
            // diagnostics.jl:189-193) and the second Hp2; the same division as
            // Hp1's when the mod left the position unchanged
            // (when the mod kept every lane's position, Hp1's cell and offset are
            // reused, and the old cell's key when no lane left its cell)
            double off2;
            int key2;
            if (__ballot(valid && !xi_new_ok) == 0) {
                off2 = rn_new;
                key2 = __ballot(valid && inw_new != ic) == 0 ? key_old : imod(inw_new, nx);
            } else {
                const double xi2 = xi_new_ok ? xi_new : div_by(x - xmin, dx, rdx);
                e_inexact |= valid && RHO_ONLY && !pos_ok(xi2);
                const int ic2 = (int)xi2;
                off2 = xi2 - (double)ic2;
                key2 = imod(ic2, nx);
            }
            double c0[W0];
            double c1[W1];
            if constexpr (FAST) bspline_pair_fast<P0, P1>(off2, c0, c1);
            else if constexpr (HP2B) bspline_pair<P0, P1>(off2, c0, c1);
            else bspline<P0>(off2, c0);
            if constexpr (HP2B) {
                if constexpr (FAST) {
                    const double bf = gather_pf<P1>(c1, sB, key2);
                    v1 = __builtin_fma(a.dtqm_hp2 * v2, bf, v1);
                } else {
                    const double bf = gather_p<P1>(c1, sB, key2);
                    v1 = v1 + a.dtqm_hp2 * v2 * bf;
                }
            }
            // rho: q * S0; Hp2's j2: (q * v2) * S0 (particle_mesh_coupling_1d.jl:120)
            const double wq = HP2B ? q * v2 : q;
            double vals[W0];
            if constexpr (FAST) {
                const double ws = wq * a.scaling0;
#pragma unroll
                for (int k = 0; k < W0; ++k) vals[k] = ws * c0[k];
            } else {
#pragma unroll
                for (int k = 0; k < W0; ++k) vals[k] = wq * c0[k] * a.scaling0;
            }
            cache_add<W0>(c_new, valid, key2, vals, myrow + NEW_ROW * nx, nx);
        }
        if (valid) {
            if constexpr (HP1) elem(px, i) = x;
            if constexpr (HE_PREV || HE || HP2A || HP2B) elem(pv1, i) = v1;
            if constexpr (!V1 && (HE_PREV || HE || HP1)) elem(pv2, i) = v2;
        }
        // a recorded fused step also keeps the state after the pass
        if constexpr (HE && HP2A && HP1 && HP2B && !V1)
            if (a.cx && valid) {
                a.cx[base + i] = x;
                a.cv1[base + i] = v1;
                a.cv2[base + i] = v2;
            }
    }
    if (e_inexact || e_far) atomicOr(a.err, (e_inexact ? (int)E_INEXACT : 0) | (e_far ? (int)E_FAR : 0));

    if constexpr (DEPMASK != 0u) {
        if constexpr (HP2A) cache_drain<W0>(c_j2a, myrow + D_J2A * nx, nx);
        if constexpr (HP1) cache_drain<W1>(c_j1, myrow + D_J1 * nx, nx);
        if constexpr (NEED_NEW) cache_drain<W0>(c_new, myrow + NEW_ROW * nx, nx);
        __syncthreads();
        const int len = NDEP * nx;
        for (int idx = tid; idx < len; idx += kBlock) {
            if (!((DEPMASK >> (idx / nx)) & 1u)) continue;
            double s = 0.0;
            for (int wv = 0; wv < kWaves; ++wv) s += rows[wv * len + idx];
            a.part[(size_t)bid * len + idx] = s;
        }
        // the cross-workgroup sum and the field solves run in k_epi1d
    } else if constexpr (HE && !V1 && !HE_PREV) {
        // operatorHE alone: compute_b_from_e! after the push (:217-218). B is
        // not read by this pass, so workgroup 0 can update it at any time.
        if (bid == 0) {
            ep_b_from_e(sB, sE2, a.dt_he, dx, nx);
            for (int i = tid; i < nx; i += kBlock) a.grid[S_B * nx + i] = sB[i];
        }
    }
}

// The batch of an ensemble: instance i's workgroups are [block_start[i], block_start[i+1])
__device__ __forceinline__ int batch_instance(const int *start, int ninst, int block)
{
    int lo = 0, hi = ninst;              // start[lo] <= block < start[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (start[mid] <= block) lo = mid;
        else hi = mid;
    }
    return lo;
}

template <int P0, int P1, int MODE>
#ifndef GEMPIC_PASS_MIN_WAVES
#define GEMPIC_PASS_MIN_WAVES 4
#endif
__global__ void __launch_bounds__(kBlock, GEMPIC_PASS_MIN_WAVES) k_pass1d(PassArgs a)
{
    extern __shared__ double smem[];
    pass1d_body<P0, P1, MODE>(a, blockIdx.x, smem);
}

// One launch runs the same pass for every instance of an ensemble; each
// workgroup works for one instance with that instance's own launch geometry,
// so every instance's result is the one its own launch would give.
template <int P0, int P1, int MODE>
__global__ void __launch_bounds__(kBlock) k_pass1d_batch(BatchArgs b)
{
    extern __shared__ double smem[];
    const int inst = batch_instance(b.block_start, b.ninst, blockIdx.x);
    pass1d_body<P0, P1, MODE>(b.args[inst], blockIdx.x - b.block_start[inst], smem);
}

// ---------------------------------------------------------------------------
// The field part of a deposit pass: sum of the workgroups' partial rows and
// the field solves of its sub-steps.
//
// It runs as a kernel of its own after the particle pass. The partial rows
// are then visible to it through the kernel boundary, so the particle pass
// needs no device-scope fence: on MI355X such a fence writes the XCD's whole
// L2 back, and with the particle stores in flight hundreds of workgroups each
// doing so cost more than the particle work. Here the L2 holds only this
// kernel's few group rows, so its own fences are cheap: G workgroups each sum
// a slice of the partial rows, and the last one to finish sums the G group
// rows and runs the field solves.
// ---------------------------------------------------------------------------

// index r*nx + k of live entry e (the e-th entry of the rows in mask)
__device__ __forceinline__ int live_index(int e, int nx, unsigned mask)
{
    const int li = e / nx, node = e - li * nx;
    int r = 0, c = -1;
#pragma unroll
    for (int rr = 0; rr < NDEP; ++rr)
        if ((mask >> rr) & 1u) {
            ++c;
            if (c == li) r = rr;
        }
    return r * nx + node;
}

// out[i] = sum of base[b*NDEP*nx + i] over the n rows b = b0 + j * stride,
// j in [0, n), for the live entries i of the rows in mask. Fixed order: the
// k threads of an entry take consecutive slices of j, and their slice sums
// are added in slice order. Called by all threads; scratch holds blockDim.x
// doubles.
__device__ void sum_rows(const double *base, int b0, int n, int stride, int nx, unsigned mask,
                         double *out, double *scratch)
{
    const int nt = blockDim.x, tid = threadIdx.x;
    const int len = NDEP * nx;
    const int E = __popc(mask) * nx;
    const size_t rs = (size_t)stride * len;
    for (int e0 = 0; e0 < E; e0 += nt) {
        const int ec = E - e0 < nt ? E - e0 : nt;
        const int k = nt / ec;
        const int h = tid / ec;
        const int e = e0 + (tid - h * ec);
        double s = 0.0;
        int idx = 0;
        if (h < k) {
            idx = live_index(e, nx, mask);
            const int r0 = (int)(((long)n * h) / k);
            const int r1 = (int)(((long)n * (h + 1)) / k);
            const double *col = base + (size_t)b0 * len + idx;
            int b = r0;
            constexpr int B = 16;
            for (; b + B <= r1; b += B) {
                double t[B];
#pragma unroll
                for (int u = 0; u < B; ++u) t[u] = col[(size_t)(b + u) * rs];
#pragma unroll
                for (int u = 0; u < B; ++u) s += t[u];
            }
            for (; b < r1; ++b) s += col[(size_t)b * rs];
        }
        scratch[tid] = s;
        __syncthreads();
        if (tid < ec) {
            double tot = 0.0;
            for (int hh = 0; hh < k; ++hh) tot += scratch[hh * ec + tid];
            out[idx] = tot;
        }
        __syncthreads();
    }
}

// sum_rows over the rows [b0, b1)
__device__ __forceinline__ void sum_rows_range(const double *base, int b0, int b1, int nx,
                                               unsigned mask, double *out, double *scratch)
{
    sum_rows(base, b0, b1 - b0, 1, nx, mask, out, scratch);
}

// The first level of the field part's sum: group g of ng sums the partial
// rows of its contiguous slice of the nblocks workgroups. (Slices of the
// rows written on the group's own XCD, b mod ng == g, measured the same.)
__device__ __forceinline__ void sum_group_rows(const double *part, int nblocks, int g, int ng, int nx,
                                               unsigned mask, double *out, double *scratch)
{
    const int b0 = (int)(((long)nblocks * g) / ng), b1 = (int)(((long)nblocks * (g + 1)) / ng);
    sum_rows_range(part, b0, b1, nx, mask, out, scratch);
}

// (C rhs)[i] for the circulant C with first column c: sum_j c[(i - j) mod n] rhs[j],
// in four interleaved partial sums
__device__ __forceinline__ double circ_apply(const double *c, const double *rhs, int i, int n)
{
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int t = i;                                    // (i - j) mod n for j = 0
    int j = 0;
    for (; j + 4 <= n; j += 4) {
        const int t1 = t == 0 ? n - 1 : t - 1;
        const int t2 = t1 == 0 ? n - 1 : t1 - 1;
        const int t3 = t2 == 0 ? n - 1 : t2 - 1;
        s0 = __builtin_fma(c[t], rhs[j], s0);
        s1 = __builtin_fma(c[t1], rhs[j + 1], s1);
        s2 = __builtin_fma(c[t2], rhs[j + 2], s2);
        s3 = __builtin_fma(c[t3], rhs[j + 3], s3);
        t = t3 == 0 ? n - 1 : t3 - 1;
    }
    for (; j < n; ++j) {
        s0 = __builtin_fma(c[t], rhs[j], s0);
        t = t == 0 ? n - 1 : t - 1;
    }
    return (s0 + s1) + (s2 + s3);
}

// (C^T rhs)[i] for the circulant C with first column c: sum_j c[(j - i) mod n] rhs[j]
__device__ __forceinline__ double circ_apply_t(const double *c, const double *rhs, int i, int n)
{
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int t = i == 0 ? 0 : n - i;                   // (j - i) mod n for j = 0
    int j = 0;
    for (; j + 4 <= n; j += 4) {
        const int t1 = t + 1 == n ? 0 : t + 1;
        const int t2 = t1 + 1 == n ? 0 : t1 + 1;
        const int t3 = t2 + 1 == n ? 0 : t2 + 1;
        s0 = __builtin_fma(c[t], rhs[j], s0);
        s1 = __builtin_fma(c[t1], rhs[j + 1], s1);
        s2 = __builtin_fma(c[t2], rhs[j + 2], s2);
        s3 = __builtin_fma(c[t3], rhs[j + 3], s3);
        t = t3 + 1 == n ? 0 : t3 + 1;
    }
    for (; j < n; ++j) {
        s0 = __builtin_fma(c[t], rhs[j], s0);
        t = t + 1 == n ? 0 : t + 1;
    }
    return (s0 + s1) + (s2 + s3);
}

// workgroups of k_epi1d for nblocks partial rows
int epi_groups(int nblocks)
{
    int g = (int)std::ceil(std::sqrt((double)nblocks));
    return g < 1 ? 1 : (g > kEpiGroupsMax ? kEpiGroupsMax : g);
}

// PHASE 0: the whole field part (the slice sums, then the last workgroup
// to finish sums the group rows and runs the solves); PHASE 1: only this
// workgroup's slice sum into its group row; PHASE 2: only the part after the
// slice sums (a later launch, so no fences or counters are needed). Phases 1
// and 2 together add the same numbers in the same order as phase 0.
template <int P0, int P1, int MODE, int PHASE = 0>
__device__ __forceinline__ void epi1d_body(const PassArgs &a, int nblocks, int g, int ng,
                                           double *smem)
{
    constexpr bool HE_PREV = (MODE & M_HE_PREV) != 0;
    constexpr bool HE = (MODE & M_HE) != 0;
    constexpr bool HP2A = (MODE & M_HP2A) != 0;
    constexpr bool HP1 = (MODE & M_HP1) != 0;
    constexpr bool HP2B = (MODE & M_HP2B) != 0;
    constexpr bool V1 = (MODE & M_V1) != 0;
    constexpr bool RHO_ONLY = (MODE & M_RHO) != 0;
    constexpr bool WITH_J2A = (MODE & M_WITH_J2A) != 0;
    constexpr bool FAST = (MODE & M_FAST) != 0;
    constexpr bool DEP_RHO = (HP1 && !V1 && !HP2B) || RHO_ONLY;
    constexpr unsigned DEPMASK = (HP2A ? 1u << D_J2A : 0u) | (HP1 ? 1u << D_J1 : 0u) |
                                 (DEP_RHO ? 1u << D_RHO : 0u) | (HP2B ? 1u << D_J2B : 0u);
    constexpr unsigned REDMASK = DEPMASK | (WITH_J2A ? 1u << D_J2A : 0u);
    static_assert(!FAST || (HE && HP2A && HP1 && HP2B && !V1), "fused arithmetic: the fused step");
    (void)HE_PREV;
    (void)HE;

    __shared__ int s_last;
    const int nx = a.nx;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int len = NDEP * nx;
    const double dx = a.dx;
    double *sE1 = smem;                     // [nx] each
    double *sE2 = sE1 + nx;
    double *sB = sE2 + nx;
    double *sRed = sB + nx;                 // [NDEP][nx]
    double *work = sRed + NDEP * nx;        // [nx]
    double *hc = work + nx;                 // [nx]
    double *spare = hc + nx;                // [nx]
    double *T = spare + nx;                 // [NTAB][nx] constant tables
    double *w3 = T + NTAB * nx;             // [6][nx]
    double *scratch = w3 + 6 * nx;          // [blockDim.x]

    if constexpr (PHASE == 1) {
        sum_group_rows(a.part, nblocks, g, ng, nx, REDMASK, a.gpart + (size_t)g * len, scratch);
        return;
    }

    // every workgroup stages the fields and tables first: the one that ends up
    // last has them in LDS without another round trip
    double *G = a.grid;
    for (int i = tid; i < nx; i += nt) {
        sE1[i] = G[S_E1 * nx + i];
        sE2[i] = G[S_E2 * nx + i];
        sB[i] = G[S_B * nx + i];
    }
    for (int i = tid; i < NTAB * nx; i += nt) T[i] = a.tabs[i];
    for (int i = tid; i < len; i += nt) sRed[i] = 0.0;

    if constexpr (PHASE == 0) {
        // group sums of a slice of the workgroups' rows
        sum_group_rows(a.part, nblocks, g, ng, nx, REDMASK, a.gpart + (size_t)g * len, scratch);
        __threadfence();
        __syncthreads();
        if (tid == 0) s_last = atomicAdd(&a.cnt[0], 1u) == (unsigned)(ng - 1);
        __syncthreads();
        if (!s_last) return;
        __threadfence();
        if (tid == 0) a.cnt[0] = 0u;            // counters are zero between launches
    } else {
        (void)s_last;
        __syncthreads();
    }

    sum_rows_range(a.gpart, 0, ng, nx, REDMASK, sRed, scratch);  // ends with a barrier
    double *r_j2a = sRed + D_J2A * nx, *r_j1 = sRed + D_J1 * nx;
    double *r_rho = sRed + D_RHO * nx, *r_j2b = sRed + D_J2B * nx;
    if constexpr (RHO_ONLY) {
        ep_e_from_rho(work, r_rho, spare, hc, T, nx);
        for (int i = tid; i < nx; i += nt) {
            G[S_RHO * nx + i] = r_rho[i];
            G[S_EPOIS * nx + i] = work[i];
        }
    } else if constexpr (V1) {
        // 1d1v Hp1: j_dofs[1] = j, j_dofs[2] = 0; E1 -= M1^{-1} j / dx
        for (int i = tid; i < nx; i += nt) {
            G[S_J1 * nx + i] = r_j1[i];
            G[S_J2 * nx + i] = 0.0;
        }
        ep_e_from_j(sE1, r_j1, T + T_INVM1 * nx, dx, work, hc, T, nx);
        for (int i = tid; i < nx; i += nt) G[S_E1 * nx + i] = sE1[i];
    } else if constexpr (HP2A && !HP1) {
        // operatorHp2 alone (hamiltonian_splitting_1d2v.jl:169-175)
        for (int i = tid; i < nx; i += nt) r_j2a[i] = r_j2a[i] * a.dt_hp2;
        __syncthreads();
        ep_e_from_j(sE2, r_j2a, T + T_INVM0 * nx, dx, work, hc, T, nx);
        for (int i = tid; i < nx; i += nt) {
            G[S_J1 * nx + i] = 0.0;
            G[S_J2 * nx + i] = r_j2a[i];
            G[S_E2 * nx + i] = sE2[i];
        }
    } else if constexpr (HP1 && !HP2A && !HP2B) {
        // operatorHp1 alone (hamiltonian_splitting_1d2v.jl:88-111)
        for (int i = tid; i < nx; i += nt) {
            G[S_J1 * nx + i] = r_j1[i];
            G[S_J2 * nx + i] = r_rho[i];
        }
        ep_e_from_j(sE1, r_j1, T + T_INVM1 * nx, dx, work, hc, T, nx);
        for (int i = tid; i < nx; i += nt) G[S_E1 * nx + i] = sE1[i];
    } else if constexpr (FAST) {
        // the fused step's field part in fused arithmetic: the two Hp2
        // currents go through one M0^{-1} solve (linear, so the two updates
        // of E2 become one), and every solve is the circulant's
        // matrix-vector product (circ_apply)
        double *rhs2 = spare;                          // dt (j2a + j2b)
        for (int i = tid; i < nx; i += nt) rhs2[i] = (r_j2a[i] + r_j2b[i]) * a.dt_hp2;
        __syncthreads();
        const double rdx = 1.0 / dx;
        for (int t = tid; t < 2 * nx; t += nt) {
            const int g = t / nx, i = t - g * nx;
            const double w = g == 0 ? circ_apply(T + T_CM0 * nx, rhs2, i, nx)
                                    : circ_apply(T + T_CM1 * nx, r_j1, i, nx);
            w3[g * nx + i] = w;
        }
        __syncthreads();
        for (int i = tid; i < nx; i += nt) {
            sE2[i] = __builtin_fma(-w3[i], rdx, sE2[i]);
            sE1[i] = __builtin_fma(-w3[nx + i], rdx, sE1[i]);
        }
        __syncthreads();
        const double coef = a.dt_he * rdx;
        for (int i = tid; i < nx; i += nt) {
            G[S_J1 * nx + i] = 0.0;
            G[S_J2 * nx + i] = r_j2b[i] * a.dt_hp2;
            G[S_E2P * nx + i] = sE2[i];
            if (a.hist) a.hist[(NHIST + H_E2P) * nx + i] = sE2[i];
            const int im = i == 0 ? nx - 1 : i - 1;
            work[i] = __builtin_fma(coef, sE2[im] - sE2[i], sB[i]);     // HE's compute_b_from_e!
        }
        __syncthreads();
        for (int i = tid; i < nx; i += nt) sB[i] = work[i];
        if (a.hist)
            for (int i = tid; i < nx; i += nt) a.hist[H_BEND * nx + i] = work[i];
        __syncthreads();
        // HB of this step and (a.more) of the next one: the weak curl of B once
        const double times = a.more ? 2.0 * coef : coef;
        for (int i = tid; i < nx; i += nt) hc[i] = circ_apply(T + T_CAMP * nx, sB, i, nx);
        __syncthreads();
        for (int i = tid; i < nx; i += nt) sE2[i] = __builtin_fma(times, hc[i], sE2[i]);
        __syncthreads();
        if (a.more) {
            for (int i = tid; i < nx; i += nt) {
                const int im = i == 0 ? nx - 1 : i - 1;
                work[i] = __builtin_fma(coef, sE2[im] - sE2[i], sB[i]);
            }
            __syncthreads();
            for (int i = tid; i < nx; i += nt) sB[i] = work[i];
            __syncthreads();
        }
        for (int i = tid; i < nx; i += nt) {
            G[S_E1 * nx + i] = sE1[i];
            G[S_E2 * nx + i] = sE2[i];
            G[S_B * nx + i] = sB[i];
        }
        if (a.hist)
            for (int i = tid; i < nx; i += nt) {
                a.hist[(NHIST + H_E1) * nx + i] = sE1[i];
                a.hist[(NHIST + H_E2) * nx + i] = sE2[i];
                a.hist[(NHIST + H_B) * nx + i] = sB[i];
            }
    } else {
        // fused Strang step: Hp2 (1/2) - Hp1 - Hp2 (1/2) field solves, then
        // HE's B update and HB of this step, and (if another step follows)
        // HB and HE's B update of the next step.
        // The three inverse-mass solves (compute_e_from_j! of Hp2, Hp1, Hp2)
        // are independent: run them side by side, then apply the updates in
        // the reference's order, E2 -= w_a / dx, E1 -= w_1 / dx, E2 -= w_b / dx.
        // j2 of each Hp2 is scaled by dt inside the solve
        // (hamiltonian_splitting_1d2v.jl:173)
        const SolveArgs sa{T + T_INVM0 * nx, r_j2a, a.dt_hp2, w3, w3 + nx};
        const SolveArgs s1{T + T_INVM1 * nx, r_j1, 1.0, w3 + 2 * nx, w3 + 3 * nx};
        const SolveArgs sb{T + T_INVM0 * nx, r_j2b, a.dt_hp2, w3 + 4 * nx, w3 + 5 * nx};
        circ_solve3(sa, s1, sb, T + T_COS * nx, T + T_SIN * nx, nx);
        for (int i = tid; i < nx; i += nt) {
            double wa = sa.out[i] / dx;
            sE2[i] -= wa;
            double w1 = s1.out[i] / dx;
            sE1[i] -= w1;
            double wb = sb.out[i] / dx;
            sE2[i] -= wb;
        }
        __syncthreads();
        for (int i = tid; i < nx; i += nt) {
            G[S_J1 * nx + i] = 0.0;
            G[S_J2 * nx + i] = r_j2b[i] * a.dt_hp2;
            G[S_E2P * nx + i] = sE2[i];
            if (a.hist) a.hist[(NHIST + H_E2P) * nx + i] = sE2[i];
        }
        ep_b_from_e(sB, sE2, a.dt_he, dx, nx);
        if (a.hist)
            for (int i = tid; i < nx; i += nt) a.hist[H_BEND * nx + i] = sB[i];
        // HB of this step and (a.more) of the next one: compute_e_from_b! twice
        // with the same B, so the weak curl W(B) is solved once and added twice
        circ_solve(T + T_AMP * nx, sB, work, hc, T + T_COS * nx, T + T_SIN * nx, nx);
        {
            const double coef = a.dt_he / dx;
            for (int i = tid; i < nx; i += nt) {
                double e = sE2[i] + coef * work[i];
                if (a.more) e = e + coef * work[i];
                sE2[i] = e;
            }
            __syncthreads();
        }
        if (a.more) ep_b_from_e(sB, sE2, a.dt_he, dx, nx);
        for (int i = tid; i < nx; i += nt) {
            G[S_E1 * nx + i] = sE1[i];
            G[S_E2 * nx + i] = sE2[i];
            G[S_B * nx + i] = sB[i];
        }
        if (a.hist)
            for (int i = tid; i < nx; i += nt) {
                a.hist[(NHIST + H_E1) * nx + i] = sE1[i];
                a.hist[(NHIST + H_E2) * nx + i] = sE2[i];
                a.hist[(NHIST + H_B) * nx + i] = sB[i];
            }
    }
}

template <int P0, int P1, int MODE>
__global__ void __launch_bounds__(kBlock) k_epi1d(PassArgs a, int nblocks)
{
    extern __shared__ double smem[];
    epi1d_body<P0, P1, MODE>(a, nblocks, blockIdx.x, gridDim.x, smem);
}

// The field parts of every instance of an ensemble in two launches: the
// slice sums (instance i's epi_groups(its pass workgroups) workgroups are
// [epi_start[i], epi_start[i+1])), then one workgroup per instance for the
// rest. In one launch, the fences and counters of hundreds of instances'
// last-workgroup handshakes serialise on the L2 write-backs.
template <int P0, int P1, int MODE>
__global__ void __launch_bounds__(kBlock) k_epi1d_batch_sum(BatchArgs b)
{
    extern __shared__ double smem[];
    const int inst = batch_instance(b.epi_start, b.ninst, blockIdx.x);
    const int g0 = b.epi_start[inst];
    epi1d_body<P0, P1, MODE, 1>(b.args[inst], b.block_start[inst + 1] - b.block_start[inst],
                                blockIdx.x - g0, b.epi_start[inst + 1] - g0, smem);
}

template <int P0, int P1, int MODE>
__global__ void __launch_bounds__(kBlock) k_epi1d_batch_fin(BatchArgs b)
{
    extern __shared__ double smem[];
    const int inst = blockIdx.x;
    epi1d_body<P0, P1, MODE, 2>(b.args[inst], b.block_start[inst + 1] - b.block_start[inst], 0,
                                b.epi_start[inst + 1] - b.epi_start[inst], smem);
}

size_t epi_shmem_bytes(int nx)
{
    return sizeof(double) * ((size_t)(3 + NDEP + 3 + NTAB + 6) * nx + kBlock);
}

size_t pass_shmem_bytes(int nx, int mode)
{
    // particle pass: padded fields, the per-wave deposit rows and, for
    // deposit modes, the cache-B slots
    const bool dep = (mode & (M_HP2A | M_HP1 | M_HP2B | M_RHO | M_WITH_J2A)) != 0;
    const size_t pass = (size_t)4 * (nx + 2 * kPad) + (size_t)kWaves * NDEP * nx +
                        (dep ? (size_t)kWaves * kSlotsPerLane * kWave : 0);
    // diagnostics: 5 fields, work, hc, 8 sums, scratch, then scratch/tables
    size_t tail = (size_t)NTAB * nx;
    if (tail < (size_t)kBlock) tail = kBlock;
    const size_t diag = (size_t)nx * 8 + 8 + tail;
    return sizeof(double) * (pass > diag ? pass : diag);
}

// ---------------------------------------------------------------------------
// write_step! particle sums (diagnostics.jl:186-250) + field diagnostics
// ---------------------------------------------------------------------------
template <int P0, int P1>
__global__ void __launch_bounds__(kBlock) k_diag1d2v(PassArgs a)
{
    extern __shared__ double smem[];
    const int nx = a.nx;
    const int tid = threadIdx.x;
    double *sE1 = smem, *sE2 = sE1 + nx, *sB = sE2 + nx, *sE1n = sB + nx, *sE2n = sE1n + nx;
    double *sWork = sE2n + nx, *sHc = sWork + nx, *sScr = sHc + nx;
    __shared__ double s_part[kWaves][8];
    __shared__ double s_red[8];
    for (int i = tid; i < nx; i += kBlock) {
        sE1[i] = a.grid[S_E1 * nx + i];
        sE2[i] = a.grid[S_E2 * nx + i];
        sB[i] = a.grid[S_B * nx + i];
        sE1n[i] = a.grid[S_E1N * nx + i];
        sE2n[i] = a.grid[S_E2N * nx + i];
    }
    __syncthreads();
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    int errs = 0;
    const int64_t base = (int64_t)blockIdx.x * a.chunk;
    const double rdx = 1.0 / a.dx;   // div_by: (x - xmin) / dx, correctly rounded
    for (int it = 0; it < a.iters; ++it) {
        const int64_t i = base + (int64_t)it * kBlock + tid;
        if (i >= a.n) continue;
        const double x = a.x[i], v1 = a.v1[i], v2 = a.v2[i], w = a.w[i];
        const double wm = (w * a.mass) * a.cw;
        acc[0] += (v1 * v1 + v2 * v2) * wm;
        acc[1] += v1 * wm;
        acc[2] += v2 * wm;
        double xi = div_by(x - a.xmin, a.dx, rdx);
        if (!pos_ok(xi)) {      // evaluate's trunc(Int, xi) throws InexactError (particle_mesh_coupling_1d.jl:438-453)
            errs |= E_INEXACT;
            xi = 0.0;
        }
        const int ic = (int)xi;
        const double off = xi - (double)ic;
        const int key = imod(ic, nx);
        double b1[P1 + 1], b0[P0 + 1];
        bspline<P1>(off, b1);
        bspline<P0>(off, b0);
        const double e1 = gather<P1>(b1, sE1, key, nx);
        const double e2 = gather<P0>(b0, sE2, key, nx);
        const double bf = gather<P1>(b1, sB, key, nx);
        const double qa = (a.charge * w) * a.cw;
        acc[3] += (v1 * e1 + v2 * e2) * qa;
        const double qb = (w * a.charge) * a.cw;
        acc[4] += qb * v1 * v2 * bf;
    }
    if (errs) atomicOr(a.err, errs);
    // deterministic block sum: wave butterflies, then waves in order
#pragma unroll
    for (int k = 0; k < 5; ++k) {
        double s = wave_sum(acc[k]);
        if (lane_id() == 0) s_part[tid / kWave][k] = s;
    }
    __syncthreads();
    const int len = 8;
    if (tid < len) {
        double s = 0.0;
        if (tid < 5)
            for (int wv = 0; wv < kWaves; ++wv) s += s_part[wv][tid];
        a.part[(size_t)blockIdx.x * len + tid] = s;
    }
    double *red = sScr;
    double *rscr = sScr + 8 + nx;     // kBlock doubles of reduction scratch, then the tables
    if (!grid_reduce(a.part, a.gpart, a.cnt, a.ngroups, len, 0x1fu, len, red, rscr)) return;
    // field diagnostics (diagnostics.jl:221-246)
    load_tabs(a.tabs, rscr, nx);
    const double *T = rscr;
    double *scratch = sScr + 8;
    for (int i = tid; i < nx; i += kBlock) scratch[i] = 0.0;
    __syncthreads();
    ep_e_from_b(scratch, sB, 1.0, a.dx, sWork, sHc, T, nx);
    const double poynting = ep_inner(sE2, scratch, T + T_M0 * nx, a.dx, sWork, sHc, T, nx);
    const double pe1 = ep_inner(sE1, sE1n, T + T_M1 * nx, a.dx, sWork, sHc, T, nx);
    const double pe2 = ep_inner(sE2, sE2n, T + T_M0 * nx, a.dx, sWork, sHc, T, nx);
    const double pb3 = ep_inner(sB, sB, T + T_M1 * nx, a.dx, sWork, sHc, T, nx);
    if (tid == 0) {
        double err = 0.0;
        for (int i = 0; i < nx; ++i) {
            double d = fabs(sE1[i] - a.grid[S_EPOIS * nx + i]);
            if (d > err) err = d;
        }
        a.diag[1] = red[0];
        a.diag[2] = red[1];
        a.diag[3] = red[2];
        a.diag[4] = pe1;
        a.diag[5] = pe2;
        a.diag[6] = pb3;
        a.diag[7] = red[3];
        a.diag[8] = red[4];
        a.diag[9] = poynting;
        a.diag[10] = err;
    }
    (void)s_red;
}

// ---------------------------------------------------------------------------
// Launch dispatch
// ---------------------------------------------------------------------------
bool mode_has_epi(int m)
{
    return (m & (M_HP2A | M_HP1 | M_HP2B | M_RHO)) != 0 && (m & M_NOEPI) == 0;
}

template <int P0, int P1>
static hipError_t launch_epi_pd(int mode, const PassArgs &a, int pass_blocks, hipStream_t s)
{
#define GK_EPI(M)                                                                             \
    case (M):                                                                                 \
        hipLaunchKernelGGL((k_epi1d<P0, P1, (M)>), dim3(epi_groups(pass_blocks)), dim3(kBlock), \
                           epi_shmem_bytes(a.nx), s, a, pass_blocks);                         \
        return hipGetLastError();
    switch (mode) {
        GK_EPI(M_HP2A)
        GK_EPI(M_HP1)
        GK_EPI(M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_EPI(M_HE_PREV | M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_EPI(M_FAST | M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_EPI(M_FAST | M_HE_PREV | M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_EPI(M_HP1 | M_HP2B | M_WITH_J2A)
        GK_EPI(M_V1 | M_HP1)
        GK_EPI(M_V1 | M_HE | M_HP1)
        GK_EPI(M_V1 | M_HE_PREV | M_HE | M_HP1)
        GK_EPI(M_RHO)
    default:
        return hipErrorInvalidValue;
    }
#undef GK_EPI
}

hipError_t launch_epi(int p0, int p1, int mode, const PassArgs &a, int pass_blocks, hipStream_t s)
{
    if (a.nx > kNxMax || pass_blocks < 1) return hipErrorInvalidValue;
    // the sub-step modes' field parts are the same in either arithmetic
    if ((mode & M_FAST) && (mode & (M_HE | M_HP2A | M_HP1 | M_HP2B)) != (M_HE | M_HP2A | M_HP1 | M_HP2B))
        mode &= ~M_FAST;
#define GK_PAIR(A, B) \
    if (p0 == (A) && p1 == (B)) return launch_epi_pd<A, B>(mode, a, pass_blocks, s);
    GK_PAIR(1, 0)
    GK_PAIR(1, 1)
    GK_PAIR(2, 1)
    GK_PAIR(2, 2)
    GK_PAIR(3, 2)
    GK_PAIR(3, 3)
#undef GK_PAIR
    return hipErrorInvalidValue;
}

template <int P0, int P1>
static hipError_t launch_batch_pd(bool epi, int mode, const BatchArgs &b, int total, size_t shmem,
                                  hipStream_t s)
{
#define GK_B(M)                                                                                  \
    case (M):                                                                                    \
        if (epi) {                                                                               \
            hipLaunchKernelGGL((k_epi1d_batch_sum<P0, P1, (M)>), dim3(total), dim3(kBlock), shmem, s, b); \
            hipLaunchKernelGGL((k_epi1d_batch_fin<P0, P1, (M)>), dim3(b.ninst), dim3(kBlock), shmem, s, b); \
        } else {                                                                                 \
            hipLaunchKernelGGL((k_pass1d_batch<P0, P1, (M)>), dim3(total), dim3(kBlock), shmem, s, b); \
        }                                                                                        \
        return hipGetLastError();
    switch (mode) {
        GK_B(M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_B(M_HE_PREV | M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_B(M_FAST | M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_B(M_FAST | M_HE_PREV | M_HE | M_HP2A | M_HP1 | M_HP2B)
    case M_HE_PREV:
        if (epi) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_pass1d_batch<P0, P1, M_HE_PREV>), dim3(total), dim3(kBlock), shmem, s, b);
        return hipGetLastError();
    case M_FAST | M_HE_PREV:
        if (epi) return hipErrorInvalidValue;
        hipLaunchKernelGGL((k_pass1d_batch<P0, P1, M_FAST | M_HE_PREV>), dim3(total), dim3(kBlock), shmem, s, b);
        return hipGetLastError();
    default:
        return hipErrorInvalidValue;
    }
#undef GK_B
}

static hipError_t launch_batch(bool epi, int p0, int p1, int mode, const BatchArgs &b, int total,
                               size_t shmem, hipStream_t s)
{
    if (total < 1 || b.ninst < 1) return hipErrorInvalidValue;
#define GK_PAIR(A, B) \
    if (p0 == (A) && p1 == (B)) return launch_batch_pd<A, B>(epi, mode, b, total, shmem, s);
    GK_PAIR(1, 0)
    GK_PAIR(1, 1)
    GK_PAIR(2, 1)
    GK_PAIR(2, 2)
    GK_PAIR(3, 2)
    GK_PAIR(3, 3)
#undef GK_PAIR
    return hipErrorInvalidValue;
}

hipError_t launch_pass_batch(int p0, int p1, int mode, const BatchArgs &b, int total_blocks,
                             size_t shmem, hipStream_t s)
{
    return launch_batch(false, p0, p1, mode, b, total_blocks, shmem, s);
}

hipError_t launch_epi_batch(int p0, int p1, int mode, const BatchArgs &b, int total_groups,
                            size_t shmem, hipStream_t s)
{
    return launch_batch(true, p0, p1, mode, b, total_groups, shmem, s);
}

template <int P0, int P1>
static hipError_t launch_pass_pd(int mode, const PassArgs &a, int blocks, size_t shmem,
                                 hipStream_t s)
{
    // blocks < 0: query how many workgroups of this kernel fit on one CU
#define GK_CASE(M)                                                                          \
    case (M):                                                                               \
        if (blocks < 0) {                                                                   \
            int nb = 0;                                                                     \
            hipError_t e_ = hipOccupancyMaxActiveBlocksPerMultiprocessor(                   \
                &nb, k_pass1d<P0, P1, (M)>, kBlock, shmem);                                 \
            *(int *)a.diag = nb;                                                            \
            return e_;                                                                      \
        }                                                                                   \
        hipLaunchKernelGGL((k_pass1d<P0, P1, (M)>), dim3(blocks), dim3(kBlock), shmem, s, a); \
        return hipGetLastError();
    switch (mode) {
        GK_CASE(M_HE)
        GK_CASE(M_HP2A)
        GK_CASE(M_HP1)
        GK_CASE(M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_CASE(M_HE_PREV | M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_CASE(M_HE_PREV)
        GK_CASE(M_HE | M_HP2A | M_NOEPI)
        GK_CASE(M_HE_PREV | M_HE | M_HP2A | M_NOEPI)
        GK_CASE(M_HP1 | M_HP2B | M_WITH_J2A)
        GK_CASE(M_V1 | M_HE)
        GK_CASE(M_V1 | M_HP1)
        GK_CASE(M_V1 | M_HE | M_HP1)
        GK_CASE(M_V1 | M_HE_PREV | M_HE | M_HP1)
        GK_CASE(M_V1 | M_HE_PREV)
        GK_CASE(M_RHO)
        GK_CASE(M_FAST | M_HE)
        GK_CASE(M_FAST | M_HP2A)
        GK_CASE(M_FAST | M_HP1)
        GK_CASE(M_FAST | M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_CASE(M_FAST | M_HE_PREV | M_HE | M_HP2A | M_HP1 | M_HP2B)
        GK_CASE(M_FAST | M_HE_PREV)
    case M_DIAG:
        if (blocks < 0) {
            int nb = 0;
            hipError_t e_ = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_diag1d2v<P0, P1>,
                                                                         kBlock, shmem);
            *(int *)a.diag = nb;
            return e_;
        }
        hipLaunchKernelGGL((k_diag1d2v<P0, P1>), dim3(blocks), dim3(kBlock), shmem, s, a);
        return hipGetLastError();
    default:
        return hipErrorInvalidValue;
    }
#undef GK_CASE
}

int pass_blocks_per_cu(int p0, int p1, int mode, size_t shmem)
{
    PassArgs a;
    int nb = 0;
    a.diag = (double *)&nb;
    if (launch_pass(p0, p1, mode, a, -1, shmem, nullptr) != hipSuccess || nb < 1) return 1;
    return nb;
}

hipError_t launch_pass(int p0, int p1, int mode, const PassArgs &a, int blocks, size_t shmem,
                       hipStream_t s)
{
#define GK_PAIR(A, B) \
    if (p0 == (A) && p1 == (B)) return launch_pass_pd<A, B>(mode, a, blocks, shmem, s);
    GK_PAIR(1, 0)
    GK_PAIR(1, 1)
    GK_PAIR(2, 1)
    GK_PAIR(2, 2)
    GK_PAIR(3, 2)
    GK_PAIR(3, 3)
#undef GK_PAIR
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// HamiltonianSplittingBoris (hamiltonian_splitting_boris.jl): the particle
// work of one strang_splitting! step is one pass (push_v_epart!(dt/2),
// push_v_bpart!(dt), push_v_epart!(dt/2), push_x_accumulate_j!(dt), :146-155):
// the E and B kicks read the mid-step fields at the old position, then the
// position moves and the currents are deposited at the midpoint. OPS selects
// the pushers for the stand-alone verbs.
// ---------------------------------------------------------------------------
template <int P0, int P1, int OPS>
__global__ void __launch_bounds__(kBlock) k_boris1d(PassArgs a)
{
    constexpr bool EPART = (OPS & BO_EPART) != 0;
    constexpr bool BPART = (OPS & BO_BPART) != 0;
    constexpr bool EPART2 = (OPS & BO_EPART2) != 0;
    constexpr bool PUSHX = (OPS & BO_PUSHX) != 0;
    constexpr bool KICK = EPART || BPART || EPART2;
    constexpr int W0 = P0 + 1, W1 = P1 + 1;
    static_assert(P0 <= kPad && P1 <= kPad, "LDS field padding must cover the spline support");
    static_assert(W0 + W1 <= kSlotsPerLane, "cache-B slots");

    extern __shared__ double smem[];
    const int nx = a.nx;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int nxp = nx + 2 * kPad;
    double *pE1 = smem, *pE2 = pE1 + nxp, *pB = pE2 + nxp;     // padded E1_mid, E2_mid, B_mid
    double *rows = pB + nxp;                                     // [kWaves][NDEP][nx]
    double *slots = rows + kWaves * NDEP * nx + wave * (kSlotsPerLane * kWave);
    const double *sE1 = pE1 + kPad, *sE2 = pE2 + kPad, *sB = pB + kPad;
    if constexpr (KICK) {
        load_padded(pE1, a.grid + S_E1M * nx, nx);
        load_padded(pE2, a.grid + S_E2M * nx, nx);
        load_padded(pB, a.grid + S_BM * nx, nx);
    }
    for (int i = tid; i < kWaves * NDEP * nx; i += kBlock) rows[i] = 0.0;
    __syncthreads();
    double *myrow = rows + wave * NDEP * nx;
    Cache2<W1> c_j1;
    Cache2<W0> c_j2;
    if constexpr (PUSHX) {
        c_j1.init(slots);
        c_j2.init(slots + W1 * kWave);
    }

    const double xmin = a.xmin, dx = a.dx;
    const double rdx = 1.0 / dx;     // div_by: (x - xmin) / dx, correctly rounded
    const PassChunk pc = pass_chunk(a, blockIdx.x);
    const int64_t base = pc.base;
    const int cnt_here = pc.cnt;
    const int iters = pc.iters;
    double *__restrict__ px = a.x + base;
    double *__restrict__ pv1 = a.v1 + base;
    double *__restrict__ pv2 = a.v2 + base;
    const double *__restrict__ pw = a.w + base;
    int errs = 0;

    double nx_x = 0.0, nx_v1 = 0.0, nx_v2 = 0.0, nx_w = 0.0;
    if (iters > 0 && tid < cnt_here) {
        nx_x = elem(px, tid);
        nx_v1 = elem(pv1, tid);
        nx_v2 = elem(pv2, tid);
        if (PUSHX) nx_w = elem(pw, tid);
    }
    for (int it = 0; it < iters; ++it) {
        const int i = it * kBlock + tid;
        const bool valid = i < cnt_here;
        double x = nx_x, v1 = nx_v1, v2 = nx_v2;
        const double w = nx_w;
        {
            const int in = i + kBlock;
            if (it + 1 < iters && in < cnt_here) {
                nx_x = elem(px, in);
                nx_v1 = elem(pv1, in);
                nx_v2 = elem(pv2, in);
                if (PUSHX) nx_w = elem(pw, in);
            }
        }
        if constexpr (KICK) {
            // evaluate (particle_mesh_coupling_1d.jl:438-453) at the old position
            const double xi = div_by(x - xmin, dx, rdx);
            if (valid && !pos_ok(xi)) errs |= E_INEXACT;
            const int ic = (int)xi;
            const double off = xi - (double)ic;
            const int key = imod(ic, nx);
            double b1[W1];
            bspline<P1>(off, b1);
            double e1 = 0.0, e2 = 0.0;
            if constexpr (EPART || EPART2) {
                double b0[W0];
                bspline<P0>(off, b0);
                e1 = gather_p<P1>(b1, sE1, key);
                e2 = gather_p<P0>(b0, sE2, key);
            }
            if constexpr (EPART) {
                // push_v_epart! (:189-204): v .+ dt * qm .* [E1, E2]
                v1 = v1 + a.dtqm_he * e1;
                v2 = v2 + a.dtqm_he * e2;
            }
            if constexpr (BPART) {
                // push_v_bpart! (:211-233)
                double bf = gather_p<P1>(b1, sB, key);
                bf = a.dtqm_hp2 * bf;
                double m11 = 1.0 / (1.0 + bf * bf);
                const double m12 = m11 * bf * 2.0;
                m11 = m11 * (1.0 - bf * bf);
                const double n1 = m11 * v1 + m12 * v2;
                const double n2 = -m12 * v1 + m11 * v2;
                v1 = n1;
                v2 = n2;
            }
            if constexpr (EPART2) {
                v1 = v1 + a.dtqm_he * e1;
                v2 = v2 + a.dtqm_he * e2;
            }
        }
        if constexpr (PUSHX) {
            // push_x_accumulate_j! (:250-288): add_charge! of q v1 (degree P1)
            // and q v2 (degree P0) at the midpoint, then x = mod(x_new, Lx)
            const double q = (w * a.charge) * a.cw;
            const double xn = x + a.dt_hp1 * v1;
            const double xm = (x + xn) * 0.5;
            const double xim = div_by(xm - xmin, dx, rdx);
            if (valid && !pos_ok(xim)) errs |= E_INEXACT;
            const int icm = (int)xim;                    // trunc(Int, xi)
            const double offm = xim - (double)icm;
            const int keym = imod(icm, nx);
            double c1[W1], c0[W0];
            bspline<P1>(offm, c1);
            bspline<P0>(offm, c0);
            const double q1 = q * v1, q2 = q * v2;
            double vals1[W1], vals0[W0];
#pragma unroll
            for (int k = 0; k < W1; ++k) vals1[k] = q1 * c1[k] * a.scaling1;
#pragma unroll
            for (int k = 0; k < W0; ++k) vals0[k] = q2 * c0[k] * a.scaling0;
            cache_add<W1>(c_j1, valid, keym, vals1, myrow + D_J1 * nx, nx);
            cache_add<W0>(c_j2, valid, keym, vals0, myrow + D_J2A * nx, nx);
            x = jl_mod(xn, a.Lx);
        }
        if (valid) {
            if constexpr (PUSHX) elem(px, i) = x;
            if constexpr (KICK) {
                elem(pv1, i) = v1;
                elem(pv2, i) = v2;
            }
        }
    }
    if (errs) atomicOr(a.err, errs);
    if constexpr (PUSHX) {
        cache_drain<W1>(c_j1, myrow + D_J1 * nx, nx);
        cache_drain<W0>(c_j2, myrow + D_J2A * nx, nx);
        __syncthreads();
        constexpr unsigned MASK = (1u << D_J1) | (1u << D_J2A);
        const int len = NDEP * nx;
        for (int idx = tid; idx < len; idx += kBlock) {
            if (!((MASK >> (idx / nx)) & 1u)) continue;
            double t = 0.0;
            for (int wv = 0; wv < kWaves; ++wv) t += rows[wv * len + idx];
            a.part[(size_t)blockIdx.x * len + idx] = t;
        }
    }
}

// The field part after a Boris x push: the deposit sum over the workgroups'
// rows (as k_epi1d: group slices, then the last workgroup), then
//  BE_J:       j_dofs = the deposits (push_x_accumulate_j!);
//  BE_STAGGER: staggering! (:103-121) with h = a.dt_hp1: E_mid = E,
//              j = h j, E1_mid -= M1^-1 j1/dx, E2_mid += h W(B),
//              E2_mid -= M0^-1 j2/dx;
//  BE_STEP:    strang_splitting! (:158-175) with dt = a.dt_hp1: E = E_mid,
//              j = dt j, the same E_mid updates with dt, and with a.more the
//              next step's B_mid = (B + b_from_e(B, dt, E2_mid)) / 2 (:138-144).
template <int KIND>
__global__ void __launch_bounds__(kBlock) k_boris_epi(PassArgs a, int nblocks)
{
    extern __shared__ double smem[];
    __shared__ int s_last;
    constexpr unsigned REDMASK = (1u << D_J1) | (1u << D_J2A);
    const int nx = a.nx;
    const int tid = threadIdx.x;
    const int nt = blockDim.x;
    const int len = NDEP * nx;
    const double dx = a.dx, h = a.dt_hp1;
    double *sE1 = smem;                     // [nx] each: E1_mid (or E1), E2_mid (or E2), B, B_mid
    double *sE2 = sE1 + nx;
    double *sB = sE2 + nx;
    double *sBM = sB + nx;
    double *sRed = sBM + nx;                // [NDEP][nx]
    double *work = sRed + NDEP * nx;
    double *hc = work + nx;
    double *T = hc + nx;                    // [NTAB][nx]
    double *w3 = T + NTAB * nx;             // [6][nx]: outputs and workspaces of the solves
    double *scratch = w3 + 6 * nx;          // [blockDim.x]
    double *G = a.grid;
    const int se1 = KIND == BE_STAGGER ? S_E1 : S_E1M, se2 = KIND == BE_STAGGER ? S_E2 : S_E2M;
    for (int i = tid; i < nx; i += nt) {
        sE1[i] = G[se1 * nx + i];
        sE2[i] = G[se2 * nx + i];
        sB[i] = G[S_B * nx + i];
    }
    for (int i = tid; i < NTAB * nx; i += nt) T[i] = a.tabs[i];
    for (int i = tid; i < len; i += nt) sRed[i] = 0.0;

    const int ng = gridDim.x, g = blockIdx.x;
    const int b0 = (int)(((long)nblocks * g) / ng), b1 = (int)(((long)nblocks * (g + 1)) / ng);
    sum_rows_range(a.part, b0, b1, nx, REDMASK, a.gpart + (size_t)g * len, scratch);
    __threadfence();
    __syncthreads();
    if (tid == 0) s_last = atomicAdd(&a.cnt[0], 1u) == (unsigned)(ng - 1);
    __syncthreads();
    if (!s_last) return;
    __threadfence();
    if (tid == 0) a.cnt[0] = 0u;
    sum_rows_range(a.gpart, 0, ng, nx, REDMASK, sRed, scratch);
    double *r_j1 = sRed + D_J1 * nx, *r_j2 = sRed + D_J2A * nx;
    if constexpr (KIND == BE_J) {
        for (int i = tid; i < nx; i += nt) {
            G[S_J1 * nx + i] = r_j1[i];
            G[S_J2 * nx + i] = r_j2[i];
        }
        return;
    } else {
        for (int i = tid; i < nx; i += nt) {
            if (KIND == BE_STEP) {               // e_dofs .= e_dofs_mid
                G[S_E1 * nx + i] = sE1[i];
                G[S_E2 * nx + i] = sE2[i];
            }
            r_j1[i] = h * r_j1[i];
            r_j2[i] = h * r_j2[i];
        }
        __syncthreads();
        // the three solves (M1^-1 j1, W(B), M0^-1 j2) are independent: run
        // them side by side, then apply them in the reference's order
        // (compute_e_from_j!(E1), compute_e_from_b!(E2), compute_e_from_j!(E2))
        const SolveArgs s1{T + T_INVM1 * nx, r_j1, 1.0, w3, w3 + nx};
        const SolveArgs sb{T + T_AMP * nx, sB, 1.0, w3 + 2 * nx, w3 + 3 * nx};
        const SolveArgs s2{T + T_INVM0 * nx, r_j2, 1.0, w3 + 4 * nx, w3 + 5 * nx};
        circ_solve3(s1, sb, s2, T + T_COS * nx, T + T_SIN * nx, nx);
        const double coef = h / dx;
        for (int i = tid; i < nx; i += nt) {
            double w1 = s1.out[i] / dx;
            sE1[i] -= w1;
            sE2[i] += coef * sb.out[i];
            double w2 = s2.out[i] / dx;
            sE2[i] -= w2;
        }
        __syncthreads();
        if (KIND == BE_STEP && a.more) {
            for (int i = tid; i < nx; i += nt) sBM[i] = sB[i];
            __syncthreads();
            ep_b_from_e(sB, sE2, h, dx, nx);
            for (int i = tid; i < nx; i += nt) sBM[i] = (sBM[i] + sB[i]) * 0.5;
            __syncthreads();
            for (int i = tid; i < nx; i += nt) {
                G[S_B * nx + i] = sB[i];
                G[S_BM * nx + i] = sBM[i];
            }
        }
        for (int i = tid; i < nx; i += nt) {
            G[S_J1 * nx + i] = r_j1[i];
            G[S_J2 * nx + i] = r_j2[i];
            G[S_E1M * nx + i] = sE1[i];
            G[S_E2M * nx + i] = sE2[i];
        }
    }
}

size_t boris_shmem_bytes(int nx)
{
    return sizeof(double) * ((size_t)3 * (nx + 2 * kPad) + (size_t)kWaves * NDEP * nx +
                             (size_t)kWaves * kSlotsPerLane * kWave);
}

static size_t boris_epi_shmem_bytes(int nx)
{
    return sizeof(double) * ((size_t)(4 + NDEP + 2 + NTAB + 6) * nx + kBlock);
}

template <int P0, int P1>
static hipError_t launch_boris_pd(int ops, const PassArgs &a, int blocks, size_t shmem, hipStream_t s)
{
#define GK_BO(O)                                                                             \
    case (O):                                                                                \
        if (blocks < 0) {                                                                    \
            int nb = 0;                                                                      \
            hipError_t e_ = hipOccupancyMaxActiveBlocksPerMultiprocessor(                    \
                &nb, k_boris1d<P0, P1, (O)>, kBlock, shmem);                                 \
            *(int *)a.diag = nb;                                                             \
            return e_;                                                                       \
        }                                                                                    \
        hipLaunchKernelGGL((k_boris1d<P0, P1, (O)>), dim3(blocks), dim3(kBlock), shmem, s, a); \
        return hipGetLastError();
    switch (ops) {
        GK_BO(BO_EPART | BO_BPART | BO_EPART2 | BO_PUSHX)
        GK_BO(BO_PUSHX)
        GK_BO(BO_EPART)
        GK_BO(BO_BPART)
    default:
        return hipErrorInvalidValue;
    }
#undef GK_BO
}

hipError_t launch_boris(int p0, int p1, int ops, const PassArgs &a, int blocks, size_t shmem,
                        hipStream_t s)
{
    if (a.nx > kNxMax) return hipErrorInvalidValue;
#define GK_PAIR(A, B) \
    if (p0 == (A) && p1 == (B)) return launch_boris_pd<A, B>(ops, a, blocks, shmem, s);
    GK_PAIR(1, 0)
    GK_PAIR(1, 1)
    GK_PAIR(2, 1)
    GK_PAIR(2, 2)
    GK_PAIR(3, 2)
    GK_PAIR(3, 3)
#undef GK_PAIR
    return hipErrorInvalidValue;
}

int boris_blocks_per_cu(int p0, int p1, size_t shmem)
{
    PassArgs a;
    int nb = 0;
    a.diag = (double *)&nb;
    a.nx = 1;
    if (launch_boris(p0, p1, BO_EPART | BO_BPART | BO_EPART2 | BO_PUSHX, a, -1, shmem, nullptr) !=
            hipSuccess || nb < 1)
        return 1;
    return nb;
}

hipError_t launch_boris_epi(int kind, const PassArgs &a, int pass_blocks, hipStream_t s)
{
    if (a.nx > kNxMax || pass_blocks < 1) return hipErrorInvalidValue;
    const dim3 grid(epi_groups(pass_blocks));
    const size_t sh = boris_epi_shmem_bytes(a.nx);
    switch (kind) {
    case BE_J: hipLaunchKernelGGL(k_boris_epi<BE_J>, grid, dim3(kBlock), sh, s, a, pass_blocks); break;
    case BE_STAGGER: hipLaunchKernelGGL(k_boris_epi<BE_STAGGER>, grid, dim3(kBlock), sh, s, a, pass_blocks); break;
    case BE_STEP: hipLaunchKernelGGL(k_boris_epi<BE_STEP>, grid, dim3(kBlock), sh, s, a, pass_blocks); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Single-workgroup field programs (operatorHB, strang pre-step, Maxwell verbs)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_grid1d(GridProg g, int nx, double dx, double *grid,
                                                   const double *tabs, double *scalar_out)
{
    __shared__ double sF[NSLOT * kNxMax];
    __shared__ double sWork[kNxMax], sHc[kNxMax];
    __shared__ double sT[NTAB * kNxMax];
    const int tid = threadIdx.x;
    for (int i = tid; i < NSLOT * nx; i += kBlock) sF[i] = grid[i];
    load_tabs(tabs, sT, nx);
    tabs = sT;
    for (int o = 0; o < g.nops; ++o) {
        const double dt = g.dt[o];
        double *A = sF + g.slot_a[o] * nx;
        double *B = sF + g.slot_b[o] * nx;
        double *O = sF + g.slot_out[o] * nx;
        switch (g.op[o]) {
        case G_HB: ep_e_from_b(O, A, dt, dx, sWork, sHc, tabs, nx); break;
        case G_HE_B: ep_b_from_e(O, A, dt, dx, nx); break;
        case G_E_FROM_J1: ep_e_from_j(O, A, tabs + T_INVM1 * nx, dx, sWork, sHc, tabs, nx); break;
        case G_E_FROM_J2: ep_e_from_j(O, A, tabs + T_INVM0 * nx, dx, sWork, sHc, tabs, nx); break;
        case G_E_FROM_RHO: ep_e_from_rho(O, A, sWork, sHc, tabs, nx); break;
        case G_INNER_M0: {
            double r = ep_inner(A, B, tabs + T_M0 * nx, dx, sWork, sHc, tabs, nx);
            if (tid == 0 && scalar_out) scalar_out[o] = r;
            break;
        }
        case G_INNER_M1: {
            double r = ep_inner(A, B, tabs + T_M1 * nx, dx, sWork, sHc, tabs, nx);
            if (tid == 0 && scalar_out) scalar_out[o] = r;
            break;
        }
        case G_COPY:
            for (int i = tid; i < nx; i += kBlock) O[i] = A[i];
            break;
        case G_AVG:
            for (int i = tid; i < nx; i += kBlock) O[i] = (O[i] + A[i]) * 0.5;
            break;
        default: break;
        }
        __syncthreads();
    }
    for (int i = tid; i < NSLOT * nx; i += kBlock) grid[i] = sF[i];
}

hipError_t launch_grid(const GridProg &g, int nx, double dx, double *grid, const double *tabs,
                       double *scalar_out, hipStream_t s)
{
    if (nx > kNxMax) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_grid1d, dim3(1), dim3(kBlock), 0, s, g, nx, dx, grid, tabs, scalar_out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Stable counting sort by (cell, sub-cell bin), wave tiles of kSortTile
// particles. Five launches: per-tile bin counts; per-group sums over
// tiles_per_group tiles; bin bases and group bases (one workgroup); per-tile
// offsets; scatter. Every step is a fixed function of the input order, so the
// permutation is deterministic; particles of one bin keep their order.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int sort_key(double x, double xmin, double dx, int nx, int sub)
{
    double xi = (x - xmin) / dx;
    if (!(fabs(xi) < 1.0e9)) xi = 0.0;
    const int ic = (int)xi;                                  // the cell the passes use
    int s = (int)((xi - (double)ic) * (double)sub);
    s = s < 0 ? 0 : (s >= sub ? sub - 1 : s);
    return imod(ic, nx) * sub + s;
}

// Each wave owns one tile and walks it in batches of kSortBatch iterations
// whose loads are issued together.
constexpr int kSortBatch = 8;

__device__ __forceinline__ void sort_count_body(const SortArgs &a, int bx, int *ish)
{
    const int wave = threadIdx.x / kWave;
    const int lane = lane_id();
    const int nb = a.nbins;
    int *h = ish + wave * nb;
    const int tile = bx * kWaves + wave;
    for (int c = lane; c < nb; c += kWave) h[c] = 0;
    __syncthreads();
    if (tile < a.ntiles) {
        const int64_t t0 = (int64_t)tile * kSortTile;
        for (int k0 = 0; k0 < kSortTile / kWave; k0 += kSortBatch) {
            double xs[kSortBatch];
#pragma unroll
            for (int u = 0; u < kSortBatch; ++u) {
                const int64_t i = t0 + (int64_t)(k0 + u) * kWave + lane;
                xs[u] = i < a.n ? a.x[i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kSortBatch; ++u) {
                const int64_t i = t0 + (int64_t)(k0 + u) * kWave + lane;
                const bool valid = i < a.n;
                const int c = valid ? sort_key(xs[u], a.xmin, a.dx, a.nx, a.sub) : -1;
                unsigned long long pending = __ballot(valid);
                while (pending) {
                    const int leader = __ffsll((long long)pending) - 1;
                    const int cc = __shfl(c, leader, kWave);
                    const unsigned long long m = __ballot(valid && c == cc);
                    if (lane == leader) h[cc] += __popcll(m);
                    pending &= ~m;
                }
            }
        }
        int *row = a.counts + (size_t)tile * nb;
        for (int c = lane; c < nb; c += kWave) row[c] = h[c];
    }
}

// gsum[g][b] = sum of counts[t][b] over the tiles t of group g
__device__ __forceinline__ void sort_gsum_body(const SortArgs &a, int g, int by)
{
    const int nb = a.nbins;
    const int b = by * kBlock + threadIdx.x;
    if (b >= nb) return;
    const int t0 = g * a.tiles_per_group;
    const int t1 = min(a.ntiles, t0 + a.tiles_per_group);
    const int *col = a.counts + b;
    int s = 0;
    int t = t0;
    for (; t + kSortBatch <= t1; t += kSortBatch) {
        int v[kSortBatch];
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) v[u] = col[(size_t)(t + u) * nb];
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) s += v[u];
    }
    for (; t < t1; ++t) s += col[(size_t)t * nb];
    a.gsum[(size_t)g * nb + b] = s;
}

// One workgroup: bin totals, their exclusive scan (the bin bases), and
// gsum[g][b] <- base[b] + sum of gsum[g'][b] over g' < g. Each thread walks
// its bin's column of group sums twice, kBasesBatch loads in flight.
constexpr int kBasesBatch = 32;
__device__ __forceinline__ void sort_bases_body(const SortArgs &a, int *s_scan)
{
    const int nb = a.nbins;
    const int b = threadIdx.x;
    int *col = a.gsum + b;
    int total = 0;
    if (b < nb) {
        int g = 0;
        for (; g + kBasesBatch <= a.ngroups; g += kBasesBatch) {
            int v[kBasesBatch];
#pragma unroll
            for (int u = 0; u < kBasesBatch; ++u) v[u] = col[(size_t)(g + u) * nb];
#pragma unroll
            for (int u = 0; u < kBasesBatch; ++u) total += v[u];
        }
        for (; g < a.ngroups; ++g) total += col[(size_t)g * nb];
    }
    s_scan[b] = b < nb ? total : 0;
    __syncthreads();
    for (int off = 1; off < kSortBinsMax; off <<= 1) {
        const int add = b >= off ? s_scan[b - off] : 0;
        __syncthreads();
        s_scan[b] += add;
        __syncthreads();
    }
    if (b >= nb) return;
    int run = s_scan[b] - total;
    int g = 0;
    for (; g + kBasesBatch <= a.ngroups; g += kBasesBatch) {
        int v[kBasesBatch];
#pragma unroll
        for (int u = 0; u < kBasesBatch; ++u) v[u] = col[(size_t)(g + u) * nb];
#pragma unroll
        for (int u = 0; u < kBasesBatch; ++u) {
            col[(size_t)(g + u) * nb] = run;
            run += v[u];
        }
    }
    for (; g < a.ngroups; ++g) {
        const int v = col[(size_t)g * nb];
        col[(size_t)g * nb] = run;
        run += v;
    }
}

// counts[t][b] <- the first output position of bin b's particles of tile t
__device__ __forceinline__ void sort_offsets_body(const SortArgs &a, int g, int by)
{
    const int nb = a.nbins;
    const int b = by * kBlock + threadIdx.x;
    if (b >= nb) return;
    const int t0 = g * a.tiles_per_group;
    const int t1 = min(a.ntiles, t0 + a.tiles_per_group);
    int *col = a.counts + b;
    int run = a.gsum[(size_t)g * nb + b];
    int t = t0;
    for (; t + kSortBatch <= t1; t += kSortBatch) {
        int v[kSortBatch];
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) v[u] = col[(size_t)(t + u) * nb];
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) {
            col[(size_t)(t + u) * nb] = run;
            run += v[u];
        }
    }
    for (; t < t1; ++t) {
        const int v = col[(size_t)t * nb];
        col[(size_t)t * nb] = run;
        run += v;
    }
}

__device__ __forceinline__ void sort_scatter_body(const SortArgs &a, int has_v2, int bx, int *ish)
{
    const int wave = threadIdx.x / kWave;
    const int lane = lane_id();
    const int nb = a.nbins;
    int *cur = ish + wave * nb;
    const int tile = bx * kWaves + wave;
    if (tile < a.ntiles) {
        const int *row = a.counts + (size_t)tile * nb;
        for (int c = lane; c < nb; c += kWave) cur[c] = row[c];
    }
    __syncthreads();
    if (tile >= a.ntiles) return;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t t0 = (int64_t)tile * kSortTile;
    for (int k0 = 0; k0 < kSortTile / kWave; k0 += kSortBatch) {
        double xs[kSortBatch], v1s[kSortBatch], v2s[kSortBatch], ws[kSortBatch];
        int64_t ids[kSortBatch];
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) {
            const int64_t i = t0 + (int64_t)(k0 + u) * kWave + lane;
            const bool valid = i < a.n;
            xs[u] = valid ? a.x[i] : 0.0;
            v1s[u] = valid ? a.v1[i] : 0.0;
            v2s[u] = (valid && has_v2) ? a.v2[i] : 0.0;
            ws[u] = valid ? a.w[i] : 0.0;
            ids[u] = valid ? a.id[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < kSortBatch; ++u) {
            const int64_t i = t0 + (int64_t)(k0 + u) * kWave + lane;
            const bool valid = i < a.n;
            const int c = valid ? sort_key(xs[u], a.xmin, a.dx, a.nx, a.sub) : -1;
            unsigned long long pending = __ballot(valid);
            int pos = 0;
            while (pending) {
                const int leader = __ffsll((long long)pending) - 1;
                const int cc = __shfl(c, leader, kWave);
                const unsigned long long m = __ballot(valid && c == cc);
                const int basepos = cur[cc];
                if (valid && c == cc) pos = basepos + __popcll(m & lt);
                if (lane == leader) cur[cc] = basepos + __popcll(m);
                pending &= ~m;
            }
            if (valid) {
                a.x2[pos] = xs[u];
                a.v12[pos] = v1s[u];
                if (has_v2) a.v22[pos] = v2s[u];
                a.w2[pos] = ws[u];
                a.id2[pos] = ids[u];
                if (a.perm) a.perm[pos] = (int)i;
            }
        }
    }
}

__global__ void __launch_bounds__(kBlock) k_sort_count(SortArgs a)
{
    extern __shared__ int ish[];
    sort_count_body(a, blockIdx.x, ish);
}

__global__ void __launch_bounds__(kBlock) k_sort_gsum(SortArgs a)
{
    sort_gsum_body(a, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(kSortBinsMax) k_sort_bases(SortArgs a)
{
    __shared__ int s_scan[kSortBinsMax];
    sort_bases_body(a, s_scan);
}

__global__ void __launch_bounds__(kBlock) k_sort_offsets(SortArgs a)
{
    sort_offsets_body(a, blockIdx.x, blockIdx.y);
}

__global__ void __launch_bounds__(kBlock) k_sort_scatter(SortArgs a, int has_v2)
{
    extern __shared__ int ish[];
    sort_scatter_body(a, has_v2, blockIdx.x, ish);
}

// The sorts of every instance of an ensemble, one launch per stage: tile
// workgroups of instance i are [tile_start[i], tile_start[i+1]), group
// columns [group_start[i], group_start[i+1]) (y: bins / kBlock, at most that
// of the largest grid), one workgroup per instance for the bases.
__global__ void __launch_bounds__(kBlock) k_sort_count_batch(SortBatch b)
{
    extern __shared__ int ish[];
    const int i = batch_instance(b.tile_start, b.ninst, blockIdx.x);
    sort_count_body(b.args[i], blockIdx.x - b.tile_start[i], ish);
}

__global__ void __launch_bounds__(kBlock) k_sort_gsum_batch(SortBatch b)
{
    const int i = batch_instance(b.group_start, b.ninst, blockIdx.x);
    sort_gsum_body(b.args[i], blockIdx.x - b.group_start[i], blockIdx.y);
}

__global__ void __launch_bounds__(kSortBinsMax) k_sort_bases_batch(SortBatch b)
{
    __shared__ int s_scan[kSortBinsMax];
    sort_bases_body(b.args[blockIdx.x], s_scan);
}

__global__ void __launch_bounds__(kBlock) k_sort_offsets_batch(SortBatch b)
{
    const int i = batch_instance(b.group_start, b.ninst, blockIdx.x);
    sort_offsets_body(b.args[i], blockIdx.x - b.group_start[i], blockIdx.y);
}

__global__ void __launch_bounds__(kBlock) k_sort_scatter_batch(SortBatch b)
{
    extern __shared__ int ish[];
    const int i = batch_instance(b.tile_start, b.ninst, blockIdx.x);
    sort_scatter_body(b.args[i], 1, blockIdx.x - b.tile_start[i], ish);
}

hipError_t launch_sort_batch(const SortBatch &b, int total_tiles, int total_groups, int max_bins,
                             hipStream_t s)
{
    if (b.ninst < 1 || total_tiles < 1 || total_groups < 1 || max_bins > kSortBinsMax)
        return hipErrorInvalidValue;
    const size_t sh = sizeof(int) * (size_t)kWaves * max_bins;
    const dim3 cols(total_groups, (max_bins + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_sort_count_batch, dim3(total_tiles), dim3(kBlock), sh, s, b);
    hipLaunchKernelGGL(k_sort_gsum_batch, cols, dim3(kBlock), 0, s, b);
    hipLaunchKernelGGL(k_sort_bases_batch, dim3(b.ninst), dim3(kSortBinsMax), 0, s, b);
    hipLaunchKernelGGL(k_sort_offsets_batch, cols, dim3(kBlock), 0, s, b);
    hipLaunchKernelGGL(k_sort_scatter_batch, dim3(total_tiles), dim3(kBlock), sh, s, b);
    return hipGetLastError();
}

SortGeometry sort_geometry(int64_t n, int nx)
{
    SortGeometry g;
    g.sub = nx > 0 ? std::max(1, std::min(32, kSortBinsMax / nx)) : 1;
    g.nbins = nx * g.sub;
    g.ntiles = (int)std::max<int64_t>(1, (n + kSortTile - 1) / kSortTile);
    int tpg = (int)std::ceil(std::sqrt((double)g.ntiles));
    g.tiles_per_group = std::max(16, tpg);
    g.ngroups = (g.ntiles + g.tiles_per_group - 1) / g.tiles_per_group;
    return g;
}

hipError_t launch_sort(const SortArgs &a, int has_v2, hipStream_t s)
{
    if (a.nx > kNxMax || a.nbins > kSortBinsMax || a.nbins != a.nx * a.sub || a.ntiles < 1 ||
        (int64_t)a.ntiles * kSortTile < a.n || a.ngroups < 1 ||
        (int64_t)a.ngroups * a.tiles_per_group < a.ntiles)
        return hipErrorInvalidValue;
    const int blocks = (a.ntiles + kWaves - 1) / kWaves;
    const size_t sh = sizeof(int) * (size_t)kWaves * a.nbins;
    const dim3 cols(a.ngroups, (a.nbins + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_sort_count, dim3(blocks), dim3(kBlock), sh, s, a);
    hipLaunchKernelGGL(k_sort_gsum, cols, dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_sort_bases, dim3(1), dim3(kSortBinsMax), 0, s, a);
    hipLaunchKernelGGL(k_sort_offsets, cols, dim3(kBlock), 0, s, a);
    hipLaunchKernelGGL(k_sort_scatter, dim3(blocks), dim3(kBlock), sh, s, a, has_v2);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Layout transposes between the reference's particle matrix and SoA
// ---------------------------------------------------------------------------
// sample!(pg, alpha, k, sigma, mesh) (particle_sampling.jl:266-311): particle
// i (1-based) from the i-th point of Sobol.jl's SobolSeq(2) (Gray-code order,
// the zero point skipped; dimension 1 van der Corput, dimension 2 from the
// polynomial x + 1), its speed sigma sqrt(-2 log((i - 1/2)/n)) at angle
// 2 pi r1, and x = newton(r2, alpha, k) (:236-245, the reference's two
// interleaved iterations, at most 200).
__device__ void sobol2(uint64_t k, double &r1, double &r2)
{
    const uint64_t g = k ^ (k >> 1);
    uint32_t x1 = 0, x2 = 0, m = 1;
    for (int j = 1; j <= 32; ++j) {
        if (j > 1) m = (m << 1) ^ m;
        if ((g >> (j - 1)) & 1u) {
            x1 ^= 1u << (32 - j);
            x2 ^= m << (32 - j);
        }
    }
    r1 = (double)x1 * 0x1p-32;
    r2 = (double)x2 * 0x1p-32;
}

__global__ void k_sample_landau(int64_t n, int dim_v, double alpha, double k, double sigma,
                                double dimx, double *x, double *v1, double *v2, double *w,
                                int64_t *id, int *err)
{
    int errs = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double v = sigma * sqrt(-2.0 * log(((double)(i + 1) - 0.5) / (double)n));
        double r1, r2;
        sobol2((uint64_t)(i + 1), r1, r2);
        const double th = r1 * 6.283185307179586;
        double x0 = 0.0, x1 = 1.0;
        const double r = r2 * (6.283185307179586 / k);
        int it = 0;
        while (fabs(x1 - x0) > 1e-12 && it < 200) {
            const double p = x0 + alpha * sin(k * x0) / k;
            const double f = 1.0 + alpha * cos(k * x0);
            const double nx1 = x0 - (p - r) / f;
            x0 = x1;
            x1 = nx1;
            ++it;
        }
        if (it >= 200) errs |= E_FAR;
        x[i] = x1;
        v1[i] = v * cos(th);
        if (dim_v == 2) v2[i] = v * sin(th);
        w[i] = dimx;
        id[i] = i;
    }
    if (errs) atomicOr(err, errs);
}

hipError_t launch_sample_landau(int64_t n, int dim_v, double alpha, double k, double sigma,
                                double dimx, double *x, double *v1, double *v2, double *w,
                                int64_t *id, int *err, hipStream_t s)
{
    const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (n + kBlock - 1) / kBlock));
    hipLaunchKernelGGL(k_sample_landau, dim3(blocks), dim3(kBlock), 0, s, n, dim_v, alpha, k, sigma,
                       dimx, x, v1, v2, w, id, err);
    return hipGetLastError();
}

__global__ void k_aos_to_soa(const double *aos, int64_t ld, int64_t n, int dim_v, double *x,
                             double *v1, double *v2, double *w, int64_t *id)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const double *p = aos + i * ld;
        x[i] = p[0];
        v1[i] = p[1];
        if (dim_v == 2) v2[i] = p[2];
        w[i] = p[1 + dim_v];
        id[i] = i;
    }
}

__global__ void k_soa_to_aos(double *aos, int64_t ld, int64_t n, int dim_v, const double *x,
                             const double *v1, const double *v2, const double *w,
                             const int64_t *id)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = id[i];
        if (j < 0 || j >= n) continue;
        double *p = aos + j * ld;
        p[0] = x[i];
        p[1] = v1[i];
        if (dim_v == 2) p[2] = v2[i];
        p[1 + dim_v] = w[i];
    }
}

hipError_t launch_aos_to_soa(const double *aos, int64_t ld, int64_t n, int dim_v, double *x,
                             double *v1, double *v2, double *w, int64_t *id, hipStream_t s)
{
    int blocks = (int)((n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_aos_to_soa, dim3(blocks), dim3(256), 0, s, aos, ld, n, dim_v, x, v1, v2,
                       w, id);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(double *aos, int64_t ld, int64_t n, int dim_v, const double *x,
                             const double *v1, const double *v2, const double *w,
                             const int64_t *id, hipStream_t s)
{
    int blocks = (int)((n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_soa_to_aos, dim3(blocks), dim3(256), 0, s, aos, ld, n, dim_v, x, v1, v2,
                       w, id);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Batched ParticleMeshCoupling1D verbs (particles in the caller's order)
// ---------------------------------------------------------------------------

// horner_1d (splinepp.jl:268-274)
__device__ __forceinline__ double horner(int degree, const double *pp, int ld, double x, int col)
{
    double res = pp[col * ld];
    for (int i = 1; i <= degree; ++i) res = res * x + pp[i + col * ld];
    return res;
}

template <int P>
__device__ __forceinline__ void pp_basis(const double *pc, double x, double (&b)[P + 1])
{
#pragma unroll
    for (int i = 0; i <= P; ++i) b[i] = horner(P, pc, P + 1, x, i);
}

template <int P, int OP>
__global__ void __launch_bounds__(kBlock) k_pmc(PmcArgs a)
{
    constexpr int W = P + 1;
    constexpr int NQ = (P + 2) / 2;
    constexpr bool DEPOSIT = OP == PMC_CHARGE || OP == PMC_CHARGE_PP || OP == PMC_CURRENT_B ||
                             OP == PMC_CURRENT_NOB || OP == PMC_CURRENT_PP;
    extern __shared__ double smem[];
    const int nx = a.nx;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    double *sRed = smem;                       // [nx]
    double *rows = sRed + nx;                  // [kWaves][nx], later reduction scratch
    const int rows_len = kWaves * nx > kBlock ? kWaves * nx : kBlock;
    double *slots = rows + rows_len + wave * (W * kWave);   // this wave's cache-B slots
    for (int i = tid; i < kWaves * nx; i += kBlock) rows[i] = 0.0;
    __syncthreads();
    double *myrow = rows + wave * nx;
    Cache2<W> cache;
    if constexpr (DEPOSIT) cache.init(slots);
    int errs = 0;
    const int64_t base = (int64_t)blockIdx.x * a.chunk;
    const double rdx = 1.0 / a.dx;   // div_by: (x - xmin) / dx, correctly rounded
    for (int it = 0; it < a.iters; ++it) {
        const int64_t i = base + (int64_t)it * kBlock + tid;
        const bool valid = i < a.n;
        const double x = valid ? a.x[i] : a.xmin;
        const double xi0 = div_by(x - a.xmin, a.dx, rdx);
        if (valid && !pos_ok(xi0)) errs |= E_INEXACT;
        if constexpr (OP == PMC_EVAL || OP == PMC_EVAL_PP || OP == PMC_CHARGE ||
                      OP == PMC_CHARGE_PP) {
            const int ic = (int)xi0;
            const double off = xi0 - (double)ic;
            double b[W];
            if constexpr (OP == PMC_EVAL_PP || OP == PMC_CHARGE_PP) pp_basis<P>(a.pc, off, b);
            else bspline<P>(off, b);
            if constexpr (OP == PMC_EVAL) {
                if (valid) a.out[i] = gather<P>(b, a.field, imod(ic, nx), nx);
            } else if constexpr (OP == PMC_EVAL_PP) {
                // evaluate_pp: horner on column trunc(xi) of the pp matrix, unwrapped
                if (valid) {
                    if (ic < 0 || ic >= nx) errs |= E_BOUNDS;
                    else a.out[i] = horner(P, a.field, P + 1, off, ic);
                }
            } else {
                const double q = valid ? a.q[i] : 0.0;
                double vals[W];
#pragma unroll
                for (int k = 0; k < W; ++k) vals[k] = q * b[k] * a.scaling;
                cache_add<W>(cache, valid, imod(ic, nx), vals, myrow, nx);
            }
        } else {
            // add_current_update_v! / _pp!
            const double xn = valid ? a.x_new[i] : x;
            const double q = valid ? a.q[i] : 0.0;
            double v = (valid && OP != PMC_CURRENT_NOB) ? a.out[i] : 0.0;
            const int io = (int)xi0;
            const double ro = xi0 - (double)io;
            const double xin = div_by(xn - a.xmin, a.dx, rdx);
            if (valid && !pos_ok(xin)) errs |= E_INEXACT;
            const int inw = (OP == PMC_CURRENT_NOB) ? (int)floor(xin) : (int)xin;
            const double rn = xin - (double)inw;
            const int dseg = inw - io;
            const bool far = !pos_ok(xin) || !pos_ok(xi0) || dseg > kMaxCellsPerStep ||
                             dseg < -kMaxCellsPerStep;
            if (valid && far) errs |= E_FAR;
            const int nseg = far ? 1 : 1 + (dseg < 0 ? -dseg : dseg);
            for (int s = 0;; ++s) {
                const bool act = valid && s < nseg;
                if (!__ballot(act)) break;
                double sv[W];
                int cell;
                if constexpr (OP == PMC_CURRENT_PP) {
                    // update_jv_pp! bounds (:155-189): reversed segments swap bounds
                    double lo, up;
                    if (io == inw) { lo = ro; up = rn; cell = io; }
                    else if (io < inw) {
                        if (s == 0) { lo = ro; up = 1.0; cell = io; }
                        else if (s == 1) { lo = 0.0; up = rn; cell = inw; }
                        else { lo = 0.0; up = 1.0; cell = io + s - 1; }
                    } else {
                        if (s == 0) { lo = 1.0; up = rn; cell = inw; }
                        else if (s == 1) { lo = ro; up = 0.0; cell = io; }
                        else { lo = 1.0; up = 0.0; cell = inw + s - 1; }
                    }
                    double svm[W];
#pragma unroll
                    for (int k = 0; k < W; ++k) sv[k] = horner(P, a.pf, W, lo, k) * lo;
#pragma unroll
                    for (int k = 0; k < W; ++k) svm[k] = horner(P, a.pf, W, up, k) * up;
#pragma unroll
                    for (int k = 0; k < W; ++k) sv[k] = (svm[k] - sv[k]) * a.dx;
                } else {
                    double lo, up, sg;
                    segment_bounds(s, io, ro, inw, rn, lo, up, sg, cell);
                    segment<P, NQ>(lo, up, sg, a.qx, a.qw, a.dx, sv);
                }
                const int key = imod(cell, nx);
                double vals[W];
                int k = first_node(key, P, nx);
                double vs = v;
#pragma unroll
                for (int m = 0; m < W; ++m) {
                    vals[m] = q * sv[m] * a.scaling;
                    if constexpr (OP != PMC_CURRENT_NOB) vs = vs - a.qm * sv[m] * a.field[k];
                    k = next_node(k, nx);
                }
                if (act) v = vs;
                cache_add<W>(cache, act, key, vals, myrow, nx);
            }
            if (valid && OP != PMC_CURRENT_NOB) a.out[i] = v;
        }
    }
    if (errs) atomicOr(a.err, errs);
    if constexpr (DEPOSIT) {
        cache_drain<W>(cache, myrow, nx);
        __syncthreads();
        for (int idx = tid; idx < nx; idx += kBlock) {
            double s = 0.0;
            for (int wv = 0; wv < kWaves; ++wv) s += rows[wv * nx + idx];
            a.part[(size_t)blockIdx.x * nx + idx] = s;
        }
        if (!grid_reduce(a.part, a.gpart, a.cnt, a.ngroups, nx, 1u, nx, sRed, rows)) return;
        for (int idx = tid; idx < nx; idx += kBlock) a.acc[idx] = a.acc[idx] + sRed[idx];
    }
}

template <int P>
static hipError_t launch_pmc_p(int op, const PmcArgs &a, int blocks, hipStream_t s)
{
    size_t rows = (size_t)kWaves * a.nx;
    if (rows < (size_t)kBlock) rows = kBlock;
    const size_t sh = sizeof(double) * ((size_t)a.nx + rows + (size_t)kWaves * (P + 1) * kWave);
    switch (op) {
#define GK_OP(O)                                                                        \
    case O:                                                                             \
        hipLaunchKernelGGL((k_pmc<P, O>), dim3(blocks), dim3(kBlock), sh, s, a);        \
        return hipGetLastError();
        GK_OP(PMC_EVAL)
        GK_OP(PMC_EVAL_PP)
        GK_OP(PMC_CHARGE)
        GK_OP(PMC_CHARGE_PP)
        GK_OP(PMC_CURRENT_B)
        GK_OP(PMC_CURRENT_NOB)
        GK_OP(PMC_CURRENT_PP)
#undef GK_OP
    default:
        return hipErrorInvalidValue;
    }
}

hipError_t launch_pmc(int op, const PmcArgs &a, int blocks, hipStream_t s)
{
    switch (a.degree) {
    case 0:
        if (op == PMC_EVAL_PP || op == PMC_CHARGE_PP || op == PMC_CURRENT_PP)
            return hipErrorInvalidValue;
        return launch_pmc_p<0>(op, a, blocks, s);
    case 1: return launch_pmc_p<1>(op, a, blocks, s);
    case 2: return launch_pmc_p<2>(op, a, blocks, s);
    case 3: return launch_pmc_p<3>(op, a, blocks, s);
    case 4: return launch_pmc_p<4>(op, a, blocks, s);
    case 5: return launch_pmc_p<5>(op, a, blocks, s);
    default: return hipErrorInvalidValue;
    }
}

// ---------------------------------------------------------------------------
// Reverse mode of the fused 1D2V Strang step.
//
// The forward pass k_pass1d<P0,P1,HE_PREV|HE|HP2A|HP1|HP2B> maps a particle
// (x, v1, v2) through HE of the previous step, HE, Hp2, Hp1, Hp2 of this one
// (hamiltonian_splitting_1d2v.jl:41-219) with fields that stay fixed during
// the pass. k_adj1d re-runs that map in registers from the state before the
// pass (bit for bit, with the forward's operations in the forward's order)
// and applies the transpose of its Jacobian to the particle's cotangents:
//
//   Hp2 (v1 += c qm v2 B(x), j2 += q v2 S0(x)):
//       lv2 += c qm B(x) lv1 + q s0 mu2(x),
//       lx  += c qm v2 lv1 B'(x) + q s0 v2 mu2'(x),   lambda_B += c qm v2 lv1 S1(x)
//   Hp1 (x_new = x + dt v1, v2 -= qm int_x^x_new B, j1 += q int S1, x = mod(x_new)):
//       lxn = lx - qm lv2 B(x_new) + q s1 mu1(x_new),
//       lx  = lxn + qm lv2 B(x) - q s1 mu1(x),  lv1 += dt lxn,  lambda_B -= qm lv2 int S1
//   HE (v += c qm E(x)):
//       lx += c qm (lv1 E1'(x) + lv2 E2'(x)),  lambda_E1 += c qm lv1 S1(x),
//       lambda_E2 += c qm lv2 S0(x)
//
// mu1 = -(1/dx) M1^-1 lambda_E1 and mu2 = -(c/dx) M0^-1 lambda_E2 are the
// cotangents of the pass's current deposits, from the field part of the
// reverse step (k_adj_epi). Endpoint values use the trunc() cells of
// add_current_update_v! (particle_mesh_coupling_1d.jl:307,313), for which the
// Gauss-Legendre line integral is exact, so its derivatives are the spline
// values there. B' = (1/dx) sum B_k S1_k'(t), S_p'[r] = S_{p-1}[r-1] - S_{p-1}[r].
// The cotangent deposits use the same cached, deterministic accumulation as
// the forward pass.
// ---------------------------------------------------------------------------

template <int P>
__device__ __forceinline__ void dbspline(double off, double (&d)[P + 1])
{
    if constexpr (P == 0) {
        d[0] = 0.0;
    } else {
        double lo[P];
        bspline<P - 1>(off, lo);
        d[0] = -lo[0];
#pragma unroll
        for (int r = 1; r < P; ++r) d[r] = lo[r - 1] - lo[r];
        d[P] = lo[P - 1];
    }
}

// waves per SIMD the reverse pass is compiled for (register budget)
#ifndef GEMPIC_ADJ_WAVES
#define GEMPIC_ADJ_WAVES 2
#endif

template <int P0, int P1>
constexpr int adj_slots() { return 2 * (P1 + 1) + (P0 + 1); }

size_t adj_shmem_bytes(int nx, int slots)
{
    return sizeof(double) *
           ((size_t)6 * (nx + 2 * kPad) + (size_t)kWaves * NDEP * nx + (size_t)kWaves * slots * kWave);
}

template <int P0, int P1, int RM>
__device__ __forceinline__ void adj1d_body(const AdjArgs &A, int bid, double *smem)
{
    constexpr bool PREV = (RM & R_PREV) != 0;
    constexpr bool FWD = (RM & R_FWD) != 0;
    constexpr bool REV = (RM & R_REV) != 0;
    constexpr bool HEONLY = (RM & R_HEONLY) != 0;
    constexpr bool POIS = (RM & R_POIS) != 0;
    constexpr bool STEP = FWD || REV;
    constexpr bool DEP = REV || HEONLY;
    constexpr int NQ = (P1 + 2) / 2;
    constexpr int W0 = P0 + 1, W1 = P1 + 1;
    constexpr int SL = adj_slots<P0, P1>();
    static_assert(P0 <= kPad && P1 <= kPad, "LDS field padding must cover the spline support");

    const PassArgs &a = A.p;
    const int nx = a.nx;
    const int tid = threadIdx.x;
    const int wave = tid / kWave;
    const int nxp = nx + 2 * kPad;
    double *pE1 = smem, *pE2 = pE1 + nxp, *pB = pE2 + nxp, *pE2p = pB + nxp;
    double *pM1 = pE2p + nxp, *pM2 = pM1 + nxp;
    double *rows = pM2 + nxp;                               // [kWaves][NDEP][nx]
    double *slots = rows + kWaves * NDEP * nx + wave * (SL * kWave);
    const double *sE1 = pE1 + kPad, *sE2 = pE2 + kPad, *sB = pB + kPad, *sE2p = pE2p + kPad;
    const double *sM1 = pM1 + kPad, *sM2 = pM2 + kPad;

    if constexpr (!POIS) {
        const double *F = A.fields;
        load_padded(pE1, F + H_E1 * nx, nx);
        load_padded(pE2, F + H_E2 * nx, nx);
        load_padded(pB, F + H_B * nx, nx);
        load_padded(pE2p, F + H_E2P * nx, nx);
    }
    if constexpr (REV) {
        load_padded(pM1, A.agrid + A_MU1 * nx, nx);
        load_padded(pM2, A.agrid + A_MU2 * nx, nx);
    }
    if constexpr (POIS) load_padded(pM1, A.agrid + A_NU * nx, nx);
    for (int i = tid; i < kWaves * NDEP * nx; i += kBlock) rows[i] = 0.0;
    __syncthreads();
    double *myrow = rows + wave * NDEP * nx;

    // HEONLY: lambda_E1 and lambda_E2 caches; REV: lambda_E1 and lambda_B
    // cached together (both are S1-basis deposits, mostly in the old cell),
    // and lambda_E2
    Cache2<W1> c_e1;
    Cache2<2 * W1> c_eb;
    Cache2<W0> c_e2;
    if constexpr (HEONLY) {
        c_e1.init(slots);
        c_e2.init(slots + W1 * kWave);
    }
    if constexpr (REV) {
        c_eb.init(slots);
        c_e2.init(slots + 2 * W1 * kWave);
    }

    const double xmin = a.xmin, dx = a.dx, qm = a.qm;
    const double rdx = 1.0 / dx;     // div_by: every quotient by dx, correctly rounded
    const double dtqm = a.dtqm_he, g = a.dtqm_hp2, dt1 = a.dt_hp1;
    const double s0 = a.scaling0, s1 = a.scaling1;
    const PassChunk pc = pass_chunk(a, bid);
    const int64_t base = pc.base;
    const int cnt_here = pc.cnt;
    const int iters = pc.iters;
    // this workgroup's slices of the particle arrays (element access by elem())
    const double *qx = A.x + base, *qw = a.w + base;
    const double *qv1 = STEP ? A.v1 + base : nullptr, *qv2 = STEP ? A.v2 + base : nullptr;
    double *qlx = A.lx + base;
    double *qlv1 = DEP ? A.lv1 + base : nullptr, *qlv2 = DEP ? A.lv2 + base : nullptr;

    // the loss's explicit dependence on the cell size dx, on the domain
    // length through the mod, and on the particle weights (summed): per lane,
    // then per wave and workgroup in a fixed order
    double acc_dx = 0.0, acc_mod = 0.0, acc_w = 0.0;
    const double cwq = a.charge * a.cw;         // d q / d w

    // software pipeline: the primal state of iteration it+1 is loaded before
    // the work of iteration it (the cotangents are loaded at the top of their
    // own iteration; the forward re-run covers their latency)
    double nx_x = xmin, nx_v1 = 0.0, nx_v2 = 0.0, nx_w = 0.0;
    if (iters > 0 && tid < cnt_here) {
        nx_x = elem(qx, tid);
        nx_w = elem(qw, tid);
        if constexpr (STEP) {
            nx_v1 = elem(qv1, tid);
            nx_v2 = elem(qv2, tid);
        }
    }
    for (int it = 0; it < iters; ++it) {
        const int i = it * kBlock + tid;
        const bool valid = i < cnt_here;
        double x = nx_x, v1 = nx_v1, v2 = nx_v2, w = nx_w;
        {
            const int in = i + kBlock;
            if (it + 1 < iters && in < cnt_here) {
                nx_x = elem(qx, in);
                nx_w = elem(qw, in);
                if constexpr (STEP) {
                    nx_v1 = elem(qv1, in);
                    nx_v2 = elem(qv2, in);
                }
            }
        }
        double lx = 0.0, lv1 = 0.0, lv2 = 0.0;
        if (valid && (DEP || POIS)) lx = elem(qlx, i);
        if (valid && DEP) {
            lv1 = elem(qlv1, i);
            lv2 = elem(qlv2, i);
        }
        const double q = (w * a.charge) * a.cw;
        const double xi = div_by(x - xmin, dx, rdx);
        const int ic = pos_ok(xi) ? (int)xi : 0;
        const double off = xi - (double)ic;
        const int key = imod(ic, nx);
        double b1[W1], b0[W0];
        bspline_pair<P0, P1>(off, b0, b1);

        if constexpr (POIS) {
            double db0[W0];
            dbspline<P0>(off, db0);
            const double dnu = div_by(gather_p<P0>(db0, sM1, key), dx, rdx);
            if (valid) elem(qlx, i) = q * s0 * dnu;
            continue;
        }
        if constexpr (HEONLY) {
            double db1[W1], db0[W0];
            dbspline<P1>(off, db1);
            dbspline<P0>(off, db0);
            const double de1 = div_by(gather_p<P1>(db1, sE1, key), dx, rdx);
            const double de2p = div_by(gather_p<P0>(db0, sE2p, key), dx, rdx);
            const double dhe = dtqm * (lv1 * de1 + lv2 * de2p);
            lx = lx + dhe;
            if (valid) acc_dx = acc_dx - xi * dhe;
            double ve[W1], vf[W0];
#pragma unroll
            for (int k = 0; k < W1; ++k) ve[k] = (dtqm * lv1) * b1[k];
#pragma unroll
            for (int k = 0; k < W0; ++k) vf[k] = (dtqm * lv2) * b0[k];
            cache_add<W1>(c_e1, valid, key, ve, myrow + AR_E1 * nx, nx);
            cache_add<W0>(c_e2, valid, key, vf, myrow + AR_E2 * nx, nx);
            if (valid) elem(qlx, i) = lx;
            continue;
        }

        // ---- the forward pass, as k_pass1d computes it ----
        const double e1 = gather_p<P1>(b1, sE1, key);
        if constexpr (PREV) {
            const double e2 = gather_p<P0>(b0, sE2p, key);
            v1 = v1 + dtqm * e1;
            v2 = v2 + dtqm * e2;
        }
        {
            const double e2 = gather_p<P0>(b0, sE2, key);
            v1 = v1 + dtqm * e1;
            v2 = v2 + dtqm * e2;
        }
        const double bA = gather_p<P1>(b1, sB, key);
        const double v2b = v2;
        v1 = v1 + g * v2 * bA;
        const double xn = x + dt1 * v1;
        const double xin = div_by(xn - xmin, dx, rdx);
        const int inw = pos_ok(xin) ? (int)xin : 0;
        const double rn = xin - (double)inw;
        const int dseg = inw - ic;
        const bool far = !pos_ok(xin) || !pos_ok(xi) || dseg > kMaxCellsPerStep ||
                         dseg < -kMaxCellsPerStep;
        const int nseg = far ? 1 : 1 + (dseg < 0 ? -dseg : dseg);
        double sv0[W1];          // the first segment's integrals, kept for the reverse
        for (int sgi = 0; sgi < nseg; ++sgi) {
            double lo, up, sg;
            int cell;
            segment_bounds(sgi, ic, off, inw, rn, lo, up, sg, cell);
            double sv[W1];
            segment<P1, NQ>(lo, up, sg, a.qx, a.qw, dx, sv);
            const double *bk = sB + (imod(cell, nx) - P1);
#pragma unroll
            for (int m = 0; m < W1; ++m) {
                v2 = v2 - qm * sv[m] * bk[m];
                if (sgi == 0) sv0[m] = sv[m];
            }
        }
        const double x1 = jl_mod(xn, a.Lx);
        const bool same = __double_as_longlong(x1) == __double_as_longlong(xn);
        const double xi2 = same ? xin : div_by(x1 - xmin, dx, rdx);
        const int ic2 = pos_ok(xi2) ? (int)xi2 : 0;
        const double off2 = xi2 - (double)ic2;
        const int key2 = imod(ic2, nx);
        double c0[W0], c1[W1];
        bspline_pair<P0, P1>(off2, c0, c1);
        const double bB = gather_p<P1>(c1, sB, key2);
        if constexpr (FWD) {
            v1 = v1 + g * v2 * bB;
            if (valid) {
                elem(A.xo + base, i) = x1;
                elem(A.v1o + base, i) = v1;
                elem(A.v2o + base, i) = v2;
            }
            continue;
        }

        // ---- reverse ----
        if constexpr (REV) {
            const double v2d = v2;
            // lambda_B lands in the particle's old cell (first Hp2, the line
            // integral's cell when it stays there, second Hp2 when the push
            // kept the cell): those parts are summed per particle first
            double vkey[W1];
            // second Hp2
            double dc0[W0], dc1[W1];
            if constexpr (P1 == P0 - 1) {
                dc0[0] = -c1[0];
#pragma unroll
                for (int r = 1; r < P0; ++r) dc0[r] = c1[r - 1] - c1[r];
                dc0[P0] = c1[P1];
            } else {
                dbspline<P0>(off2, dc0);
            }
            dbspline<P1>(off2, dc1);
            const double gl2 = g * v2d * lv1;
            double lx1;                 // cotangent of the position after the mod
            {
                const double mu = gather_p<P0>(c0, sM2, key2);
                const double dmu = div_by(gather_p<P0>(dc0, sM2, key2), dx, rdx);
                const double db = div_by(gather_p<P1>(dc1, sB, key2), dx, rdx);
                const double d2 = gl2 * db + q * s0 * v2d * dmu;
                lx = lx + d2;
                lv2 = lv2 + g * bB * lv1 + q * s0 * mu;
                lx1 = lx;
                if (valid) {
                    acc_dx = acc_dx - xi2 * d2 - div_by(q * s0 * v2d * mu, dx, rdx);
                    acc_w = acc_w + cwq * s0 * v2d * mu;
                }
            }
            const bool new_here = key2 == key;
#pragma unroll
            for (int k = 0; k < W1; ++k) vkey[k] = new_here ? gl2 * c1[k] : 0.0;
            // Hp1: the basis at the pushed position is c1 when the mod kept it
            {
                double bn[W1];
                if (same) {
#pragma unroll
                    for (int k = 0; k < W1; ++k) bn[k] = c1[k];
                } else {
                    bspline<P1>(rn, bn);
                }
                const int keyn = imod(inw, nx);
                const double bN = same ? bB : gather_p<P1>(bn, sB, keyn);
                const double muN = gather_p<P1>(bn, sM1, keyn);
                const double muO = gather_p<P1>(b1, sM1, key);
                const double lxn = lx - qm * lv2 * bN + q * s1 * muN;
                lx = lxn + qm * lv2 * bA - q * s1 * muO;
                lv1 = lv1 + dt1 * lxn;
                const double wl = -qm * lv2;
                const double lv2h = lv2;
                double ib = 0.0, im1 = 0.0;     // int B and int mu1 along the path
                for (int sgi = 0;; ++sgi) {
                    const bool act = valid && sgi < nseg;
                    if (!__ballot(act)) break;
                    double lo, up, sg;
                    int cell;
                    segment_bounds(sgi, ic, off, inw, rn, lo, up, sg, cell);
                    double sv[W1];
                    if (sgi == 0) {
#pragma unroll
                        for (int m = 0; m < W1; ++m) sv[m] = sv0[m];
                    } else {
                        segment<P1, NQ>(lo, up, sg, a.qx, a.qw, dx, sv);
                    }
                    const int kc = imod(cell, nx);
                    const bool here = kc == key;
                    const double *bk = sB + (kc - P1);
                    const double *mk = sM1 + (kc - P1);
                    double vb[2 * W1];
#pragma unroll
                    for (int m = 0; m < W1; ++m) {
                        vb[m] = 0.0;
                        vb[W1 + m] = wl * sv[m];
                        if (act && here) vkey[m] = vkey[m] + vb[W1 + m];
                        if (act) {
                            ib = ib + sv[m] * bk[m];
                            im1 = im1 + sv[m] * mk[m];
                        }
                    }
                    cache_add_rows2<W1>(c_eb, act && !here, kc, vb, myrow + AR_E1 * nx,
                                        myrow + AR_B * nx, nx);
                }
                if (valid) {
                    // d/d dx of int_x^xn f = (1/dx) int f - xin f(xn) + xi f(x)
                    acc_dx = acc_dx - qm * lv2h * (div_by(ib, dx, rdx) - xin * bN + xi * bA) +
                             q * s1 * (xi * muO - xin * muN);
                    acc_w = acc_w + cwq * s1 * im1;
                    // x = mod(xn, Lx): d x / d Lx = the number of periods added
                    if (!same) acc_mod = acc_mod + lx1 * (x1 > xn ? 1.0 : -1.0);
                }
            }
            // first Hp2
            double db1[W1], db0[W0];
            dbspline<P1>(off, db1);
            if constexpr (P1 == P0 - 1) {
                db0[0] = -b1[0];
#pragma unroll
                for (int r = 1; r < P0; ++r) db0[r] = b1[r - 1] - b1[r];
                db0[P0] = b1[P1];
            } else {
                dbspline<P0>(off, db0);
            }
            {
                const double mu = gather_p<P0>(b0, sM2, key);
                const double dmu = div_by(gather_p<P0>(db0, sM2, key), dx, rdx);
                const double db = div_by(gather_p<P1>(db1, sB, key), dx, rdx);
                const double gl1 = g * v2b * lv1;
                const double d1 = gl1 * db + q * s0 * v2b * dmu;
                lx = lx + d1;
                lv2 = lv2 + g * bA * lv1 + q * s0 * mu;
                if (valid) {
                    acc_dx = acc_dx - xi * d1 - div_by(q * s0 * v2b * mu, dx, rdx);
                    acc_w = acc_w + cwq * s0 * v2b * mu;
                }
#pragma unroll
                for (int k = 0; k < W1; ++k) vkey[k] = vkey[k] + gl1 * b1[k];
            }
            {
                // lambda_E1 of the two HE (the same cotangent and cell) with lambda_B
                double veb[2 * W1];
#pragma unroll
                for (int k = 0; k < W1; ++k) {
                    veb[k] = (dtqm * lv1) * b1[k];
                    veb[W1 + k] = vkey[k];
                }
                cache_add_rows2<W1>(c_eb, valid, key, veb, myrow + AR_E1 * nx, myrow + AR_B * nx, nx);
                double vb[2 * W1];
#pragma unroll
                for (int k = 0; k < W1; ++k) {
                    vb[k] = 0.0;
                    vb[W1 + k] = gl2 * c1[k];
                }
                cache_add_rows2<W1>(c_eb, valid && !new_here, key2, vb, myrow + AR_E1 * nx,
                                    myrow + AR_B * nx, nx);
            }
            // HE of this step and of the previous one: same x and cotangents
            {
                const double de1 = div_by(gather_p<P1>(db1, sE1, key), dx, rdx);
                const double de2 = div_by(gather_p<P0>(db0, sE2, key), dx, rdx);
                const double dh = dtqm * (lv1 * de1 + lv2 * de2);
                lx = lx + dh;
                double dhs = dh;
                if constexpr (PREV) {
                    const double de2p = div_by(gather_p<P0>(db0, sE2p, key), dx, rdx);
                    const double dhp = dtqm * (lv1 * de1 + lv2 * de2p);
                    lx = lx + dhp;
                    dhs = dhs + dhp;
                }
                if (valid) acc_dx = acc_dx - xi * dhs;
                double vf[W0];
#pragma unroll
                for (int k = 0; k < W0; ++k) vf[k] = (dtqm * lv2) * b0[k];
                cache_add<W0>(c_e2, valid, key, vf, myrow + AR_E2 * nx, nx);
            }
            if (valid) {
                elem(qlx, i) = lx;
                elem(qlv1, i) = lv1;
                elem(qlv2, i) = lv2;
            }
        }
    }

    if constexpr (DEP) {
        {
            // wave sums (xor butterfly: identical in every lane) into the
            // wave's row AR_SCAL; the loop below adds the waves in order
            const double sdx = wave_sum(acc_dx), smod = wave_sum(acc_mod), sw = wave_sum(acc_w);
            if ((tid & (kWave - 1)) == 0) {
                myrow[AR_SCAL * nx + 0] = sdx;
                myrow[AR_SCAL * nx + 1] = smod;
                myrow[AR_SCAL * nx + 2] = sw;
            }
        }
        if constexpr (HEONLY) cache_drain<W1>(c_e1, myrow + AR_E1 * nx, nx);
        if constexpr (REV) cache_drain_rows2<W1>(c_eb, myrow + AR_E1 * nx, myrow + AR_B * nx, nx);
        cache_drain<W0>(c_e2, myrow + AR_E2 * nx, nx);
        __syncthreads();
        constexpr unsigned MASK = (1u << AR_E1) | (1u << AR_E2) | (REV ? 1u << AR_B : 0u) |
                                  (1u << AR_SCAL);
        const int len = NDEP * nx;
        for (int idx = tid; idx < len; idx += kBlock) {
            if (!((MASK >> (idx / nx)) & 1u)) continue;
            double s = 0.0;
            for (int wv = 0; wv < kWaves; ++wv) s += rows[wv * len + idx];
            a.part[(size_t)bid * len + idx] = s;
        }
    }
}

template <int P0, int P1, int RM>
__global__ void __launch_bounds__(kBlock, GEMPIC_ADJ_WAVES) k_adj1d(AdjArgs A)
{
    extern __shared__ double smem[];
    adj1d_body<P0, P1, RM>(A, blockIdx.x, smem);
}

// every instance of an ensemble in one launch (as k_pass1d_batch)
template <int P0, int P1, int RM>
__global__ void __launch_bounds__(kBlock, GEMPIC_ADJ_WAVES) k_adj1d_batch(AdjBatch b)
{
    extern __shared__ double smem[];
    const int inst = batch_instance(b.block_start, b.ninst, blockIdx.x);
    adj1d_body<P0, P1, RM>(b.args[inst], blockIdx.x - b.block_start[inst], smem);
}

// lambda_e[j] += coef (lambda_b[j+1] - lambda_b[j]): the transpose of
// compute_b_from_e! (maxwell_1d_fem.jl:407-420), by the whole workgroup
__device__ void adj_b_from_e(double *le, double coef, const double *lb, int n)
{
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const int jp = j + 1 == n ? 0 : j + 1;
        le[j] = le[j] + coef * (lb[jp] - lb[j]);
    }
    __syncthreads();
}

// The field part of a reverse step: the transpose of the grid work between
// two forward passes, in reverse order, with the deposit rows of the reverse
// pass just run. With c = dt/2, U = compute_b_from_e!(c), H = compute_e_from_b!(c),
// L = the loss at a step's end:
//   AE_TAIL  (after the trailing HE):     L, H, U, then the HE rows at E_c
//   AE_STEP  (between steps s and s+1):   the pass's rows at (E_f, B_b), U, H, L, H, U,
//            then the HE rows again at E_c (HE of step s)
//   AE_HEAD  (the prologue):              the pass's rows at (E_f, B_b), U, H
// and then, for AE_TAIL and AE_STEP, mu1 = -(1/dx) M1^-1 lambda_E1 and
// mu2 = -(c/dx) M0^-1 lambda_E2 for the next reverse pass. AE_POIS computes
// nu = C_pois D^T lambda_E1 for the solve_poisson! gradient.
template <int KIND, int PHASE = 0>
__device__ __forceinline__ void adj_epi_body(const AdjArgs &A, int nblocks, int gb, int ng,
                                             double *smem)
{
    const PassArgs &a = A.p;
    __shared__ int s_last;
    const int nx = a.nx;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int len = NDEP * nx;
    const double dx = a.dx;
    double *lE1 = smem, *lE2 = lE1 + nx, *lB = lE2 + nx;
    double *red = lB + nx;                  // [NDEP][nx]
    double *w3 = red + NDEP * nx;           // [6][nx]: three outputs, three workspaces
    double *T = w3 + 6 * nx;                // [NTAB][nx]
    double *bend = T + NTAB * nx;
    double *scratch = bend + nx;            // [blockDim.x]
    double *G = A.agrid;

    for (int i = tid; i < nx; i += nt) {
        lE1[i] = G[A_E1 * nx + i];
        lE2[i] = G[A_E2 * nx + i];
        lB[i] = G[A_B * nx + i];
        bend[i] = A.bend ? A.bend[i] : 0.0;
    }
    for (int i = tid; i < NTAB * nx; i += nt) T[i] = a.tabs[i];
    for (int i = tid; i < len; i += nt) red[i] = 0.0;
    __syncthreads();

    constexpr unsigned MASK = KIND == AE_TAIL ? (1u << AR_E1) | (1u << AR_E2) | (1u << AR_SCAL)
                              : KIND == AE_POIS ? 0u
                                                : (1u << AR_E1) | (1u << AR_E2) | (1u << AR_B) |
                                                      (1u << AR_SCAL);
    if constexpr (MASK != 0u) {
        // PHASE 1: only this workgroup's slice sum; PHASE 2: only the rest (a
        // later launch, no fences or counters); PHASE 0: both, as one launch
        if constexpr (PHASE != 2) {
            const int b0 = (int)(((long)nblocks * gb) / ng), b1 = (int)(((long)nblocks * (gb + 1)) / ng);
            sum_rows_range(a.part, b0, b1, nx, MASK, a.gpart + (size_t)gb * len, scratch);
            if constexpr (PHASE == 1) return;
            __threadfence();
            __syncthreads();
            if (tid == 0) s_last = atomicAdd(&a.cnt[0], 1u) == (unsigned)(ng - 1);
            __syncthreads();
            if (!s_last) return;
            __threadfence();
            if (tid == 0) a.cnt[0] = 0u;
        } else {
            (void)s_last;
        }
        sum_rows_range(a.gpart, 0, ng, nx, MASK, red, scratch);
    }
    __syncthreads();
    const double *dE1 = red + AR_E1 * nx, *dE2 = red + AR_E2 * nx, *dB = red + AR_B * nx;
    const double c = a.dt_he;
    const double coef = c / dx;
    // explicit derivatives of the loss in dx: the particle sums of the pass,
    // and (thread 0, below) the grid terms
    //   U = compute_b_from_e!:  -(1/dx) lambda_B . (c/dx) D E2
    //   H = compute_e_from_b!:  -(1/dx) lambda_E2 . (c/dx) W B = -(1/dx) (c/dx) (W^T lambda_E2) . B
    //   L = w/2 dx B.M1 B:      w/2 B.M1 B
    double gdx = 0.0;
    const double *E2f = A.fields ? A.fields + H_E2 * nx : nullptr;
    const double *E2c = A.fields ? A.fields + H_E2P * nx : nullptr;
    // grid dot products: wave 0's lanes take nodes j = lane, lane + 64, ...,
    // and a butterfly adds the lanes in a fixed order (the sum is the same in
    // every lane of wave 0; call from all of wave 0 only)
    auto dot_d = [&](const double *lb, const double *e) {      // lb . D e, D e[j] = e[j-1] - e[j]
        double r = 0.0;
        for (int j = tid; j < nx; j += kWave) r += lb[j] * (e[j == 0 ? nx - 1 : j - 1] - e[j]);
        return wave_sum(r);
    };
    auto dot = [&](const double *u, const double *v) {
        double r = 0.0;
        for (int j = tid; j < nx; j += kWave) r += u[j] * v[j];
        return wave_sum(r);
    };
    if (KIND != AE_POIS && tid == 0) {
        G[A_LOSS * nx + AS_DX] = G[A_LOSS * nx + AS_DX] + red[AR_SCAL * nx + 0];
        G[A_LOSS * nx + AS_MOD] = G[A_LOSS * nx + AS_MOD] + red[AR_SCAL * nx + 1];
        G[A_LOSS * nx + AS_W] = G[A_LOSS * nx + AS_W] + red[AR_SCAL * nx + 2];
    }
    // the field part's circulant operators are matrix-vector products with
    // their first columns (circ_apply, circ_apply_t): the reverse mode agrees
    // with the transform form to rounding, and has no transform phases
    double *wA = w3, *wM = w3 + nx, *w1 = w3 + 2 * nx;

    if constexpr (KIND == AE_POIS) {
        // the cotangent of E arrives in the A_NU slot and nu replaces it
        double *tmp = w3 + 3 * nx;
        for (int i = tid; i < nx; i += nt) {
            tmp[i] = 0.0;
            bend[i] = G[A_NU * nx + i];
        }
        __syncthreads();
        adj_b_from_e(tmp, 1.0, bend, nx);
        for (int i = tid; i < nx; i += nt) G[A_NU * nx + i] = circ_apply(T + T_CPOIS * nx, tmp, i, nx);
        return;
    }
    if constexpr (KIND == AE_STEP || KIND == AE_HEAD) {
        // cotangents of the fields the pass read: B_b, E1 and E2 of its first HE
        for (int i = tid; i < nx; i += nt) {
            lB[i] = lB[i] + dB[i];
            lE1[i] = lE1[i] + dE1[i];
            lE2[i] = lE2[i] + dE2[i];
        }
        __syncthreads();
        if (tid < kWave) {
            const double r = dot_d(lB, E2f);
            if (tid == 0) gdx = gdx - (coef / dx) * r;
        }
        adj_b_from_e(lE2, coef, lB, nx);                       // U of the next step's HE
    }
    if constexpr (KIND == AE_HEAD) {
        for (int i = tid; i < nx; i += nt) wA[i] = circ_apply_t(T + T_CAMP * nx, lE2, i, nx);
        __syncthreads();                                        // H of the first step
        for (int i = tid; i < nx; i += nt) {
            lB[i] = lB[i] + coef * wA[i];
            G[A_E1 * nx + i] = lE1[i];
            G[A_E2 * nx + i] = lE2[i];
            G[A_B * nx + i] = lB[i];
        }
        __syncthreads();
        if (tid < kWave) {
            const double r = dot(wA, bend);                    // (W^T lambda_E2) . B_0
            if (tid == 0) {
                gdx = gdx - (coef / dx) * r;
                G[A_LOSS * nx + AS_DX] = G[A_LOSS * nx + AS_DX] + gdx;
            }
        }
        return;
    }
    if constexpr (KIND == AE_STEP || KIND == AE_TAIL) {
        // lambda_E1 is final once the rows of this step's last HE are added
        for (int i = tid; i < nx; i += nt) lE1[i] = lE1[i] + dE1[i];
        __syncthreads();
        // W^T lambda_E2, M1 B_end and M1^-1 lambda_E1, each a matrix-vector
        // product with the circulant's first column
        for (int t = tid; t < 3 * nx; t += nt) {
            const int k = t / nx, i = t - k * nx;
            if (k == 0) wA[i] = circ_apply_t(T + T_CAMP * nx, lE2, i, nx);
            else if (k == 1) wM[i] = circ_apply(T + T_CMASS1 * nx, bend, i, nx);
            else w1[i] = circ_apply(T + T_CM1 * nx, lE1, i, nx);
        }
        __syncthreads();
        const double wl = A.wloss;
        for (int i = tid; i < nx; i += nt) {
            double b = lB[i];
            if (KIND == AE_STEP) b = b + coef * wA[i];         // H of the next step
            b = b + wl * dx * wM[i];                           // L at this step's end
            b = b + coef * wA[i];                              // H of this step
            lB[i] = b;
            G[A_MU1 * nx + i] = -w1[i] / dx;
        }
        __syncthreads();
        if (tid < kWave) {
            const double r = dot(bend, wM), rh = dot(wA, bend);
            const double rd = dot_d(lB, E2c);
            if (tid == 0) {
                G[A_LOSS * nx] = G[A_LOSS * nx] + wl * 0.5 * (r * dx);
                gdx = gdx + wl * 0.5 * r - (KIND == AE_STEP ? 2.0 : 1.0) * (coef / dx) * rh;
                gdx = gdx - (coef / dx) * rd;
                G[A_LOSS * nx + AS_DX] = G[A_LOSS * nx + AS_DX] + gdx;
            }
        }
        adj_b_from_e(lE2, coef, lB, nx);                       // U of this step's HE
        for (int i = tid; i < nx; i += nt) lE2[i] = lE2[i] + dE2[i];
        __syncthreads();
        for (int i = tid; i < nx; i += nt) {
            G[A_MU2 * nx + i] = -(c / dx) * circ_apply(T + T_CM0 * nx, lE2, i, nx);
            G[A_E1 * nx + i] = lE1[i];
            G[A_E2 * nx + i] = lE2[i];
            G[A_B * nx + i] = lB[i];
        }
    }
}

template <int KIND>
__global__ void __launch_bounds__(kBlock) k_adj_epi(AdjArgs A, int nblocks)
{
    extern __shared__ double smem[];
    adj_epi_body<KIND>(A, nblocks, blockIdx.x, gridDim.x, smem);
}

template <int KIND>
__global__ void __launch_bounds__(kBlock) k_adj_epi_batch_sum(AdjBatch b)
{
    extern __shared__ double smem[];
    const int inst = batch_instance(b.epi_start, b.ninst, blockIdx.x);
    const int g0 = b.epi_start[inst];
    adj_epi_body<KIND, 1>(b.args[inst], b.block_start[inst + 1] - b.block_start[inst],
                          blockIdx.x - g0, b.epi_start[inst + 1] - g0, smem);
}

template <int KIND>
__global__ void __launch_bounds__(kBlock) k_adj_epi_batch_fin(AdjBatch b)
{
    extern __shared__ double smem[];
    const int inst = blockIdx.x;
    adj_epi_body<KIND, 2>(b.args[inst], b.block_start[inst + 1] - b.block_start[inst], 0,
                          b.epi_start[inst + 1] - b.epi_start[inst], smem);
}

size_t adj_epi_shmem_bytes(int nx)
{
    return sizeof(double) * ((size_t)(3 + NDEP + 6 + NTAB + 1) * nx + kBlock);
}

template <int P0, int P1>
static hipError_t launch_adj_pd(int rm, const AdjArgs &a, int blocks, hipStream_t s)
{
    const size_t sh = adj_shmem_bytes(a.p.nx, adj_slots<P0, P1>());
#define GK_CASE(M)                                                                          \
    case (M):                                                                               \
        hipLaunchKernelGGL((k_adj1d<P0, P1, (M)>), dim3(blocks), dim3(kBlock), sh, s, a); \
        return hipGetLastError();
    switch (rm) {
        GK_CASE(R_FWD)
        GK_CASE(R_FWD | R_PREV)
        GK_CASE(R_REV)
        GK_CASE(R_REV | R_PREV)
        GK_CASE(R_HEONLY)
        GK_CASE(R_POIS)
    default:
        return hipErrorInvalidValue;
    }
#undef GK_CASE
}

hipError_t launch_adj(int p0, int p1, int rm, const AdjArgs &a, int blocks, hipStream_t s)
{
    if (a.p.nx > kAdjNxMax || blocks < 1) return hipErrorInvalidValue;
#define GK_PAIR(A, B) \
    if (p0 == (A) && p1 == (B)) return launch_adj_pd<A, B>(rm, a, blocks, s);
    GK_PAIR(1, 0)
    GK_PAIR(1, 1)
    GK_PAIR(2, 1)
    GK_PAIR(2, 2)
    GK_PAIR(3, 2)
    GK_PAIR(3, 3)
#undef GK_PAIR
    return hipErrorInvalidValue;
}

// workgroups of the reverse pass resident on one CU, at the largest supported
// grid (one launch geometry serves every nx, so the deposit sums of a run
// always add the same rows in the same order)
int adj_blocks_per_cu(int p0, int p1, int rm)
{
    int nb = 0;
#define GK_Q(A, B)                                                                              \
    if (p0 == (A) && p1 == (B)) {                                                               \
        const size_t s_ = adj_shmem_bytes(kAdjNxMax, adj_slots<A, B>());                        \
        if (rm & R_REV)                                                                         \
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adj1d<A, B, R_REV | R_PREV>, \
                                                               kBlock, s_);                      \
        else                                                                                    \
            (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_adj1d<A, B, R_FWD | R_PREV>, \
                                                               kBlock, s_);                      \
    }
    GK_Q(1, 0)
    GK_Q(1, 1)
    GK_Q(2, 1)
    GK_Q(2, 2)
    GK_Q(3, 2)
    GK_Q(3, 3)
#undef GK_Q
    return nb > 0 ? nb : 1;
}

template <int P0, int P1>
static hipError_t launch_adj_batch_pd(int rm, const AdjBatch &b, int total, hipStream_t s)
{
    const size_t sh = adj_shmem_bytes(b.max_nx, adj_slots<P0, P1>());
#define GK_CASE(M)                                                                              \
    case (M):                                                                                   \
        hipLaunchKernelGGL((k_adj1d_batch<P0, P1, (M)>), dim3(total), dim3(kBlock), sh, s, b); \
        return hipGetLastError();
    switch (rm) {
        GK_CASE(R_FWD)
        GK_CASE(R_FWD | R_PREV)
        GK_CASE(R_REV)
        GK_CASE(R_REV | R_PREV)
        GK_CASE(R_HEONLY)
    default:
        return hipErrorInvalidValue;
    }
#undef GK_CASE
}

hipError_t launch_adj_batch(int p0, int p1, int rm, const AdjBatch &b, int total_blocks,
                            hipStream_t s)
{
    if (b.max_nx > kAdjNxMax || total_blocks < 1 || b.ninst < 1) return hipErrorInvalidValue;
#define GK_PAIR(A, B) \
    if (p0 == (A) && p1 == (B)) return launch_adj_batch_pd<A, B>(rm, b, total_blocks, s);
    GK_PAIR(1, 0)
    GK_PAIR(1, 1)
    GK_PAIR(2, 1)
    GK_PAIR(2, 2)
    GK_PAIR(3, 2)
    GK_PAIR(3, 3)
#undef GK_PAIR
    return hipErrorInvalidValue;
}

hipError_t launch_adj_epi_batch(int kind, const AdjBatch &b, int total_groups, hipStream_t s)
{
    if (b.max_nx > kAdjNxMax || total_groups < 1 || b.ninst < 1) return hipErrorInvalidValue;
    const size_t sh = adj_epi_shmem_bytes(b.max_nx);
#define GK_K(K)                                                                                        \
    case (K):                                                                                          \
        hipLaunchKernelGGL(k_adj_epi_batch_sum<K>, dim3(total_groups), dim3(kBlock), sh, s, b);    \
        hipLaunchKernelGGL(k_adj_epi_batch_fin<K>, dim3(b.ninst), dim3(kBlock), sh, s, b);          \
        return hipGetLastError();
    switch (kind) {
        GK_K(AE_TAIL)
        GK_K(AE_STEP)
        GK_K(AE_HEAD)
    default:
        return hipErrorInvalidValue;
    }
#undef GK_K
}

hipError_t launch_adj_epi(int kind, const AdjArgs &a, int pass_blocks, hipStream_t s)
{
    if (a.p.nx > kAdjNxMax) return hipErrorInvalidValue;
    const size_t sh = adj_epi_shmem_bytes(a.p.nx);
    const dim3 grid(kind == AE_POIS ? 1 : epi_groups(pass_blocks));
    switch (kind) {
    case AE_TAIL: hipLaunchKernelGGL(k_adj_epi<AE_TAIL>, grid, dim3(kBlock), sh, s, a, pass_blocks); break;
    case AE_STEP: hipLaunchKernelGGL(k_adj_epi<AE_STEP>, grid, dim3(kBlock), sh, s, a, pass_blocks); break;
    case AE_HEAD: hipLaunchKernelGGL(k_adj_epi<AE_HEAD>, grid, dim3(kBlock), sh, s, a, pass_blocks); break;
    case AE_POIS: hipLaunchKernelGGL(k_adj_epi<AE_POIS>, grid, dim3(kBlock), sh, s, a, 0); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

__global__ void k_adj_rows_in(const double *aos, int64_t ld, int64_t n, const int64_t *id,
                              double *r0, double *r1, double *r2)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = id[i];
        if (j < 0 || j >= n) continue;
        const double *p = aos + j * ld;
        r0[i] = p[0];
        r1[i] = p[1];
        r2[i] = p[2];
    }
}

__global__ void k_adj_rows_out(double *aos, int64_t ld, int64_t n, const int64_t *id,
                               const double *r0, const double *r1, const double *r2)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = id[i];
        if (j < 0 || j >= n) continue;
        double *p = aos + j * ld;
        p[0] = r0[i];
        p[1] = r1[i];
        p[2] = r2[i];
        for (int64_t r = 3; r < ld; ++r) p[r] = 0.0;
    }
}

__global__ void k_adj_unsort(const int *perm, int64_t n, const double *src, double *dst)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t j = perm[i];
        if (j < 0 || j >= n) continue;
        dst[j] = src[i];
        dst[n + j] = src[n + i];
        dst[2 * n + j] = src[2 * n + i];
    }
}

hipError_t launch_adj_unsort(const int *perm, int64_t n, const double *src, double *dst,
                             hipStream_t s)
{
    int blocks = (int)std::min<int64_t>(4096, std::max<int64_t>(1, (n + 255) / 256));
    hipLaunchKernelGGL(k_adj_unsort, dim3(blocks), dim3(256), 0, s, perm, n, src, dst);
    return hipGetLastError();
}

hipError_t launch_adj_rows_in(const double *aos, int64_t ld, int64_t n, const int64_t *id,
                              double *r0, double *r1, double *r2, hipStream_t s)
{
    int blocks = (int)std::min<int64_t>(4096, std::max<int64_t>(1, (n + 255) / 256));
    hipLaunchKernelGGL(k_adj_rows_in, dim3(blocks), dim3(256), 0, s, aos, ld, n, id, r0, r1, r2);
    return hipGetLastError();
}

hipError_t launch_adj_rows_out(double *aos, int64_t ld, int64_t n, const int64_t *id,
                               const double *r0, const double *r1, const double *r2, hipStream_t s)
{
    int blocks = (int)std::min<int64_t>(4096, std::max<int64_t>(1, (n + 255) / 256));
    hipLaunchKernelGGL(k_adj_rows_out, dim3(blocks), dim3(256), 0, s, aos, ld, n, id, r0, r1, r2);
    return hipGetLastError();
}

}  // namespace gl
