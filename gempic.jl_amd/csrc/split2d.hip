// split2d.hip -- the 2D3V Hamiltonian splitting on the GPU (FP64).
//
// The reference includes hamiltonian_splitting_2d3v.jl (src/GEMPIC.jl:33)
// but the file is empty. This is the splitting of the GEMPIC paper
// (H = Hp1 + Hp2 + Hp3 + HE + HB) on the reference's 2D pieces: the 2D
// coupling's shape factor (particle_mesh_coupling_2d.jl:54-82) with a spline
// degree per direction, and TwoDMaxwell's field operations
// (maxwell_2d_fem.jl:370-499, run by pic2d.hip on device arrays). Fields:
// E1 (p-1, p), E2 (p, p-1), E3 (p, p); B1 (p, p-1), B2 (p-1, p),
// B3 (p-1, p-1). Particles: SoA x1, x2, v1, v2, v3, w, kept sorted by cell.
//
// Each operator is one particle pass. A workgroup stages the fields it reads
// in LDS, with a periodic halo of p nodes on the low side of each direction
// (so a particle's stencil is constant offsets from its cell, no wrap), and
// deposits its current into an LDS grid: each lane keeps the node sums of
// the patch its particles fall in; when a lane's next contribution starts
// another patch, the wave flushes (DPP sums into one patch when many
// flushing lanes hold the same one, LDS atomics otherwise). The workgroup
// grids are summed over workgroups in a fixed order. The shape factors are
// evaluated in closed form and the gathers and deposits summed with fused
// multiply-adds (the oracle: de Boor, every product and sum rounded); Hp1
// and Hp2 take the path integrals of the degree p-1 splines from their
// antiderivative (suffix sums of the degree-p values at the two ends) where
// the oracle integrates piece by piece: the same quantities, rounded
// differently; the positions follow the oracle's operations exactly. The
// deposit's additions happen in another order (and, between lanes that
// flush different patches, in a timing-dependent one).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/gempic_hip.h"
#include "device_common.hpp"
#include "host_setup.hpp"
#include "lib_internal.hpp"

using namespace gl;
using namespace gk;

namespace {

// threads per workgroup of a particle pass: 1024 (4 waves per SIMD, at most
// 128 VGPRs per lane), 768 (3 waves per SIMD, up to 168) or 512 (2 waves
// per SIMD, up to 256). The passes with a deposit and B gathers hold the
// lane's patch cache, the shape factors and the staged-field reads at once:
// Hp2 fits 127 registers, Hp1 needs 160 (at 1024 threads it spilled 28
// registers to scratch and took 634 us a pass on 16,777,216 particles,
// the step 3.33 ms; at 768 threads the step takes 3.13 ms) and Hp3 136 (at
// 768 threads the step takes 3.03 ms, at 512 3.13). The LDS of a
// workgroup (staged fields and the deposit grid, above 80 KB) allows one
// workgroup per CU in every case.
#ifndef GEMPIC_S2_THREADS_HP1
#define GEMPIC_S2_THREADS_HP1 768
#endif
#ifndef GEMPIC_S2_THREADS_HP2
#define GEMPIC_S2_THREADS_HP2 1024
#endif
#ifndef GEMPIC_S2_THREADS_HP3
#define GEMPIC_S2_THREADS_HP3 768
#endif
constexpr int s2_threads(int op)
{
    return op == 1 ? GEMPIC_S2_THREADS_HP1 : op == 2 ? GEMPIC_S2_THREADS_HP2 : (op == 3 ? GEMPIC_S2_THREADS_HP3 : 1024);
}
constexpr int kNn2Max = 4096;              // nx * ny of the device path (LDS grids)
#ifndef GEMPIC_S2_HE_PREFETCH
#define GEMPIC_S2_HE_PREFETCH 1
#endif
#ifndef GEMPIC_S2_GEN_SKEW
#define GEMPIC_S2_GEN_SKEW 0.25
#endif
constexpr double kS2GenSkew = GEMPIC_S2_GEN_SKEW;   // chunk weights of a workgroup's waves (s2_args)
#ifndef GEMPIC_S2_SPLITS
#define GEMPIC_S2_SPLITS 16
#endif
constexpr int kSplits2 = GEMPIC_S2_SPLITS;  // first-level groups of the workgroup-grid sum

enum S2Op { S2_HE = 0, S2_HP1, S2_HP2, S2_HP3, S2_RHO };
static_assert(S2_HP1 == 1 && S2_HP2 == 2 && S2_HP3 == 3, "s2_threads names the passes by value");

constexpr int kS2MaxWaves = 16;            // waves of a 2D3V particle-pass workgroup, at most
static_assert(s2_threads(S2_HE) <= kS2MaxWaves * 64 && s2_threads(S2_HP1) <= kS2MaxWaves * 64 &&
                  s2_threads(S2_HP2) <= kS2MaxWaves * 64 && s2_threads(S2_HP3) <= kS2MaxWaves * 64 &&
                  s2_threads(S2_HP1) % 256 == 0 && s2_threads(S2_HP2) % 256 == 0 &&
                  s2_threads(S2_HP3) % 256 == 0,
              "a particle pass runs whole waves on each of the 4 SIMDs, at most kS2MaxWaves");

struct S2Args {
    int64_t n;
    int64_t chunk;                  // particles per wave on average (multiple of 64)
    // wave w of a workgroup takes wch[w] particles from the workgroup's first
    // (blockIdx.x * waves * chunk) plus woff[w] (both multiples of 64)
    int64_t wch[kS2MaxWaves], woff[kS2MaxWaves];
    int nx, ny;
    double xmin, ymin, dx, dy, rdx, rdy, Lx, Ly;
    double charge, cw, qm, dt;
    double *x1, *x2, *v1, *v2, *v3;
    const double *w;
    const double *f[3];             // fields read by the pass (staged in LDS)
    double *part;                   // [workgroups][nx*ny] deposit grids
    int *err;
};

// compute_shape_factor (particle_mesh_coupling_2d.jl:54-65) in one direction:
// ceil() cell, offset in (0, 1], the degree-P and degree-(P-1) values there
// (the same de Boor recurrence); `cell` wrapped into [0, n)
template <int P>
__device__ __forceinline__ bool shape_dir(double pos, double lo, double d, double rd, int n,
                                          int &cell, double (&b0)[P + 1], double (&b1)[P])
{
    const double xp = div_by(pos - lo, d, rd);
    const bool ok = fabs(xp) < 1.0e9;
    const int ip = ok ? (int)ceil(xp) : 1;
    const double off = ok ? xp - (double)(ip - 1) : 1.0;
    bspline_pair_fast<P, P - 1>(off, b0, b1);
    cell = imod(ip - 1, n);
    return ok;
}

// staged field: node (i, j) of the periodic grid, i in [-P, nx), j in [-P, ny),
// at fs[(j + P) * (nx + P) + i + P]
template <int P>
__device__ void stage_field(double *fs, const double *f, int nx, int ny)
{
    const int W = nx + P, H = ny + P;
    for (int k = threadIdx.x; k < W * H; k += blockDim.x) {
        const int r = k / W, c = k - r * W;
        int i = c - P, j = r - P;
        if (i < 0) i += nx;
        if (j < 0) j += ny;
        fs[k] = f[i + j * nx];
    }
}

// evaluate (particle_mesh_coupling_2d.jl:189-203) of a staged field of
// degrees (DX, DY) at cell (cx, cy): sum over x then y of f * Sx * Sy
template <int P, int DX, int DY>
__device__ __forceinline__ double gather2(const double *fs, int W, int cx, int cy,
                                          const double *bx, const double *by)
{
    // sum_c by[c] (sum_a f[c][a] bx[a]), with fused multiply-adds
    const double *base = fs + (cy - DY + P) * W + (cx - DX + P);
    double v = 0.0;
#pragma unroll
    for (int c = 0; c <= DY; ++c) {
        double r = 0.0;
#pragma unroll
        for (int a = 0; a <= DX; ++a) r = __builtin_fma(base[c * W + a], bx[a], r);
        v = __builtin_fma(r, by[c], v);
    }
    return v;
}

// Lane cache of one deposit patch of WX x WY nodes (x slowest: value a*WY + c)
// starting at node `key` = sx + sy * nx; flushed into the workgroup's LDS grid.
template <int WX, int WY>
struct Patch {
    int key;
    double acc[WX * WY];
};

// Row stride of the workgroup's LDS deposit grid, in doubles: congruent to 4
// modulo 16, so that the 16 nodes of a 4 x 4 patch (column a, row c at
// a + 4 c modulo 16) fall in 16 different pairs of the 32 four-byte banks.
// With rows of nx = 64 doubles every node of a column shared one bank.
__host__ __device__ __forceinline__ int grid_stride(int nx) { return nx + ((20 - nx % 16) % 16); }

// LDS index of node j of the patch that starts at node `key` = sx + sy * nx
template <int WX, int WY>
__device__ __forceinline__ int patch_node(int key, int j, int nx, int ny)
{
    const int sy = key / nx, sx = key - sy * nx;
    const int a = j / WY, c = j - a * WY;
    int gx = sx + a, gy = sy + c;
    if (gx >= nx) gx -= nx;
    if (gy >= ny) gy -= ny;
    return gx + gy * grid_stride(nx);
}

// Flush of the lanes with `evict` set. A patch held by at least kDppLanes of
// them (checked for the patches of the first kFlushKeys evicting lanes in
// turn) is summed over those lanes with DPP and added by 16 lanes at once;
// the lanes of the other patches add their own nodes with LDS atomics
// (same-node atomics in one instruction serialize, so a patch shared by
// many lanes is cheaper as one sum).
#ifndef GEMPIC_S2_DPP_LANES
#define GEMPIC_S2_DPP_LANES 64
#endif
constexpr int kDppLanes = GEMPIC_S2_DPP_LANES;
constexpr int kFlushKeys = 4;

template <int WX, int WY>
__device__ __forceinline__ void patch_flush(Patch<WX, WY> &pc, bool evict, double *grid,
                                            double *stage, int nx, int ny)
{
    constexpr int WW = WX * WY;
    const int lane = lane_id();
    unsigned long long pending = __ballot(evict);
    if (!pending) return;
    const unsigned long long all = pending;
    for (int r = 0; r < kFlushKeys && pending; ++r) {
        const int k0 = __shfl(pc.key, __ffsll((long long)pending) - 1, kWave);
        const unsigned long long mk = __ballot(evict && pc.key == k0);
        pending &= ~mk;
        if (__popcll(mk) < (kDppLanes < __popcll(all) ? kDppLanes : __popcll(all))) continue;
        const bool mine = evict && pc.key == k0;
#pragma unroll
        for (int j = 0; j < WW; ++j) {
            const double sj = wave_sum_lane63(mine ? pc.acc[j] : 0.0);
            if (lane == kWave - 1) stage[j] = sj;
        }
        if (lane < WW) atomicAdd(&grid[patch_node<WX, WY>(k0, lane, nx, ny)], stage[lane]);
        if (mine) evict = false;
    }
    if (evict) {
#pragma unroll
        for (int j = 0; j < WW; ++j) atomicAdd(&grid[patch_node<WX, WY>(pc.key, j, nx, ny)], pc.acc[j]);
    }
}

// make patch `key` the lane's cached patch for the lanes with `act`
// (flushing the one they held); all lanes call. The caller then adds its
// values to pc.acc directly.
template <int WX, int WY>
__device__ __forceinline__ void patch_begin(Patch<WX, WY> &pc, bool act, int key, double *grid,
                                            double *stage, int nx, int ny)
{
    const bool miss = act && key != pc.key;
    patch_flush<WX, WY>(pc, miss && pc.key >= 0, grid, stage, nx, ny);
    if (miss) {
        pc.key = key;
#pragma unroll
        for (int j = 0; j < WX * WY; ++j) pc.acc[j] = 0.0;
    }
}

// add vals (WX x WY) at patch `key` for the lanes with `act`; all lanes call
template <int WX, int WY>
__device__ __forceinline__ void patch_add(Patch<WX, WY> &pc, bool act, int key, const double *vals,
                                          double *grid, double *stage, int nx, int ny)
{
    patch_begin<WX, WY>(pc, act, key, grid, stage, nx, ny);
    if (act) {
#pragma unroll
        for (int j = 0; j < WX * WY; ++j) pc.acc[j] += vals[j];
    }
}

// the largest number of cells a particle may cross in one push
constexpr int kMaxCross2 = 1 << 16;

template <int P, int OP>
__global__ void __launch_bounds__(s2_threads(OP)) k_split2d(S2Args a)
{
    constexpr bool DEP = OP != S2_HE;
    constexpr int NF = OP == S2_HE ? 3 : (OP == S2_RHO ? 0 : 2);
    // deposit patch shape: j1 (P-1, P), j2 (P, P-1), j3 and rho (P, P)
    constexpr int WX = P + 1;
    constexpr int WY = P + 1;
    extern __shared__ __align__(16) double smem[];
    const int nx = a.nx, ny = a.ny, nn = nx * ny;
    const int W = nx + P;
    const int fsz = W * (ny + P);
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    double *fs0 = smem, *fs1 = smem + fsz, *fs2 = smem + 2 * fsz;
    double *grid = smem + NF * fsz;
    const int gsz = grid_stride(nx) * ny;
    double *stage = grid + (DEP ? gsz : 0) + wave * 16;
    if (NF > 0) stage_field<P>(fs0, a.f[0], nx, ny);
    if (NF > 1) stage_field<P>(fs1, a.f[1], nx, ny);
    if (NF > 2) stage_field<P>(fs2, a.f[2], nx, ny);
    if (DEP)
        for (int k = threadIdx.x; k < gsz; k += blockDim.x) grid[k] = 0.0;
    __syncthreads();

    Patch<WX, WY> pc;
    pc.key = -1;
#pragma unroll
    for (int j = 0; j < WX * WY; ++j) pc.acc[j] = 0.0;
    int errs = 0;
    const double qm = a.qm, dt = a.dt;
    constexpr int kW2 = s2_threads(OP) / kWave;
    const int64_t wbase = (int64_t)blockIdx.x * kW2 * a.chunk + a.woff[wave];
    const int64_t wlen = a.wch[wave];
    if constexpr (OP == S2_HE && GEMPIC_S2_HE_PREFETCH) {
        // HE reads the state and the staged E only: the next particle's
        // state is loaded while this one's gathers run
        int64_t i = wbase + lane;
        bool valid = 0 < wlen && i < a.n;
        int64_t ii = valid ? i : 0;
        double nx1 = a.x1[ii], nx2 = a.x2[ii], nv1 = a.v1[ii], nv2 = a.v2[ii], nv3 = a.v3[ii];
        for (int64_t off = 0; off < wlen; off += kWave) {
            if (!__ballot(valid)) break;
            const double x1 = nx1, x2 = nx2;
            double v1 = nv1, v2 = nv2, v3 = nv3;
            const int64_t ic = i;
            const bool vc = valid;
            i += kWave;
            valid = off + kWave < wlen && i < a.n;
            ii = valid ? i : ii;
            nx1 = a.x1[ii];
            nx2 = a.x2[ii];
            nv1 = a.v1[ii];
            nv2 = a.v2[ii];
            nv3 = a.v3[ii];
            int cx, cy;
            double bx0[P + 1], bx1[P], by0[P + 1], by1[P];
            bool ok = shape_dir<P>(x1, a.xmin, a.dx, a.rdx, nx, cx, bx0, bx1);
            ok = shape_dir<P>(x2, a.ymin, a.dy, a.rdy, ny, cy, by0, by1) && ok;
            if (vc && !ok) errs |= E_INEXACT;
            const double e1 = gather2<P, P - 1, P>(fs0, W, cx, cy, bx1, by0);
            const double e2 = gather2<P, P, P - 1>(fs1, W, cx, cy, bx0, by1);
            const double e3 = gather2<P, P, P>(fs2, W, cx, cy, bx0, by0);
            const double dtqm = dt * qm;
            if (vc) {
                a.v1[ic] = __builtin_fma(dtqm, e1, v1);
                a.v2[ic] = __builtin_fma(dtqm, e2, v2);
                a.v3[ic] = __builtin_fma(dtqm, e3, v3);
            }
        }
    } else
    for (int64_t off = 0; off < wlen; off += kWave) {
        const int64_t i = wbase + off + lane;
        const bool valid = i < a.n;
        if (!__ballot(valid)) break;
        const int64_t ii = valid ? i : 0;
        double x1 = a.x1[ii], x2 = a.x2[ii];
        double v1 = 0.0, v2 = 0.0, v3 = 0.0;
        if (OP != S2_RHO) {
            v1 = a.v1[ii];
            v2 = a.v2[ii];
            v3 = a.v3[ii];
        }
        const double q = (a.w[ii] * a.charge) * a.cw;
        int cx, cy;
        double bx0[P + 1], bx1[P], by0[P + 1], by1[P];
        bool ok = shape_dir<P>(x1, a.xmin, a.dx, a.rdx, nx, cx, bx0, bx1);
        ok = shape_dir<P>(x2, a.ymin, a.dy, a.rdy, ny, cy, by0, by1) && ok;
        if (valid && !ok) errs |= E_INEXACT;
        if constexpr (OP == S2_HE) {
            const double e1 = gather2<P, P - 1, P>(fs0, W, cx, cy, bx1, by0);
            const double e2 = gather2<P, P, P - 1>(fs1, W, cx, cy, bx0, by1);
            const double e3 = gather2<P, P, P>(fs2, W, cx, cy, bx0, by0);
            const double dtqm = dt * qm;
            v1 = __builtin_fma(dtqm, e1, v1);
            v2 = __builtin_fma(dtqm, e2, v2);
            v3 = __builtin_fma(dtqm, e3, v3);
        } else if constexpr (OP == S2_HP3 || OP == S2_RHO) {
            double wq = q;
            if constexpr (OP == S2_HP3) {
                const double b1 = gather2<P, P, P - 1>(fs0, W, cx, cy, bx0, by1);
                const double b2 = gather2<P, P - 1, P>(fs1, W, cx, cy, bx1, by0);
                const double dv = (dt * qm) * v3;
                v1 = __builtin_fma(-dv, b2, v1);
                v2 = __builtin_fma(dv, b1, v2);
                wq = q * v3;
            }
            patch_begin<WX, WY>(pc, valid, imod(cx - P, nx) + imod(cy - P, ny) * nx, grid, stage,
                                nx, ny);
            if (valid) {
#pragma unroll
                for (int aa = 0; aa <= P; ++aa) {
                    const double wb = wq * bx0[aa];
#pragma unroll
                    for (int c = 0; c <= P; ++c) pc.acc[aa * WY + c] = __builtin_fma(wb, by0[c], pc.acc[aa * WY + c]);
                }
            }
        } else {
            // Hp1 moves x1 (fields fs0 = B3, fs1 = B2), Hp2 moves x2 (fs0 = B3, fs1 = B1)
            constexpr bool ALONG_X = OP == S2_HP1;
            const double pos = ALONG_X ? x1 : x2;
            const double vel = ALONG_X ? v1 : v2;
            const double lo = ALONG_X ? a.xmin : a.ymin;
            const double d = ALONG_X ? a.dx : a.dy, rd = ALONG_X ? a.rdx : a.rdy;
            const double pn = pos + dt * vel;
            const double xa = div_by(pos - lo, d, rd), xb = div_by(pn - lo, d, rd);
            const bool far = !(fabs(xa) < 1.0e9) || !(fabs(xb) < 1.0e9) ||
                             fabs(floor(xb) - floor(xa)) > (double)kMaxCross2;
            if (valid && far) errs |= E_FAR;
            const double fa = far ? 0.0 : floor(xa), fb = far ? 0.0 : floor(xb);
            const int ca = (int)fa, cb = (int)fb;
            const int clo = ca < cb ? ca : cb, chi = ca < cb ? cb : ca;
            // The path's integrals of the degree P-1 splines from their
            // antiderivative F_k = sum_{m >= k} N^P_m (d/dx N^P_m = N^{P-1}_m -
            // N^{P-1}_{m+1}): int_xa^xb N^{P-1}_k = F_k(xb) - F_k(xa), where F_k
            // at an end in cell c is 1 for k <= c - P, 0 for k > c and otherwise
            // the suffix sum from k - (c - P) of the degree-P values there. The
            // nodes clo - P + 1 .. chi (P of them within one cell, P + 1 across
            // one boundary) go in chunks of P + 1, one patch each: one chunk for
            // nearly every particle, and no pieces.
            const int nch = far ? 0 : (chi - clo + P + P) / (P + 1);
            double sa[P + 1], sb[P + 1];        // suffix sums sa[j] = sum_{i >= j} ia[i]
            {
                double ia[P + 1], ib[P + 1];
                bspline_fast<P>(xa - fa, ia);
                bspline_fast<P>(xb - fb, ib);
                sa[P] = ia[P];
                sb[P] = ib[P];
#pragma unroll
                for (int j = P - 1; j >= 1; --j) {
                    sa[j] = sa[j + 1] + ia[j];
                    sb[j] = sb[j + 1] + ib[j];
                }
                sa[0] = 1.0;
                sb[0] = 1.0;
            }
            // F at node k for an end in cell c: index j = k - c + P
            auto F = [](const double (&sfx)[P + 1], int j) {
                double f = j <= 0 ? 1.0 : 0.0;
#pragma unroll
                for (int t = 1; t <= P; ++t) f = j == t ? sfx[t] : f;
                return f;
            };
            double u = ALONG_X ? v2 : v1;         // Hp1: v2 (with B3); Hp2: v1 (with B3)
            double t3 = v3;                        // both: v3 (with B2 or B1)
            const int nd = ALONG_X ? nx : ny;
            for (int m = 0;; ++m) {
                const bool act = valid && m < nch;
                if (!__ballot(act)) break;
                const int kb = clo - P + 1 + m * (P + 1);          // first node of the chunk
                double val[P + 1];
#pragma unroll
                for (int r = 0; r <= P; ++r)
                    val[r] = d * (F(sb, kb + r - cb + P) - F(sa, kb + r - ca + P));
                int k0 = imod(kb, nd);             // in [-P, nd - 1 - P]: within the staging
                if (k0 + P > nd - 1) k0 -= nd;
                if constexpr (ALONG_X) {
                    // j1 += q (dx int S^{p-1}) S^p(x2); v3 += qm int B2 ..; v2 -= qm int B3 ..
                    patch_begin<WX, WY>(pc, act, imod(k0, nx) + imod(cy - P, ny) * nx, grid, stage,
                                        nx, ny);
                    if (act) {
                        // per node r of the path: the deposit q val[r] S^p(x2), and
                        // the B2 and B3 gathers along x2 weighted by val[r]
                        const double *b2 = fs1 + (cy - P + P) * W + (k0 + P);
                        const double *b3 = fs0 + (cy - (P - 1) + P) * W + (k0 + P);
                        double s2 = 0.0, s3 = 0.0;
#pragma unroll
                        for (int r = 0; r <= P; ++r) {
                            const double qv = q * val[r];
                            double g2 = 0.0, g3 = 0.0;
#pragma unroll
                            for (int cc = 0; cc <= P; ++cc) {
                                pc.acc[r * WY + cc] = __builtin_fma(qv, by0[cc], pc.acc[r * WY + cc]);
                                g2 = __builtin_fma(b2[cc * W + r], by0[cc], g2);
                            }
#pragma unroll
                            for (int cc = 0; cc < P; ++cc) g3 = __builtin_fma(b3[cc * W + r], by1[cc], g3);
                            s2 = __builtin_fma(val[r], g2, s2);
                            s3 = __builtin_fma(val[r], g3, s3);
                        }
                        t3 = __builtin_fma(qm, s2, t3);
                        u = __builtin_fma(-qm, s3, u);
                    }
                } else {
                    // j2 += q S^p(x1) (dy int S^{p-1}); v3 -= qm int B1 ..; v1 += qm int B3 ..
                    patch_begin<WX, WY>(pc, act, imod(cx - P, nx) + imod(k0, ny) * nx, grid, stage,
                                        nx, ny);
                    if (act) {
                        // per node r of the path: the deposit q S^p(x1) val[r], and
                        // the B1 and B3 gathers along x1 weighted by val[r]
                        const double *b1 = fs1 + (k0 + P) * W + (cx - P + P);
                        const double *b3 = fs0 + (k0 + P) * W + (cx - (P - 1) + P);
                        double s1 = 0.0, s3 = 0.0;
#pragma unroll
                        for (int r = 0; r <= P; ++r) {
                            const double qv = q * val[r];
                            double g1 = 0.0, g3 = 0.0;
#pragma unroll
                            for (int aa = 0; aa <= P; ++aa) {
                                pc.acc[aa * WY + r] = __builtin_fma(qv, bx0[aa], pc.acc[aa * WY + r]);
                                g1 = __builtin_fma(b1[r * W + aa], bx0[aa], g1);
                            }
#pragma unroll
                            for (int aa = 0; aa < P; ++aa) g3 = __builtin_fma(b3[r * W + aa], bx1[aa], g3);
                            s1 = __builtin_fma(val[r], g1, s1);
                            s3 = __builtin_fma(val[r], g3, s3);
                        }
                        t3 = __builtin_fma(-qm, s1, t3);
                        u = __builtin_fma(qm, s3, u);
                    }
                }
            }
            if (ALONG_X) {
                v2 = u;
                x1 = a.xmin + jl_mod(pn - a.xmin, a.Lx);
            } else {
                v1 = u;
                x2 = a.ymin + jl_mod(pn - a.ymin, a.Ly);
            }
            v3 = t3;
        }
        if (valid && OP != S2_RHO) {
            if (OP == S2_HP1) a.x1[i] = x1;
            if (OP == S2_HP2) a.x2[i] = x2;
            if (OP != S2_HP1) a.v1[i] = v1;
            if (OP != S2_HP2) a.v2[i] = v2;
            if (OP != S2_HP3) a.v3[i] = v3;
        }
    }
    if (errs) atomicOr(a.err, errs);
    if constexpr (DEP) {
        patch_flush<WX, WY>(pc, pc.key >= 0, grid, stage, nx, ny);
        __syncthreads();
        double *out = a.part + (size_t)blockIdx.x * nn;
        const int gs = grid_stride(nx);
        for (int k = threadIdx.x; k < nn; k += blockDim.x) {
            const int gy = k / nx;
            out[k] = grid[k + gy * (gs - nx)];
        }
    }
}

// part2[s][k] = sum over the workgroups of split s of part[b][k], in order
__global__ void k_s2_reduce1(const double *part, int nblocks, int nn, double *part2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const int splits = gridDim.y, sp = blockIdx.y;
    const int b0 = (int)((int64_t)sp * nblocks / splits), b1 = (int)((int64_t)(sp + 1) * nblocks / splits);
    double t = 0.0;
    for (int b = b0; b < b1; ++b) t += part[(size_t)b * nn + k];
    part2[(size_t)sp * nn + k] = t;
}

// out[k] = (sum of part2[s][k] over the splits, in order) * scale
__global__ void k_s2_reduce2(const double *part2, int splits, int nn, double scale, double *out)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    double t = 0.0;
    for (int s = 0; s < splits; ++s) t += part2[(size_t)s * nn + k];
    out[k] = t * scale;
}

// ---- sort by cell (counting sort over the nx*ny cells) ----

__device__ __forceinline__ int cell_key(double x1, double x2, const S2Args &a)
{
    const double xp = div_by(x1 - a.xmin, a.dx, a.rdx), yp = div_by(x2 - a.ymin, a.dy, a.rdy);
    if (!(fabs(xp) < 1.0e9) || !(fabs(yp) < 1.0e9)) return 0;
    return imod((int)ceil(xp) - 1, a.nx) + imod((int)ceil(yp) - 1, a.ny) * a.nx;
}

// Wave-aggregated LDS counts: the lanes of a wave with one key take
// consecutive ranks in lane order from one atomic of the key's first lane
// (particles kept sorted by cell give a wave few distinct keys); after
// kAggKeys keys the remaining lanes add their own. Returns the lane's rank
// (its slot among this key's counts so far); call from every lane.
constexpr int kAggKeys = 8;

__device__ __forceinline__ int wave_agg_rank(bool valid, int key, int *cnt)
{
    const int lane = lane_id();
    const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (kWave - lane));
    unsigned long long pending = __ballot(valid);
    int rank = 0;
    for (int r = 0; r < kAggKeys && pending; ++r) {
        const int leader = __ffsll((long long)pending) - 1;
        const int k = __shfl(key, leader, kWave);
        const unsigned long long m = __ballot(valid && key == k);
        int b = 0;
        if (lane == leader) b = atomicAdd(&cnt[k], __popcll(m));
        b = __shfl(b, leader, kWave);
        if (valid && key == k) rank = b + __popcll(m & below);
        pending &= ~m;
    }
    if ((pending >> lane) & 1ull) rank = atomicAdd(&cnt[key], 1);
    return rank;
}

__global__ void __launch_bounds__(1024) k_s2_count(S2Args a, int *count)
{
    __shared__ int h[kNn2Max];
    const int nn = a.nx * a.ny;
    for (int k = threadIdx.x; k < nn; k += blockDim.x) h[k] = 0;
    __syncthreads();
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x; i0 < a.n; i0 += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = i0 + threadIdx.x;
        const bool valid = i < a.n;
        (void)wave_agg_rank(valid, valid ? cell_key(a.x1[i], a.x2[i], a) : 0, h);
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nn; k += blockDim.x)
        if (h[k]) atomicAdd(&count[k], h[k]);
}

// exclusive scan of count[0..nn) into cursor (one workgroup of 1024)
__global__ void __launch_bounds__(1024) k_s2_scan(const int *count, int nn, int *cursor)
{
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (nn + 1023) / 1024;
    int s = 0;
    for (int k = t * per; k < min(nn, (t + 1) * per); ++k) s += count[k];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int base = t > 0 ? part[t - 1] : 0;
    for (int k = t * per; k < min(nn, (t + 1) * per); ++k) {
        cursor[k] = base;
        base += count[k];
    }
}

struct S2Perm {
    const double *src[6];
    double *dst[6];
    const int64_t *id;
    int64_t *id_dst;
};

// each round: 1024 particles of this workgroup take consecutive slots of
// their cell, reserved per cell with one global atomic; within a wave the
// lanes of a cell take consecutive slots in lane order (coalesced stores)
__global__ void __launch_bounds__(1024) k_s2_scatter(S2Args a, int *cursor, S2Perm pm)
{
    __shared__ int cnt[kNn2Max];
    __shared__ int base[kNn2Max];
    const int nn = a.nx * a.ny;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t r0 = (int64_t)blockIdx.x * blockDim.x; r0 < a.n; r0 += stride) {
        for (int k = threadIdx.x; k < nn; k += blockDim.x) cnt[k] = 0;
        __syncthreads();
        const int64_t i = r0 + threadIdx.x;
        const bool valid = i < a.n;
        const int key = valid ? cell_key(pm.src[0][i], pm.src[1][i], a) : 0;
        const int rank = wave_agg_rank(valid, key, cnt);
        __syncthreads();
        for (int k = threadIdx.x; k < nn; k += blockDim.x)
            if (cnt[k]) base[k] = atomicAdd(&cursor[k], cnt[k]);
        __syncthreads();
        if (valid) {
            const int64_t d = (int64_t)base[key] + rank;
#pragma unroll
            for (int r = 0; r < 6; ++r) pm.dst[r][d] = pm.src[r][i];
            pm.id_dst[d] = pm.id[i];
        }
        __syncthreads();
    }
}

// AoS (ld x n, rows x1 x2 v1 v2 v3 w) <-> SoA; the device order is id's
__global__ void k_s2_in(const double *aos, int64_t ld, int64_t n, S2Perm pm)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int r = 0; r < 6; ++r) pm.dst[r][i] = aos[r + i * ld];
        pm.id_dst[i] = i;
    }
}

__global__ void k_s2_out(double *aos, int64_t ld, int64_t n, S2Perm pm)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t o = pm.id[i];
#pragma unroll
        for (int r = 0; r < 6; ++r) aos[r + o * ld] = pm.src[r][i];
    }
}

}  // namespace

struct gempic_splitting2d {
    gempic_maxwell2d *m = nullptr;
    hipStream_t stream = nullptr;           // the solver's stream
    int nx = 0, ny = 0, p = 3;
    int64_t n = 0;
    double xmin = 0, ymin = 0, Lx = 0, Ly = 0, dx = 0, dy = 0;
    double charge = 1, mass = 1, cw = 1, qm = 1;
    double *d_p[6] = {}, *d_p2[6] = {};     // x1 x2 v1 v2 v3 w, and the sort's second buffer
    int64_t *d_id = nullptr, *d_id2 = nullptr;
    double *d_e[3] = {}, *d_b[3] = {}, *d_j[3] = {}, *d_rho = nullptr;
    double *d_part = nullptr;               // workgroup grids, then the split sums
    double *d_stage = nullptr;              // AoS staging of set/get_particles
    int *d_err = nullptr, *d_count = nullptr, *d_cursor = nullptr;
    int blocks = 0;
    // Off by default. Right after a sort the deposit passes are at their
    // fastest, but for the next steps, while each wave's particles still lie
    // in a few neighbouring cells, their lanes flush patches of the same
    // nodes and the LDS atomics collide: Hp2 on 16,777,216 particles (64 x
    // 64, degree 3) takes 324 us after a sort, 530 us three steps later and
    // 360 us once the order is gone. Over 200 steps: 2.33 ms per step
    // unsorted, 2.96 ms sorting every 20 steps, 2.54 every 50.
    int64_t sort_interval = 0, steps_since_sort = 0;
};

namespace {

int s2_fail_hip(const char *what, hipError_t e)
{
    return abi_fail(GEMPIC_ERR_DEVICE, (std::string(what) + ": " + hipGetErrorString(e)).c_str());
}

#define S2_TRY(what, expr)                                       \
    do {                                                         \
        hipError_t e_ = (expr);                                  \
        if (e_ != hipSuccess) return s2_fail_hip((what), e_);    \
    } while (0)

size_t s2_shmem(int op, int nx, int ny, int p)
{
    const size_t fsz = (size_t)(nx + p) * (ny + p);
    const int nf = op == S2_HE ? 3 : (op == S2_RHO ? 0 : 2);
    const bool dep = op != S2_HE;
    return sizeof(double) * (nf * fsz + (dep ? (size_t)grid_stride(nx) * ny : 0) +
                             (size_t)(s2_threads(op) / kWave) * 16);
}

const void *s2_kernel(int p, int op)
{
#define S2K(P)                                                                   \
    switch (op) {                                                                \
    case S2_HE: return (const void *)k_split2d<P, S2_HE>;                        \
    case S2_HP1: return (const void *)k_split2d<P, S2_HP1>;                      \
    case S2_HP2: return (const void *)k_split2d<P, S2_HP2>;                      \
    case S2_HP3: return (const void *)k_split2d<P, S2_HP3>;                      \
    default: return (const void *)k_split2d<P, S2_RHO>;                          \
    }
    if (p == 1) { S2K(1) }
    if (p == 2) { S2K(2) }
    S2K(3)
#undef S2K
}

S2Args s2_args(gempic_splitting2d *h, double dt, int op)
{
    S2Args a{};
    a.n = h->n;
    a.nx = h->nx;
    a.ny = h->ny;
    a.xmin = h->xmin;
    a.ymin = h->ymin;
    a.dx = h->dx;
    a.dy = h->dy;
    a.rdx = 1.0 / h->dx;
    a.rdy = 1.0 / h->dy;
    a.Lx = h->Lx;
    a.Ly = h->Ly;
    a.charge = h->charge;
    a.cw = h->cw;
    a.qm = h->qm;
    a.dt = dt;
    a.x1 = h->d_p[0];
    a.x2 = h->d_p[1];
    a.v1 = h->d_p[2];
    a.v2 = h->d_p[3];
    a.v3 = h->d_p[4];
    a.w = h->d_p[5];
    a.part = h->d_part;
    a.err = h->d_err;
    const int kw = s2_threads(op) / kWave;
    const int64_t waves = (int64_t)h->blocks * kw;
    a.chunk = ((h->n + waves - 1) / waves + kWave - 1) / kWave * kWave;
    if (a.chunk < kWave) a.chunk = kWave;
    // The waves of a workgroup share its CU's 4 SIMDs, kw / 4 to a SIMD, and
    // a SIMD issues from its oldest wave first: with equal chunks the later
    // waves of each SIMD run the end of the pass with the SIMD to themselves.
    // Wave w is generation g = w / 4 on its SIMD; chunks weighted 1 + skew
    // (G - 1 - 2g) / (G - 1) let the generations end together (the
    // workgroup's total stays kw * chunk).
    const int G = kw / 4;
    const int64_t per = a.chunk / kWave;              // waves of 64 particles per wave
    int64_t acc = 0;
    for (int w = 0; w < kw; ++w) {
        int64_t c = per;
        if (G >= 2 && kS2GenSkew > 0.0) {
            const int g = w / 4;
            const double wt = 1.0 + kS2GenSkew * (double)(G - 1 - 2 * g) / (double)(G - 1);
            c = w + 1 < kw ? std::max<int64_t>(0, std::llround((double)per * wt)) : per * kw - acc;
            if (c < 0) c = 0;
        }
        a.woff[w] = acc * kWave;
        a.wch[w] = c * kWave;
        acc += c;
    }
    return a;
}

// one particle pass; with a deposit, its workgroup grids summed into `out`
// (times scale), and with e_comp > 0 then E_{e_comp} -= the inverse 1-form
// mass of it (compute_e_from_j!)
int s2_pass(gempic_splitting2d *h, int op, double dt, const double *f0, const double *f1,
            const double *f2, double *out, double scale, int e_comp = 0)
{
    if (h->n == 0) {
        if (out) S2_TRY("2D3V deposit", hipMemsetAsync(out, 0, sizeof(double) * h->nx * h->ny, h->stream));
        return GEMPIC_OK;
    }
    S2Args a = s2_args(h, dt, op);
    a.f[0] = f0;
    a.f[1] = f1;
    a.f[2] = f2;
    const int nt = s2_threads(op), kW2 = nt / kWave;
    const int nblk = (int)std::min<int64_t>(h->blocks, (h->n + kW2 * a.chunk - 1) / (kW2 * a.chunk));
    const size_t sh = s2_shmem(op, h->nx, h->ny, h->p);
    void *args[] = {&a};
    S2_TRY("2D3V particle pass", hipLaunchKernel(s2_kernel(h->p, op), dim3(nblk), dim3(nt), args, sh,
                                                 h->stream));
    if (op != S2_HE) {
        const int nn = h->nx * h->ny;
        const int splits = std::min(nblk, kSplits2);
        double *part2 = h->d_part + (size_t)nblk * nn;
        hipLaunchKernelGGL(k_s2_reduce1, dim3((unsigned)((nn + 255) / 256), (unsigned)splits), dim3(256), 0,
                           h->stream, (const double *)h->d_part, nblk, nn, part2);
        if (!(e_comp > 0 && m2d_e_from_parts(h->m, part2, splits, scale, out, e_comp, h->d_e[e_comp - 1]))) {
            hipLaunchKernelGGL(k_s2_reduce2, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, h->stream,
                               (const double *)part2, splits, nn, scale, out);
            if (e_comp > 0) m2d_e_from_j(h->m, out, e_comp, h->d_e[e_comp - 1]);
        }
        S2_TRY("2D3V deposit sum", hipGetLastError());
    }
    return GEMPIC_OK;
}

int s2_sort(gempic_splitting2d *h)
{
    if (h->n == 0) return GEMPIC_OK;
    const int nn = h->nx * h->ny;
    S2Args a = s2_args(h, 0.0, S2_RHO);
    S2_TRY("2D3V sort", hipMemsetAsync(h->d_count, 0, sizeof(int) * nn, h->stream));
    const unsigned gb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, (h->n + 1023) / 1024));
    hipLaunchKernelGGL(k_s2_count, dim3(gb), dim3(1024), 0, h->stream, a, h->d_count);
    hipLaunchKernelGGL(k_s2_scan, dim3(1), dim3(1024), 0, h->stream, (const int *)h->d_count, nn,
                       h->d_cursor);
    S2Perm pm{};
    for (int r = 0; r < 6; ++r) {
        pm.src[r] = h->d_p[r];
        pm.dst[r] = h->d_p2[r];
    }
    pm.id = h->d_id;
    pm.id_dst = h->d_id2;
    hipLaunchKernelGGL(k_s2_scatter, dim3(gb), dim3(1024), 0, h->stream, a, h->d_cursor, pm);
    S2_TRY("2D3V sort", hipGetLastError());
    for (int r = 0; r < 6; ++r) std::swap(h->d_p[r], h->d_p2[r]);
    std::swap(h->d_id, h->d_id2);
    h->steps_since_sort = 0;
    return GEMPIC_OK;
}

int s2_finish(gempic_splitting2d *h)
{
    S2_TRY("2D3V splitting", hipStreamSynchronize(h->stream));
    int err = 0;
    S2_TRY("2D3V splitting", hipMemcpy(&err, h->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (err) {
        S2_TRY("2D3V splitting", hipMemset(h->d_err, 0, sizeof(int)));
        if (err & E_INEXACT)
            return abi_fail(GEMPIC_ERR_INEXACT, "InexactError: ceil(Int, x) of a non-finite or out-of-range position");
        return abi_fail(GEMPIC_ERR_INEXACT, "a particle crossed more than 65536 cells in one push");
    }
    return GEMPIC_OK;
}

int s2_he(gempic_splitting2d *h, double dt)
{
    int rc = s2_pass(h, S2_HE, dt, h->d_e[0], h->d_e[1], h->d_e[2], nullptr, 1.0);
    if (rc) return rc;
    m2d_b_from_e(h->m, dt, h->d_e, h->d_b);          // compute_b_from_e!
    return GEMPIC_OK;
}

int s2_hb(gempic_splitting2d *h, double dt)
{
    m2d_e_from_b(h->m, dt, h->d_b, h->d_e);          // compute_e_from_b!
    return GEMPIC_OK;
}

int s2_hp(gempic_splitting2d *h, int comp, double dt)
{
    int rc;
    // the pass, then compute_e_from_j!
    if (comp == 1) rc = s2_pass(h, S2_HP1, dt, h->d_b[2], h->d_b[1], nullptr, h->d_j[0], 1.0, 1);
    else if (comp == 2) rc = s2_pass(h, S2_HP2, dt, h->d_b[2], h->d_b[0], nullptr, h->d_j[1], 1.0, 2);
    else rc = s2_pass(h, S2_HP3, dt, h->d_b[0], h->d_b[1], nullptr, h->d_j[2], dt, 3);
    return rc;
}

}  // namespace

#define CHECK_H2(h) \
    do { if (!(h)) return abi_fail(GEMPIC_ERR_ARGUMENT, "null 2D3V splitting handle"); } while (0)

extern "C" {

int gempic_splitting2d_create(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                              double ymax, int degree, int64_t n_particles, double charge,
                              double mass, double common_weight, gempic_splitting2d **out)
{
    if (!out) return abi_fail(GEMPIC_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (degree < 1 || degree > 3)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: the 2D3V splitting takes degree 1, 2 or 3");
    if (nx < 2 * degree + 1 || ny < 2 * degree + 1 || nx * ny > kNn2Max)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: nx, ny >= 2*degree+1 and nx*ny <= 4096");
    if (n_particles < 0 || n_particles > ((int64_t)1 << 31) - 1024)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: number of particles out of range");
    if (!(mass != 0.0)) return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: mass must be nonzero");
    int rc = abi_check_device();
    if (rc) return rc;
    gempic_maxwell2d *m = nullptr;
    rc = gempic_maxwell2d_create(nx, xmin, xmax, ny, ymin, ymax, degree, &m);
    if (rc) return rc;
    gempic_splitting2d *h = new gempic_splitting2d();
    h->m = m;
    h->stream = m2d_stream(m);
    h->nx = (int)nx;
    h->ny = (int)ny;
    h->p = degree;
    h->n = n_particles;
    h->xmin = xmin;
    h->ymin = ymin;
    h->Lx = xmax - xmin;
    h->Ly = ymax - ymin;
    h->dx = h->Lx / (double)nx;         // TwoDGrid (mesh.jl:38-39)
    h->dy = h->Ly / (double)ny;
    h->charge = charge;
    h->mass = mass;
    h->cw = common_weight == 0.0 ? 1.0 / (double)std::max<int64_t>(n_particles, 1) : common_weight;
    h->qm = charge / mass;
    int dev = 0, cus = 256, maxsh = 65536;
    if (hipGetDevice(&dev) == hipSuccess) {
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    }
    h->blocks = std::max(1, cus);
    const size_t nn = (size_t)nx * ny;
    hipError_t e = hipSuccess;
    auto al = [&](void **ptr, size_t bytes) {
        if (e == hipSuccess) e = hipMalloc(ptr, std::max<size_t>(bytes, 8));
    };
    const size_t np = (size_t)std::max<int64_t>(n_particles, 1);
    for (int r = 0; r < 6; ++r) {
        al((void **)&h->d_p[r], sizeof(double) * np);
        al((void **)&h->d_p2[r], sizeof(double) * np);
    }
    al((void **)&h->d_id, sizeof(int64_t) * np);
    al((void **)&h->d_id2, sizeof(int64_t) * np);
    for (int c = 0; c < 3; ++c) {
        al((void **)&h->d_e[c], sizeof(double) * nn);
        al((void **)&h->d_b[c], sizeof(double) * nn);
        al((void **)&h->d_j[c], sizeof(double) * nn);
    }
    al((void **)&h->d_rho, sizeof(double) * nn);
    al((void **)&h->d_part, sizeof(double) * nn * (size_t)(h->blocks + kSplits2));
    al((void **)&h->d_err, sizeof(int));
    al((void **)&h->d_count, sizeof(int) * nn);
    al((void **)&h->d_cursor, sizeof(int) * nn);
    if (e == hipSuccess) e = hipMemset(h->d_err, 0, sizeof(int));
    for (int c = 0; c < 3 && e == hipSuccess; ++c) {
        e = hipMemset(h->d_e[c], 0, sizeof(double) * nn);
        if (e == hipSuccess) e = hipMemset(h->d_b[c], 0, sizeof(double) * nn);
        if (e == hipSuccess) e = hipMemset(h->d_j[c], 0, sizeof(double) * nn);
    }
    // LDS of the passes above 64 KiB (p = 3 at 64 x 64: 3 staged fields of 36 KiB)
    for (int p = 1; p <= 3 && e == hipSuccess; ++p)
        for (int op = S2_HE; op <= S2_RHO && e == hipSuccess; ++op) {
            const size_t sh = s2_shmem(op, h->nx, h->ny, p);
            if (p == degree && sh > (size_t)std::max(maxsh, 65536)) {
                gempic_splitting2d_destroy(h);
                return abi_fail(GEMPIC_ERR_ARGUMENT,
                                "ArgumentError: the 2D3V passes' staged fields do not fit in LDS");
            }
            if (p == degree && sh > 65536)
                e = hipFuncSetAttribute(s2_kernel(p, op), hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
        }
    if (e != hipSuccess) {
        gempic_splitting2d_destroy(h);
        return s2_fail_hip("2D3V splitting allocation", e);
    }
    *out = h;
    return GEMPIC_OK;
}

int gempic_splitting2d_destroy(gempic_splitting2d *h)
{
    if (!h) return GEMPIC_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (int r = 0; r < 6; ++r) {
        if (h->d_p[r]) (void)hipFree(h->d_p[r]);
        if (h->d_p2[r]) (void)hipFree(h->d_p2[r]);
    }
    void *bufs[] = {h->d_id, h->d_id2, h->d_e[0], h->d_e[1], h->d_e[2], h->d_b[0], h->d_b[1], h->d_b[2],
                    h->d_j[0], h->d_j[1], h->d_j[2], h->d_rho, h->d_part, h->d_stage, h->d_err,
                    h->d_count, h->d_cursor};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    if (h->m) gempic_maxwell2d_destroy(h->m);
    delete h;
    return GEMPIC_OK;
}

int gempic_splitting2d_set_particles(gempic_splitting2d *h, const double *aos, int64_t ld, int64_t n)
{
    CHECK_H2(h);
    if (!aos && n > 0) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL particle array");
    if (ld != 6) return abi_fail(GEMPIC_ERR_ASSERTION, "AssertionError: leading dimension must be 6 (D+V+n_weights)");
    if (n != h->n) return abi_fail(GEMPIC_ERR_ASSERTION, "AssertionError: number of particles differs");
    if (n == 0) return GEMPIC_OK;
    if (!h->d_stage) S2_TRY("2D3V staging", hipMalloc((void **)&h->d_stage, sizeof(double) * 6 * (size_t)n));
    S2_TRY("2D3V upload", hipMemcpyAsync(h->d_stage, aos, sizeof(double) * 6 * (size_t)n,
                                         hipMemcpyHostToDevice, h->stream));
    S2Perm pm{};
    for (int r = 0; r < 6; ++r) pm.dst[r] = h->d_p[r];
    pm.id_dst = h->d_id;
    const unsigned gb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256));
    hipLaunchKernelGGL(k_s2_in, dim3(gb), dim3(256), 0, h->stream, (const double *)h->d_stage, (int64_t)6, n, pm);
    S2_TRY("2D3V upload", hipGetLastError());
    h->steps_since_sort = 0;
    int rc = h->sort_interval > 0 ? s2_sort(h) : GEMPIC_OK;   // sorted on upload where sorting is on
    if (rc) return rc;
    return s2_finish(h);
}

int gempic_splitting2d_get_particles(gempic_splitting2d *h, double *aos, int64_t ld, int64_t n)
{
    CHECK_H2(h);
    if (!aos && n > 0) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL particle array");
    if (ld != 6) return abi_fail(GEMPIC_ERR_ASSERTION, "AssertionError: leading dimension must be 6 (D+V+n_weights)");
    if (n != h->n) return abi_fail(GEMPIC_ERR_ASSERTION, "AssertionError: number of particles differs");
    if (n == 0) return GEMPIC_OK;
    if (!h->d_stage) S2_TRY("2D3V staging", hipMalloc((void **)&h->d_stage, sizeof(double) * 6 * (size_t)n));
    S2Perm pm{};
    for (int r = 0; r < 6; ++r) pm.src[r] = h->d_p[r];
    pm.id = h->d_id;
    const unsigned gb = (unsigned)std::max<int64_t>(1, std::min<int64_t>(4096, (n + 255) / 256));
    hipLaunchKernelGGL(k_s2_out, dim3(gb), dim3(256), 0, h->stream, h->d_stage, (int64_t)6, n, pm);
    S2_TRY("2D3V download", hipGetLastError());
    S2_TRY("2D3V download", hipMemcpyAsync(aos, h->d_stage, sizeof(double) * 6 * (size_t)n,
                                           hipMemcpyDeviceToHost, h->stream));
    return s2_finish(h);
}

int gempic_splitting2d_set_fields(gempic_splitting2d *h, const double *e1, const double *e2,
                                  const double *e3, const double *b1, const double *b2,
                                  const double *b3)
{
    CHECK_H2(h);
    const size_t nb = sizeof(double) * h->nx * h->ny;
    const double *src[6] = {e1, e2, e3, b1, b2, b3};
    double *dst[6] = {h->d_e[0], h->d_e[1], h->d_e[2], h->d_b[0], h->d_b[1], h->d_b[2]};
    for (int k = 0; k < 6; ++k)
        if (src[k]) S2_TRY("2D3V fields", hipMemcpyAsync(dst[k], src[k], nb, hipMemcpyHostToDevice, h->stream));
    return s2_finish(h);
}

int gempic_splitting2d_get_fields(gempic_splitting2d *h, double *e1, double *e2, double *e3,
                                  double *b1, double *b2, double *b3)
{
    CHECK_H2(h);
    const size_t nb = sizeof(double) * h->nx * h->ny;
    double *dst[6] = {e1, e2, e3, b1, b2, b3};
    const double *src[6] = {h->d_e[0], h->d_e[1], h->d_e[2], h->d_b[0], h->d_b[1], h->d_b[2]};
    for (int k = 0; k < 6; ++k)
        if (dst[k]) S2_TRY("2D3V fields", hipMemcpyAsync(dst[k], src[k], nb, hipMemcpyDeviceToHost, h->stream));
    return s2_finish(h);
}

int gempic_splitting2d_get_j(gempic_splitting2d *h, double *j1, double *j2, double *j3)
{
    CHECK_H2(h);
    const size_t nb = sizeof(double) * h->nx * h->ny;
    double *dst[3] = {j1, j2, j3};
    for (int k = 0; k < 3; ++k)
        if (dst[k]) S2_TRY("2D3V currents", hipMemcpyAsync(dst[k], h->d_j[k], nb, hipMemcpyDeviceToHost, h->stream));
    return s2_finish(h);
}

int gempic_splitting2d_operator_hb(gempic_splitting2d *h, double dt)
{
    CHECK_H2(h);
    int rc = s2_hb(h, dt);
    return rc ? rc : s2_finish(h);
}

int gempic_splitting2d_operator_he(gempic_splitting2d *h, double dt)
{
    CHECK_H2(h);
    int rc = s2_he(h, dt);
    return rc ? rc : s2_finish(h);
}

int gempic_splitting2d_operator_hp1(gempic_splitting2d *h, double dt)
{
    CHECK_H2(h);
    int rc = s2_hp(h, 1, dt);
    return rc ? rc : s2_finish(h);
}

int gempic_splitting2d_operator_hp2(gempic_splitting2d *h, double dt)
{
    CHECK_H2(h);
    int rc = s2_hp(h, 2, dt);
    return rc ? rc : s2_finish(h);
}

int gempic_splitting2d_operator_hp3(gempic_splitting2d *h, double dt)
{
    CHECK_H2(h);
    int rc = s2_hp(h, 3, dt);
    return rc ? rc : s2_finish(h);
}

int gempic_splitting2d_strang_splitting(gempic_splitting2d *h, double dt, int64_t n_steps, int flags)
{
    CHECK_H2(h);
    if (n_steps < 0) return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: number_steps must be >= 0");
    const double hdt = 0.5 * dt;
    int rc = GEMPIC_OK;
    for (int64_t s = 0; s < n_steps && !rc; ++s) {
        if (h->sort_interval > 0 && h->steps_since_sort >= h->sort_interval) rc = s2_sort(h);
        if (!rc) rc = s2_hb(h, hdt);
        if (!rc) rc = s2_he(h, hdt);
        if (!rc) rc = s2_hp(h, 3, hdt);
        if (!rc) rc = s2_hp(h, 2, hdt);
        if (!rc) rc = s2_hp(h, 1, dt);
        if (!rc) rc = s2_hp(h, 2, hdt);
        if (!rc) rc = s2_hp(h, 3, hdt);
        if (!rc) rc = s2_he(h, hdt);
        if (!rc) rc = s2_hb(h, hdt);
        h->steps_since_sort++;
    }
    if (rc) return rc;
    if (flags & GEMPIC_ASYNC) return GEMPIC_OK;
    return s2_finish(h);
}

int gempic_splitting2d_synchronize(gempic_splitting2d *h)
{
    CHECK_H2(h);
    return s2_finish(h);
}

int gempic_splitting2d_set_sort_interval(gempic_splitting2d *h, int64_t interval)
{
    CHECK_H2(h);
    if (interval < 0) return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: sort interval must be >= 0");
    h->sort_interval = interval;
    return GEMPIC_OK;
}

int gempic_splitting2d_rho(gempic_splitting2d *h, double *rho)
{
    CHECK_H2(h);
    if (!rho) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL rho");
    int rc = s2_pass(h, S2_RHO, 0.0, nullptr, nullptr, nullptr, h->d_rho, 1.0);
    if (rc) return rc;
    S2_TRY("2D3V rho", hipMemcpyAsync(rho, h->d_rho, sizeof(double) * h->nx * h->ny,
                                      hipMemcpyDeviceToHost, h->stream));
    return s2_finish(h);
}

int gempic_splitting2d_rho_from_e(gempic_splitting2d *h, double *rho)
{
    CHECK_H2(h);
    if (!rho) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL rho");
    m2d_rho_from_e(h->m, h->d_e, h->d_rho);
    S2_TRY("2D3V rho", hipMemcpyAsync(rho, h->d_rho, sizeof(double) * h->nx * h->ny,
                                      hipMemcpyDeviceToHost, h->stream));
    return s2_finish(h);
}

}  // extern "C"
