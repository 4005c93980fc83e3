// pic2d.hip -- the 2D particle-mesh coupling and the 2D Maxwell solver on
// the GPU (FP64): batched ParticleMeshCoupling2D verbs and the TwoDMaxwell /
// TwoDPoisson / TwoDLinearSolverSplineMass operations, with their C ABI.
//
// Grid arrays are nx*ny doubles, x fastest (the reference's
// vec(reshape(..., nx, ny))); node (i, j) at i + j*nx.
//
// The transforms of the reference (FFTW) are direct DFT sums over twiddle
// tables computed in long double on the host, one output per thread, summed
// in index order; the field operations apply the reference's formulas with
// its rounding order. Deposits are deterministic: each lane caches the 2D
// patch of its current cell; evicted patches are staged in LDS and added to
// the wave's private grid in lane order, one node per lane; the waves' grids
// are summed over workgroups in a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gempic_hip.h"
#include "device_common.hpp"
#include "host_setup.hpp"
#include "lib_internal.hpp"

using namespace gl;
using namespace gk;

namespace {

constexpr int kGrid2Max = 4096;            // nx*ny of the device deposit (LDS grid)

// launch shape of the evaluation kernels: threads per workgroup (the waves
// of a workgroup share its staged field), workgroups per CU, and the largest
// field (doubles, both fields of evaluate_multiple together) staged in LDS
#ifndef GEMPIC_EVAL_THREADS
#define GEMPIC_EVAL_THREADS 1024
#endif
#ifndef GEMPIC_EVAL_BLOCKS_PER_CU
#define GEMPIC_EVAL_BLOCKS_PER_CU 4
#endif
#ifndef GEMPIC_EVAL_STAGE_MAX
#define GEMPIC_EVAL_STAGE_MAX 8192
#endif
// deposit: waves per workgroup (at most; the fast mode's waves share one
// LDS grid, the deterministic mode's own one each) and workgroups per CU (0:
// as many as fit)
#ifndef GEMPIC_DEP_WAVES
#define GEMPIC_DEP_WAVES 4
#endif
#ifndef GEMPIC_DEP_WAVES_FAST
#define GEMPIC_DEP_WAVES_FAST 4
#endif
#ifndef GEMPIC_DEP_BLOCKS_PER_CU
#define GEMPIC_DEP_BLOCKS_PER_CU 0
#endif

// ---------------------------------------------------------------------------
// device: ParticleMeshCoupling2D (particle_mesh_coupling_2d.jl)
// ---------------------------------------------------------------------------
struct Pmc2Args {
    int64_t n;
    int64_t chunk;              // particles per workgroup (multiple of 64)
    int nx, ny;
    double xmin, ymin, dx, dy, scaling;
    double rdx, rdy;            // 1 / dx, 1 / dy (correctly rounded): div_by
    const double *x, *y, *w;    // w: wp = weight * charge (add_charge! argument)
    const double *f1, *f2;
    double *out1, *out2;
    double *part;               // [workgroups][nx*ny]
    int *err;
    const double *pp;           // evaluate_pp: the ((P+1)^2) x (nx*ny) pp matrix
    double pc[16];              // SplinePP coefficient table ((P+1) x (P+1), column-major)
};

// compute_shape_factor (particle_mesh_coupling_2d.jl:54-65): ceil() index, offset (0, 1]; returns the
// first node (sx, sy) of the (P+1)^2 patch, wrapped into the grid
template <int P>
__device__ __forceinline__ bool shape2(const Pmc2Args &a, double x, double y, int &sx, int &sy,
                                       double (&vx)[P + 1], double (&vy)[P + 1])
{
    const double xp = div_by(x - a.xmin, a.dx, a.rdx);
    const double yp = div_by(y - a.ymin, a.dy, a.rdy);
    const bool ok = fabs(xp) < 1.0e9 && fabs(yp) < 1.0e9;
    const double cx = ok ? ceil(xp) : 0.0, cy = ok ? ceil(yp) : 0.0;
    const int ip = (int)cx, jp = (int)cy;
    bspline<P>(ok ? xp - (double)(ip - 1) : 1.0, vx);
    bspline<P>(ok ? yp - (double)(jp - 1) : 1.0, vy);
    // index_1dto2d_column_major (:76-82): nodes ip - P - 1 + a, mod n
    sx = imod(ip - P - 1, a.nx);
    sy = imod(jp - P - 1, a.ny);
    return ok;
}

// horner_1d (splinepp.jl:268-274) of column `col` of the (P+1) x (P+1) table
template <int P>
__device__ __forceinline__ double horner_col(const double *pc, double x, int col)
{
    double r = pc[col * (P + 1)];
#pragma unroll
    for (int i = 1; i <= P; ++i) r = r * x + pc[i + col * (P + 1)];
    return r;
}

// add_charge_pp! 2D (particle_mesh_coupling_2d.jl:127-148): floor() + 1 index
// and the horner_m_2d! weights (splinepp.jl:295-300)
template <int P>
__device__ __forceinline__ bool shape2_pp(const Pmc2Args &a, double x, double y, int &sx, int &sy,
                                          double (&vx)[P + 1], double (&vy)[P + 1])
{
    const double xp = div_by(x - a.xmin, a.dx, a.rdx);
    const double yp = div_by(y - a.ymin, a.dy, a.rdy);
    const bool ok = fabs(xp) < 1.0e9 && fabs(yp) < 1.0e9;
    const int ip = ok ? (int)floor(xp) + 1 : 1, jp = ok ? (int)floor(yp) + 1 : 1;
    const double dxp = ok ? xp - (double)(ip - 1) : 0.0;
    const double dyp = ok ? yp - (double)(jp - 1) : 0.0;
#pragma unroll
    for (int i = 0; i <= P; ++i) {
        vx[i] = horner_col<P>(a.pc, dxp, i);
        vy[i] = horner_col<P>(a.pc, dyp, i);
    }
    sx = imod(ip - P - 1, a.nx);
    sy = imod(jp - P - 1, a.ny);
    return ok;
}

// Deposit (add_charge!, :94-105). Each lane keeps the (P+1)^2 node sums of
// the patch its particles fall in, in registers; when its next particle
// starts another patch, the cached sums are flushed into an LDS grid:
//  * if every flushing lane of the wave holds the same patch (particles
//    sorted by cell), the wave sums them (DPP, fixed order) and 16 lanes add
//    the patch to the grid;
//  * otherwise, in the fast mode (DET = false), each flushing lane adds its
//    patch with LDS atomics into the workgroup's shared grid;
//  * in the deterministic mode (DET = true) every wave owns its grid and the
//    flushing lanes add in passes: per pass, the lowest lane of each hash
//    slot of its patch adds its whole patch (distinct patches, so no two
//    lanes of one instruction hit one node), in lane order of the slots.
// In the deterministic mode every addition's order is a fixed function of
// the particle order; the waves' grids are summed in wave order at the end.
template <int P, bool DET, bool PP>
__global__ void __launch_bounds__(1024) k_pmc2d_deposit(Pmc2Args a)
{
    constexpr int W = P + 1, WW = W * W;
    static_assert(WW <= 16, "one node per lane of a 16-lane row");
    extern __shared__ double smem[];
    const int nn = a.nx * a.ny;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int nw = blockDim.x / kWave;
    const int ngrids = DET ? nw : 1;
    double *grid = smem + (DET ? (size_t)wave * nn : 0);
    double *stage = smem + (size_t)ngrids * nn + wave * 16;
    int *hash = (int *)(smem + (size_t)ngrids * nn + nw * 16) + wave * 256;
    for (int k = threadIdx.x; k < ngrids * nn; k += blockDim.x) smem[k] = 0.0;
    __syncthreads();

    // node j (= ia * W + ib) of the patch starting at node `key`
    auto node = [&](int key, int j) {
        const int sy = key / a.nx, sx = key - sy * a.nx;
        const int ia = j / W, ib = j - ia * W;
        int gx = sx + ia, gy = sy + ib;
        if (gx >= a.nx) gx -= a.nx;
        if (gy >= a.ny) gy -= a.ny;
        return gx + gy * a.nx;
    };

    int ckey = -1;
    double acc[WW];
#pragma unroll
    for (int j = 0; j < WW; ++j) acc[j] = 0.0;
    int errs = 0;

    auto flush = [&](bool evict) {
        const unsigned long long m = __ballot(evict);
        if (!m) return;
        const int k0 = __shfl(ckey, __ffsll((long long)m) - 1, kWave);
        if (__ballot(evict && ckey != k0) == 0) {
#pragma unroll
            for (int j = 0; j < WW; ++j) {
                const double sj = wave_sum_lane63(evict ? acc[j] : 0.0);
                if (lane == kWave - 1) stage[j] = sj;
            }
            if (lane < WW) atomicAdd(&grid[node(k0, lane)], stage[lane]);
        } else if constexpr (!DET) {
            if (evict) {
#pragma unroll
                for (int j = 0; j < WW; ++j) atomicAdd(&grid[node(ckey, j)], acc[j]);
            }
        } else {
            bool rem = evict;
            while (__ballot(rem)) {
                const int slot = ckey & 255;
                if (rem) hash[slot] = 0x7fffffff;
                if (rem) atomicMin(&hash[slot], lane);
                const bool lead = rem && hash[slot] == lane;
                if (lead) {
#pragma unroll
                    for (int j = 0; j < WW; ++j) atomicAdd(&grid[node(ckey, j)], acc[j]);
                }
                rem = rem && !lead;
            }
        }
    };

    const int64_t wbase = ((int64_t)blockIdx.x * nw + wave) * a.chunk;
    // the next particle's inputs are loaded while the current one is deposited
    double xn = a.xmin, yn = a.ymin, wn = 0.0;
    if (wbase + lane < a.n) {
        xn = a.x[wbase + lane];
        yn = a.y[wbase + lane];
        wn = a.w[wbase + lane];
    }
    for (int64_t off = 0; off < a.chunk; off += kWave) {
        const int64_t i = wbase + off + lane;
        const bool valid = i < a.n;
        const double x = xn, y = yn, wp = wn;
        if (off + kWave < a.chunk && i + kWave < a.n) {
            xn = a.x[i + kWave];
            yn = a.y[i + kWave];
            wn = a.w[i + kWave];
        }
        double vx[W], vy[W];
        int sx = 0, sy = 0;
        const bool ok = PP ? shape2_pp<P>(a, x, y, sx, sy, vx, vy) : shape2<P>(a, x, y, sx, sy, vx, vy);
        if (valid && !ok) errs |= E_INEXACT;
        const int key = sx + sy * a.nx;
        const bool miss = valid && key != ckey;
        flush(miss && ckey >= 0);
        if (miss) {
            ckey = key;
#pragma unroll
            for (int j = 0; j < WW; ++j) acc[j] = 0.0;
        }
        if (valid) {
            // wp * scaling * Sx[a] * Sy[b]
            const double ws = wp * a.scaling;
#pragma unroll
            for (int ia = 0; ia < W; ++ia) {
                const double wx = ws * vx[ia];
#pragma unroll
                for (int ib = 0; ib < W; ++ib) acc[ia * W + ib] += wx * vy[ib];
            }
        }
    }
    flush(ckey >= 0);
    if (errs) atomicOr(a.err, errs);
    __syncthreads();
    double *out = a.part + (size_t)blockIdx.x * nn;
    for (int k = threadIdx.x; k < nn; k += blockDim.x) {
        double t = smem[k];
        for (int w = 1; w < ngrids; ++w) t += smem[(size_t)w * nn + k];
        out[k] = t;
    }
}

// sum of col[b * nn] for b in [b0, b1), added in order; the loads are issued
// eight at a time
__device__ __forceinline__ double sum_column(const double *col, int b0, int b1, int nn)
{
    double t = 0.0;
    int b = b0;
    for (; b + 8 <= b1; b += 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = col[(size_t)(b + u) * nn];
#pragma unroll
        for (int u = 0; u < 8; ++u) t += v[u];
    }
    for (; b < b1; ++b) t += col[(size_t)b * nn];
    return t;
}

// part2[s][k] = sum of part[b][k] over the workgroups b of split s, in order
__global__ void k_pmc2d_reduce1(const double *part, int nblocks, int nn, double *part2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    const int splits = gridDim.y, sp = blockIdx.y;
    const int b0 = (int)((int64_t)sp * nblocks / splits), b1 = (int)((int64_t)(sp + 1) * nblocks / splits);
    part2[(size_t)sp * nn + k] = sum_column(part + k, b0, b1, nn);
}

// rho[k] += sum of part2[s][k] over the splits, in order
__global__ void k_pmc2d_reduce2(const double *part2, int splits, int nn, double *rho)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nn) return;
    rho[k] = rho[k] + sum_column(part2 + k, 0, splits, nn);
}

// evaluate (:189-203) and evaluate_multiple (:214-231). With STAGE the
// field is copied into LDS first: evaluate's with a periodic halo of P
// columns and rows ((nx+P) x (ny+P), so the 16 reads of a particle are
// constant offsets from its first node), evaluate_multiple's two fields
// interleaved (one 16-byte read per node). The next particle's position is
// loaded while the current one is evaluated.
template <int P, bool MULTI, bool STAGE>
__global__ void __launch_bounds__(1024) k_pmc2d_eval(Pmc2Args a)
{
    constexpr int W = P + 1;
    extern __shared__ __align__(16) double sfield[];
    const int sw = STAGE && !MULTI ? a.nx + P : a.nx;       // row stride of the staged field
    if constexpr (STAGE) {
        if constexpr (MULTI) {
            const int nn = a.nx * a.ny;
            for (int k = threadIdx.x; k < nn; k += blockDim.x) {
                sfield[2 * k] = a.f1[k];
                sfield[2 * k + 1] = a.f2[k];
            }
        } else {
            const int ns = sw * (a.ny + P);
            for (int k = threadIdx.x; k < ns; k += blockDim.x) {
                const int gy = k / sw, gx = k - gy * sw;
                sfield[k] = a.f1[(gx < a.nx ? gx : gx - a.nx) + (gy < a.ny ? gy : gy - a.ny) * a.nx];
            }
        }
        __syncthreads();
    }
    const double2 *f12 = (const double2 *)sfield;
    int errs = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double xn = i < a.n ? a.x[i] : 0.0, yn = i < a.n ? a.y[i] : 0.0;
    for (; i < a.n; i += stride) {
        const double x = xn, y = yn;
        if (i + stride < a.n) {
            xn = a.x[i + stride];
            yn = a.y[i + stride];
        }
        double vx[W], vy[W];
        int sx, sy;
        if (!shape2<P>(a, x, y, sx, sy, vx, vy)) errs |= E_INEXACT;
        double v1 = 0.0, v2 = 0.0;
        if constexpr (STAGE && !MULTI) {
            const double *f = sfield + sx + sy * sw;
#pragma unroll
            for (int ia = 0; ia < W; ++ia)
#pragma unroll
                for (int ib = 0; ib < W; ++ib) v1 += f[ia + ib * sw] * vx[ia] * vy[ib];
        } else {
            int gx = sx;
#pragma unroll
            for (int ia = 0; ia < W; ++ia) {
                int gy = sy;
#pragma unroll
                for (int ib = 0; ib < W; ++ib) {
                    const int k = gx + gy * a.nx;
                    if constexpr (MULTI) {
                        const double c = vx[ia] * vy[ib];
                        double g1, g2;
                        if constexpr (STAGE) {
                            const double2 g = f12[k];
                            g1 = g.x;
                            g2 = g.y;
                        } else {
                            g1 = a.f1[k];
                            g2 = a.f2[k];
                        }
                        v1 += g1 * c;
                        v2 += g2 * c;
                    } else {
                        v1 += a.f1[k] * vx[ia] * vy[ib];
                    }
                    gy = gy + 1 == a.ny ? 0 : gy + 1;
                }
                gx = gx + 1 == a.nx ? 0 : gx + 1;
            }
        }
        a.out1[i] = v1;
        if constexpr (MULTI) a.out2[i] = v2;
    }
    if (errs) atomicOr(a.err, errs);
}

// evaluate_pp 2D (particle_mesh_coupling_2d.jl:162-178) with horner_2d
// (splinepp.jl:412-434): floor() + 1 indices, not wrapped; a position outside
// the pp table is a BoundsError (the value is NaN)
template <int P>
__global__ void __launch_bounds__(256) k_pmc2d_eval_pp(Pmc2Args a)
{
    constexpr int W = P + 1, LD = W * W;
    int errs = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n;
         i += (int64_t)gridDim.x * blockDim.x) {
        double xi = div_by(a.x[i] - a.xmin, a.dx, a.rdx);
        double yi = div_by(a.y[i] - a.ymin, a.dy, a.rdy);
        const bool fin = fabs(xi) < 1.0e9 && fabs(yi) < 1.0e9;
        const int idx = fin ? (int)floor(xi) + 1 : 0, idy = fin ? (int)floor(yi) + 1 : 0;
        xi = xi - (double)(idx - 1);
        yi = yi - (double)(idy - 1);
        double res = __longlong_as_double(0x7ff8000000000000ll);
        if (idx >= 1 && idx <= a.nx && idy >= 1 && idy <= a.ny) {
            const double *col = a.pp + ((size_t)(idx - 1) + (size_t)(idy - 1) * a.nx) * LD;
            double p1[W];
#pragma unroll
            for (int r = 0; r < W; ++r) {
                const double *c = col + r * W;
                double t = c[0];
#pragma unroll
                for (int k = 1; k < W; ++k) t = t * xi + c[k];
                p1[r] = t;
            }
            res = p1[0];
#pragma unroll
            for (int r = 1; r < W; ++r) res = res * yi + p1[r];
        } else {
            errs |= E_BOUNDS;
        }
        a.out1[i] = res;
    }
    if (errs) atomicOr(a.err, errs);
}

// b_to_pp_2d! (splinepp.jl:329-399): the pp column of cell (i, j) from the
// spline coefficients b[(i - W + k) mod nx, (j - W + l) mod ny], first along x
// for each l (b_to_pp_1d_cell!, :310-316), then along y for each k
template <int P>
__global__ void __launch_bounds__(256) k_b_to_pp_2d(const double *b, int nx, int ny, Pmc2Args a,
                                                    double *pp)
{
    constexpr int W = P + 1, LD = W * W;
    const int cell = blockIdx.x * blockDim.x + threadIdx.x;
    if (cell >= nx * ny) return;
    const int j = cell / nx, i = cell - j * nx;
    double col[LD];
#pragma unroll
    for (int l = 0; l < W; ++l) {
        int gy = j - W + 1 + l;
        gy = gy < 0 ? gy + ny : gy;
        double bx[W];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            int gx = i - W + 1 + k;
            gx = gx < 0 ? gx + nx : gx;
            bx[k] = b[gx + (size_t)gy * nx];
        }
#pragma unroll
        for (int q = 0; q < W; ++q) col[l * W + q] = 0.0;
#pragma unroll
        for (int r = 0; r < W; ++r)
#pragma unroll
            for (int q = 0; q < W; ++q) col[l * W + q] += bx[r] * a.pc[q + r * W];
    }
#pragma unroll
    for (int k = 0; k < W; ++k) {
        double by[W], p2[W];
#pragma unroll
        for (int l = 0; l < W; ++l) by[l] = col[k + W * l];
#pragma unroll
        for (int q = 0; q < W; ++q) p2[q] = 0.0;
#pragma unroll
        for (int r = 0; r < W; ++r)
#pragma unroll
            for (int q = 0; q < W; ++q) p2[q] += by[r] * a.pc[q + r * W];
#pragma unroll
        for (int l = 0; l < W; ++l) col[k + W * l] = p2[l];
    }
    double *out = pp + (size_t)cell * LD;
#pragma unroll
    for (int q = 0; q < LD; ++q) out[q] = col[q];
}

// ---------------------------------------------------------------------------
// device: TwoDMaxwell, TwoDPoisson, TwoDLinearSolverSplineMass
// ---------------------------------------------------------------------------

// one complex DFT per row (x direction, DIR_X) or per column (y direction):
// X[k] = sum_j x[j] e^{-+2 pi i jk/n}, backward divided by n; in place
template <bool DIR_X>
__global__ void __launch_bounds__(256) k_dft(double *re, double *im, int nx, int ny,
                                             const double *cs, const double *sn, int backward)
{
    __shared__ double sr[kNxMax], si[kNxMax];
    const int n = DIR_X ? nx : ny;
    const int line = blockIdx.x;
    const int64_t b0 = DIR_X ? (int64_t)line * nx : line;
    const int64_t st = DIR_X ? 1 : nx;
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
        sr[j] = re[b0 + j * st];
        si[j] = im[b0 + j * st];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double ar = 0.0, ai = 0.0;
        int q = 0;
        for (int j = 0; j < n; ++j) {
            const double xr = sr[j], xi = si[j], c = cs[q], s = sn[q];
            if (!backward) {
                ar += xr * c + xi * s;
                ai += xi * c - xr * s;
            } else {
                ar += xr * c - xi * s;
                ai += xi * c + xr * s;
            }
            q += k;
            if (q >= n) q -= n;
        }
        if (backward) {
            ar = ar / (double)n;
            ai = ai / (double)n;
        }
        re[b0 + k * st] = ar;
        im[b0 + k * st] = ai;
    }
}

struct Field2 {
    int nx, ny;
    double dx, dy;
};

#define FOR_NODES(f)                                                        \
    const int64_t nn_ = (int64_t)(f).nx * (f).ny;                           \
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nn_; \
         k += (int64_t)gridDim.x * blockDim.x)

// re = in, im = 0
__global__ void k_load_real(Field2 f, const double *in, double *re, double *im)
{
    FOR_NODES(f) {
        re[k] = in[k];
        im[k] = 0.0;
    }
}

// linear_solver_spline_mass_2d.jl:23-27: wk[i, j] /= (eig1[i] * eig2[j] * factor)
__global__ void k_kron_div(Field2 f, double *re, double *im, const double *e1, const double *e2)
{
    FOR_NODES(f) {
        const int i = (int)(k % f.nx), j = (int)(k / f.nx);
        const double d = e1[i] * e2[j] * 1.0;
        re[k] = re[k] / d;
        im[k] = im[k] / d;
    }
}

// spline_fem_multiply_mass (maxwell_2d_fem.jl:300-336) along x (DIR_X) or y
template <bool DIR_X>
__global__ void k_mass(Field2 f, const double *in, double *out, const double *line, int deg)
{
    FOR_NODES(f) {
        const int i = (int)(k % f.nx), j = (int)(k / f.nx);
        const int n = DIR_X ? f.nx : f.ny;
        const int r = DIR_X ? i : j;
        const int64_t b0 = DIR_X ? (int64_t)j * f.nx : i;
        const int64_t st = DIR_X ? 1 : f.nx;
        double o = line[0] * in[b0 + r * st];
        for (int c = 1; c <= deg; ++c) {
            int rp = r + c, rm = r - c;
            if (rp >= n) rp -= n;
            if (rm < 0) rm += n;
            o += line[c] * (in[b0 + rp * st] + in[b0 + rm * st]);
        }
        out[k] = o;
    }
}

// curl of the weak Ampere step (compute_e_from_b!, maxwell_2d_fem.jl:389-402)
__global__ void k_curl_b(Field2 f, const double *w1, const double *w2, const double *w3,
                         double *c1, double *c2, double *c3)
{
    FOR_NODES(f) {
        const int i = (int)(k % f.nx), j = (int)(k / f.nx);
        const int64_t k1 = (i + 1 == f.nx ? 0 : i + 1) + (int64_t)j * f.nx;
        const int64_t k2 = i + (int64_t)(j + 1 == f.ny ? 0 : j + 1) * f.nx;
        c1[k] = -(w3[k] - w3[k2]) / f.dy;
        c2[k] = (w3[k] - w3[k1]) / f.dx;
        c3[k] = (w1[k] - w1[k2]) / f.dy - (w2[k] - w2[k1]) / f.dx;
    }
}

// e = e + dt * w  (:409-411); e = e - w (:458)
__global__ void k_axpy(Field2 f, double *e, const double *w, double dt)
{
    FOR_NODES(f) e[k] = e[k] + dt * w[k];
}

__global__ void k_sub(Field2 f, double *e, const double *w)
{
    FOR_NODES(f) e[k] = e[k] - w[k];
}

// compute_b_from_e! 2D (maxwell_2d_fem.jl:423-444)
__global__ void k_b_from_e(Field2 f, double dt, const double *e1, const double *e2,
                           const double *e3, double *b1, double *b2, double *b3)
{
    FOR_NODES(f) {
        const int i = (int)(k % f.nx), j = (int)(k / f.nx);
        const int64_t k1 = (i == 0 ? f.nx - 1 : i - 1) + (int64_t)j * f.nx;
        const int64_t k2 = i + (int64_t)(j == 0 ? f.ny - 1 : j - 1) * f.nx;
        b1[k] += -dt * (e3[k] - e3[k2]) / f.dy;
        b2[k] += dt * (e3[k] - e3[k1]) / f.dx;
        b3[k] += -dt * ((e2[k] - e2[k1]) / f.dx - (e1[k] - e1[k2]) / f.dy);
    }
}

// compute_rho_from_e! divergence (maxwell_2d_fem.jl:484-498)
__global__ void k_div(Field2 f, const double *w1, const double *w2, double *rho)
{
    FOR_NODES(f) {
        const int i = (int)(k % f.nx), j = (int)(k / f.nx);
        const int64_t k1 = (i + 1 == f.nx ? 0 : i + 1) + (int64_t)j * f.nx;
        const int64_t k2 = i + (int64_t)(j + 1 == f.ny ? 0 : j + 1) * f.nx;
        const double r = (w1[k] - w1[k1]) / f.dx + (w2[k] - w2[k2]) / f.dy;
        rho[k] = r * -1.0;
    }
}

// ---- field updates of the 2D3V splitting in two kernels each: the first
// takes blocks of kFieldRows rows (the mass products, the curl and the
// current sum row by row, then the inverse mass along x), the second blocks
// of kFieldCols columns (the inverse mass along y and the update of E). The
// inverse mass of the DFT solve is a circulant matrix; its rows are the
// kernels of host_inv_mass_kernel, applied as circular convolutions.

constexpr int kFieldRows = 4, kFieldCols = 4, kFieldThreads = 256;

struct FieldTabs {
    const double *l0x, *l1x, *l0y, *l1y;   // mass lines of degree p and p - 1 (times dx, dy)
    const double *k0x, *k1x, *k0y, *k1y;   // inverse-mass convolution kernels
    int degree;
};

// out[rr][i] = sum_m k[m] in[rr][(i - m) mod nx] for `rows` rows of nx in LDS
__device__ void conv_rows(const double *in, const double *k, int rows, int nx, double *out, int64_t ostride)
{
    for (int t = threadIdx.x; t < rows * nx; t += blockDim.x) {
        const int rr = t / nx, i = t - rr * nx;
        const double *line = in + rr * nx;
        double acc = 0.0;
        int u = i;
        for (int m = 0; m < nx; ++m) {
            acc = __builtin_fma(k[m], line[u], acc);
            u = u == 0 ? nx - 1 : u - 1;
        }
        out[(int64_t)rr * ostride + i] = acc;
    }
}

// compute_e_from_j! (maxwell_2d_fem.jl:455-459), first kernel: the rows
// r0 .. r0 + rows - 1 of the current, j = (sum of the splits' partial grids
// in order) * scale as k_s2_reduce2 forms it (written to jout), then their
// inverse mass along x into tmp
__global__ void __launch_bounds__(kFieldThreads) k_fj_rows(Field2 f, const double *part2, int splits,
                                                           double scale, const double *kx, double *jout,
                                                           double *tmp)
{
    extern __shared__ __align__(16) double lds[];
    const int nx = f.nx, nn = nx * f.ny;
    const int r0 = blockIdx.x * kFieldRows, rows = min(kFieldRows, f.ny - r0);
    double *jr = lds, *kk = lds + kFieldRows * nx;
    for (int m = threadIdx.x; m < nx; m += blockDim.x) kk[m] = kx[m];
    for (int t = threadIdx.x; t < rows * nx; t += blockDim.x) {
        const int k = r0 * nx + t;
        double v = 0.0;
#pragma unroll 8
        for (int sp = 0; sp < splits; ++sp) v += part2[(size_t)sp * nn + k];
        v = v * scale;
        jout[k] = v;
        jr[t] = v;
    }
    __syncthreads();
    conv_rows(jr, kk, rows, nx, tmp + (size_t)r0 * nx, nx);
}

// second kernel of both updates: the columns c0 .. c0 + cols - 1 of tmp
// (component blockIdx.y), their inverse mass along y, e = e + coef w
// (coef = dt: k_axpy; coef = -1: k_sub, bitwise e - w)
struct FieldCols {
    const double *tmp[3];
    const double *ky[3];
    double *e[3];
};

__global__ void __launch_bounds__(kFieldThreads) k_f_cols(Field2 f, FieldCols a, double coef)
{
    extern __shared__ __align__(16) double lds[];
    const int nx = f.nx, ny = f.ny, c = blockIdx.y;
    const int c0 = blockIdx.x * kFieldCols, cols = min(kFieldCols, nx - c0);
    double *col = lds, *kk = lds + kFieldCols * ny;
    const double *tmp = a.tmp[c], *ky = a.ky[c];
    double *e = a.e[c];
    for (int m = threadIdx.x; m < ny; m += blockDim.x) kk[m] = ky[m];
    for (int t = threadIdx.x; t < cols * ny; t += blockDim.x) {
        const int j = t / cols, cc = t - j * cols;
        col[cc * ny + j] = tmp[(size_t)j * nx + c0 + cc];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < cols * ny; t += blockDim.x) {
        const int cc = t / ny, j = t - cc * ny;
        const double *line = col + cc * ny;
        double acc = 0.0;
        int u = j;
        for (int m = 0; m < ny; ++m) {
            acc = __builtin_fma(kk[m], line[u], acc);
            u = u == 0 ? ny - 1 : u - 1;
        }
        const size_t k = (size_t)j * nx + c0 + cc;
        e[k] = e[k] + coef * acc;
    }
}

// rows of a mass product M(x: lx, dgx; y: ly, dgy) b for the output rows
// r0 .. r0 + nout - 1 (mod ny), with k_mass's arithmetic: the rows of b
// from r0 - p (H = nout + 2p of them) staged in sb, their products along x
// in mx, then along y into u
__device__ void mass_rows(const double *b, const double *lx, int dgx, const double *ly, int dgy, int p,
                          int r0, int nout, int nx, int ny, double *sb, double *mx, double *u)
{
    const int H = nout + 2 * p;
    for (int t = threadIdx.x; t < H * nx; t += blockDim.x) {
        const int hr = t / nx, i = t - hr * nx;
        int j = r0 - p + hr;
        j = j < 0 ? j + ny : (j >= ny ? j - ny : j);
        sb[t] = b[(size_t)j * nx + i];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < H * nx; t += blockDim.x) {
        const int hr = t / nx, i = t - hr * nx;
        const double *in = sb + hr * nx;
        double o = lx[0] * in[i];
        for (int c = 1; c <= dgx; ++c) {
            int ip = i + c, im = i - c;
            if (ip >= nx) ip -= nx;
            if (im < 0) im += nx;
            o += lx[c] * (in[ip] + in[im]);
        }
        mx[t] = o;
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nout * nx; t += blockDim.x) {
        const int rr = t / nx, i = t - rr * nx;
        const double *in = mx + (rr + p) * nx + i;
        double o = ly[0] * in[0];
        for (int c = 1; c <= dgy; ++c) o += ly[c] * (in[c * nx] + in[-c * nx]);
        u[t] = o;
    }
    __syncthreads();
}

// compute_e_from_b! (maxwell_2d_fem.jl:370-411), first kernel: component
// blockIdx.y of curl(M2 b) on the rows r0 .. r0 + rows - 1 (the mass
// products and the curl of k_mass and k_curl_b), then its inverse mass
// along x into tmp[c]
struct FieldRowsB {
    const double *b[3];
    double *tmp[3];
};

__global__ void __launch_bounds__(kFieldThreads) k_fb_rows(Field2 f, FieldTabs t, FieldRowsB a)
{
    extern __shared__ __align__(16) double lds[];
    const int nx = f.nx, ny = f.ny, p = t.degree, c = blockIdx.y;
    const int r0 = blockIdx.x * kFieldRows, rows = min(kFieldRows, ny - r0);
    const int H = kFieldRows + 1 + 2 * p;
    double *sb = lds, *mx = sb + H * nx, *u1 = mx + H * nx, *u2 = u1 + (kFieldRows + 1) * nx;
    double *cu = u2 + (kFieldRows + 1) * nx, *kk = cu + kFieldRows * nx;
    const double *kx = c == 0 ? t.k1x : t.k0x;
    for (int m = threadIdx.x; m < nx; m += blockDim.x) kk[m] = kx[m];
    // u at rows r0 .. r0 + rows (the curl's j + 1 included)
    if (c == 2) {
        mass_rows(a.b[0], t.l0x, p, t.l1y, p - 1, p, r0, rows + 1, nx, ny, sb, mx, u1);
        mass_rows(a.b[1], t.l1x, p - 1, t.l0y, p, p, r0, rows + 1, nx, ny, sb, mx, u2);
    } else {
        mass_rows(a.b[2], t.l1x, p - 1, t.l1y, p - 1, p, r0, rows + 1, nx, ny, sb, mx, u1);
    }
    for (int q = threadIdx.x; q < rows * nx; q += blockDim.x) {
        const int rr = q / nx, i = q - rr * nx;
        const int i1 = i + 1 == nx ? 0 : i + 1;
        const double *w = u1 + rr * nx;
        if (c == 0) cu[q] = -(w[i] - w[i + nx]) / f.dy;
        else if (c == 1) cu[q] = (w[i] - w[i1]) / f.dx;
        else cu[q] = (w[i] - w[i + nx]) / f.dy - (u2[rr * nx + i] - u2[rr * nx + i1]) / f.dx;
    }
    __syncthreads();
    conv_rows(cu, kk, rows, nx, a.tmp[c] + (size_t)r0 * nx, nx);
}

struct PoissonTabs {
    const double *eig0x, *eig0y, *dtm1x, *dtm1y, *d1re, *d1im, *d2re, *d2im;
};

// TwoDPoisson mode update (poisson_2d_fem.jl:243-256)
__global__ void k_poisson_modes(Field2 f, PoissonTabs t, const double *sr, const double *si,
                                double *xr, double *xi, double *yr, double *yi)
{
    FOR_NODES(f) {
        const int i = (int)(k % f.nx), j = (int)(k / f.nx);
        double r, m;
        if (i == 0 && j == 0) {
            r = 0.0;
            m = 0.0;
        } else {
            const double ev = t.dtm1x[i] * t.eig0y[j] + t.eig0x[i] * t.dtm1y[j];
            r = sr[k] / ev;
            m = si[k] / ev;
        }
        const double ar = -r, ai = -m;
        xr[k] = ar * t.d1re[i] - ai * t.d1im[i];
        xi[k] = ar * t.d1im[i] + ai * t.d1re[i];
        yr[k] = ar * t.d2re[j] - ai * t.d2im[j];
        yi[k] = ar * t.d2im[j] + ai * t.d2re[j];
    }
}

// sum(c1 .* w) in a fixed order: contiguous chunks per thread, then the
// chunk sums in thread order (one workgroup)
__global__ void __launch_bounds__(1024) k_dot(Field2 f, const double *c1, const double *w,
                                              double *out)
{
    __shared__ double part[1024];
    const int64_t nn = (int64_t)f.nx * f.ny;
    const int nt = blockDim.x, t = threadIdx.x;
    const int64_t per = (nn + nt - 1) / nt;
    const int64_t a0 = std::min<int64_t>(nn, per * t), a1 = std::min<int64_t>(nn, a0 + per);
    double s = 0.0;
    for (int64_t k = a0; k < a1; ++k) s += c1[k] * w[k];
    part[t] = s;
    __syncthreads();
    if (t == 0) {
        double r = 0.0;
        for (int k = 0; k < nt; ++k) r += part[k];
        out[0] = r;
    }
}

// ---------------------------------------------------------------------------
// host set-up (poisson_2d_fem.jl:23-126, maxwell_2d_fem.jl:26-87)
// ---------------------------------------------------------------------------
void host_bspline(int degree, double off, double *b)
{
    b[0] = 1.0;
    for (int j = 1; j <= degree; ++j) {
        double xx = -off;
        const double jr = (double)j;
        const double inv_j = 1.0 / jr;
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            xx = xx + 1.0;
            const double temp = b[r] * inv_j;
            b[r] = saved + xx * temp;
            saved = (jr - xx) * temp;
        }
        b[j] = saved;
    }
}

// spline_fem_mass_line (poisson_2d_fem.jl:87-106)
void host_mass_line(int degree, double *line)
{
    const int n = degree + 1;
    double x[8], w[8], val[8 * 8];
    gh::gauss_legendre(n, x, w);
    for (int k = 0; k < n; ++k) {
        x[k] += 1.0;
        x[k] *= 0.5;
        w[k] *= 0.5;
    }
    for (int j = 0; j < n; ++j) host_bspline(degree, x[j], val + j * (degree + 1));
    for (int j = 0; j <= degree; ++j) line[j] = 0.0;
    for (int j = 1; j <= degree + 1; ++j)
        for (int i = j; i <= degree + 1; ++i)
            for (int k = 1; k <= n; ++k)
                line[j - 1] += val[(i - 1) + (k - 1) * (degree + 1)] *
                               val[(i - j) + (k - 1) * (degree + 1)] * w[k - 1];
}

// spline_fem_compute_mass_eig (poisson_2d_fem.jl:113-126)
void host_mass_eig(int64_t n, int degree, const double *line, double *eig)
{
    const double factor = 6.283185307179586 / (double)n;
    for (int64_t k = 1; k <= n; ++k) {
        double e = line[0];
        for (int j = 1; j <= degree; ++j)
            e += line[j] * 2.0 * std::cos(factor * (double)(k - 1) * (double)j);
        eig[k - 1] = e;
    }
}

// The inverse of a circulant mass matrix with eigenvalues eig[k] (its DFT
// symbol, real and even: eig[k] = eig[n - k]) is the circulant matrix of
// kinv[m] = (1/n) sum_k cos(2 pi k m / n) / eig[k], in long double.
void host_inv_mass_kernel(int64_t n, const double *eig, double *kinv)
{
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int64_t m = 0; m < n; ++m) {
        long double t = 0.0L;
        for (int64_t k = 0; k < n; ++k)
            t += cosl(twopi * (long double)((k * m) % n) / (long double)n) / (long double)eig[k];
        kinv[m] = (double)(t / (long double)n);
    }
}

void host_twiddles(int64_t n, double *c, double *s)
{
    for (int64_t i = 0; i < n; ++i) {
        long double ang = 2.0L * 3.14159265358979323846264338327950288L * (long double)i / (long double)n;
        c[i] = (double)cosl(ang);
        s[i] = (double)sinl(ang);
    }
}

template <typename T>
hipError_t dalloc2(T **p, size_t count)
{
    return hipMalloc((void **)p, sizeof(T) * std::max<size_t>(count, 1));
}

unsigned grid_blocks(int64_t nn) { return (unsigned)std::min<int64_t>(1024, (nn + 255) / 256); }

}  // namespace

// table slots of a TwoDMaxwell handle (each of length kNxMax)
enum Tab2 {
    T2_CX = 0, T2_SX, T2_CY, T2_SY, T2_E0X, T2_E1X, T2_E0Y, T2_E1Y,
    T2_DTMX, T2_DTMY, T2_D1R, T2_D1I, T2_D2R, T2_D2I, T2_L0X, T2_L1X, T2_L0Y, T2_L1Y,
    T2_K0X, T2_K1X, T2_K0Y, T2_K1Y, NT2
};

struct gempic_maxwell2d {
    hipStream_t stream = nullptr;
    int nx = 0, ny = 0, degree = 3;
    double dx = 0, dy = 0;
    double *d_tab = nullptr;            // NT2 x kNxMax
    double *d_buf = nullptr;            // 16 work arrays of nx*ny
    double *d_scalar = nullptr;
    bool fused = false;                 // the row and column field kernels take this grid (k_fb_rows ..)
    const double *tab(int s) const { return d_tab + (size_t)s * kNxMax; }
    double *buf(int i) const { return d_buf + (size_t)i * nx * ny; }
    Field2 f() const { return Field2{nx, ny, dx, dy}; }
};

namespace {

// out = TwoDLinearSolverSplineMass(e1, e2).solve(in) (buffers on the device)
void kron_solve(gempic_maxwell2d *m, const double *e1, const double *e2, const double *in,
                double *out)
{
    const Field2 f = m->f();
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    double *re = m->buf(14), *im = m->buf(15);
    hipLaunchKernelGGL(k_load_real, dim3(gb), dim3(256), 0, m->stream, f, in, re, im);
    hipLaunchKernelGGL(k_dft<true>, dim3(m->ny), dim3(256), 0, m->stream, re, im, m->nx, m->ny,
                       m->tab(T2_CX), m->tab(T2_SX), 0);
    hipLaunchKernelGGL(k_dft<false>, dim3(m->nx), dim3(256), 0, m->stream, re, im, m->nx, m->ny,
                       m->tab(T2_CY), m->tab(T2_SY), 0);
    hipLaunchKernelGGL(k_kron_div, dim3(gb), dim3(256), 0, m->stream, f, re, im, e1, e2);
    hipLaunchKernelGGL(k_dft<true>, dim3(m->ny), dim3(256), 0, m->stream, re, im, m->nx, m->ny,
                       m->tab(T2_CX), m->tab(T2_SX), 1);
    hipLaunchKernelGGL(k_dft<false>, dim3(m->nx), dim3(256), 0, m->stream, re, im, m->nx, m->ny,
                       m->tab(T2_CY), m->tab(T2_SY), 1);
    (void)hipMemcpyAsync(out, re, sizeof(double) * m->nx * m->ny, hipMemcpyDeviceToDevice,
                         m->stream);
}

// the inverse mass of a 1-form (form 1) or 2-form (form 2) component
// (maxwell_2d_fem.jl:56-66)
void inv_mass(gempic_maxwell2d *m, int form, int component, const double *in, double *out)
{
    const double *ex, *ey;
    if (form == 1) {
        ex = m->tab(component == 1 ? T2_E1X : T2_E0X);
        ey = m->tab(component == 2 ? T2_E1Y : T2_E0Y);
    } else {
        ex = m->tab(component == 1 ? T2_E0X : T2_E1X);
        ey = m->tab(component == 2 ? T2_E0Y : T2_E1Y);
    }
    kron_solve(m, ex, ey, in, out);
}

// multiply_mass_2dkron! (maxwell_2d_fem.jl:343-363) with lines (a: mass 1 along x, b: along y)
void mass2(gempic_maxwell2d *m, bool onex, bool oney, const double *in, double *out)
{
    const Field2 f = m->f();
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    double *tmp = m->buf(13);
    hipLaunchKernelGGL(k_mass<true>, dim3(gb), dim3(256), 0, m->stream, f, in, tmp,
                       m->tab(onex ? T2_L1X : T2_L0X), onex ? m->degree - 1 : m->degree);
    hipLaunchKernelGGL(k_mass<false>, dim3(gb), dim3(256), 0, m->stream, f, tmp, out,
                       m->tab(oney ? T2_L1Y : T2_L0Y), oney ? m->degree - 1 : m->degree);
}

// which mass lines a (form, component) uses (inner_product, maxwell_2d_fem.jl:511-575)
void form_lines(int form, int component, bool &onex, bool &oney)
{
    onex = oney = false;
    if (form == 1) { onex = component == 1; oney = component == 2; }
    else if (form == 2) { onex = component != 1; oney = component != 2; }
    else if (form == 3) { onex = oney = true; }
}

size_t fb_rows_lds(const gempic_maxwell2d *m)
{
    const size_t H = kFieldRows + 1 + 2 * (size_t)m->degree;
    return sizeof(double) * (2 * H + 2 * (kFieldRows + 1) + kFieldRows + 1) * (size_t)m->nx;
}

size_t f_cols_lds(const gempic_maxwell2d *m) { return sizeof(double) * (kFieldCols + 1) * (size_t)m->ny; }

FieldTabs field_tabs(const gempic_maxwell2d *m)
{
    return FieldTabs{m->tab(T2_L0X), m->tab(T2_L1X), m->tab(T2_L0Y), m->tab(T2_L1Y),
                     m->tab(T2_K0X), m->tab(T2_K1X), m->tab(T2_K0Y), m->tab(T2_K1Y), m->degree};
}

// the row and column field kernels hold a few rows or columns in LDS; they
// take the grids of the 2D3V splitting (at most 4096 nodes)
constexpr int64_t kFusedNodes = 4096;

bool fused_setup(gempic_maxwell2d *m)
{
    if ((int64_t)m->nx * m->ny > kFusedNodes) return false;
    const size_t shb = fb_rows_lds(m), shc = f_cols_lds(m);
    if (shb > 65536 &&
        hipFuncSetAttribute((const void *)k_fb_rows, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shb) !=
            hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (shc > 65536 &&
        hipFuncSetAttribute((const void *)k_f_cols, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shc) !=
            hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return sizeof(double) * (size_t)(kFieldRows + 1) * m->nx <= 65536;
}

int finish2(gempic_maxwell2d *m)
{
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
    if (e != hipSuccess) return abi_fail(GEMPIC_ERR_DEVICE, (std::string("2D field kernels: ") + hipGetErrorString(e)).c_str());
    return GEMPIC_OK;
}

size_t nn2(const gempic_maxwell2d *m) { return (size_t)m->nx * m->ny; }

}  // namespace

// TwoDMaxwell operations on device arrays, queued on the solver's stream
// (used by the 2D3V splitting, split2d.hip)
hipStream_t gl::m2d_stream(gempic_maxwell2d *m) { return m->stream; }

void gl::m2d_e_from_b(gempic_maxwell2d *m, double dt, double *const *b, double *const *e)
{
    const Field2 f = m->f();
    if (m->fused) {
        const FieldTabs t = field_tabs(m);
        FieldRowsB ra{{b[0], b[1], b[2]}, {m->buf(9), m->buf(10), m->buf(11)}};
        hipLaunchKernelGGL(k_fb_rows, dim3((m->ny + kFieldRows - 1) / kFieldRows, 3), dim3(kFieldThreads),
                           fb_rows_lds(m), m->stream, f, t, ra);
        FieldCols ca{{m->buf(9), m->buf(10), m->buf(11)}, {t.k0y, t.k1y, t.k0y}, {e[0], e[1], e[2]}};
        hipLaunchKernelGGL(k_f_cols, dim3((m->nx + kFieldCols - 1) / kFieldCols, 3), dim3(kFieldThreads),
                           f_cols_lds(m), m->stream, f, ca, dt);
        return;
    }
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    mass2(m, false, true, b[0], m->buf(6));
    mass2(m, true, false, b[1], m->buf(7));
    mass2(m, true, true, b[2], m->buf(8));
    hipLaunchKernelGGL(k_curl_b, dim3(gb), dim3(256), 0, m->stream, f, m->buf(6), m->buf(7),
                       m->buf(8), m->buf(9), m->buf(10), m->buf(11));
    for (int c = 0; c < 3; ++c) {
        inv_mass(m, 1, c + 1, m->buf(9 + c), m->buf(12));
        hipLaunchKernelGGL(k_axpy, dim3(gb), dim3(256), 0, m->stream, f, e[c], m->buf(12), dt);
    }
}

void gl::m2d_b_from_e(gempic_maxwell2d *m, double dt, double *const *e, double *const *b)
{
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    hipLaunchKernelGGL(k_b_from_e, dim3(gb), dim3(256), 0, m->stream, m->f(), dt, e[0], e[1], e[2],
                       b[0], b[1], b[2]);
}

void gl::m2d_e_from_j(gempic_maxwell2d *m, const double *j, int component, double *e)
{
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    inv_mass(m, 1, component, j, m->buf(2));
    hipLaunchKernelGGL(k_sub, dim3(gb), dim3(256), 0, m->stream, m->f(), e, m->buf(2));
}

bool gl::m2d_e_from_parts(gempic_maxwell2d *m, const double *part2, int splits, double scale, double *j,
                          int component, double *e)
{
    if (!m->fused) return false;
    const Field2 f = m->f();
    const FieldTabs t = field_tabs(m);
    hipLaunchKernelGGL(k_fj_rows, dim3((m->ny + kFieldRows - 1) / kFieldRows), dim3(kFieldThreads),
                       sizeof(double) * (size_t)(kFieldRows + 1) * m->nx, m->stream, f, part2, splits, scale,
                       component == 1 ? t.k1x : t.k0x, j, m->buf(12));
    FieldCols ca{{m->buf(12), nullptr, nullptr}, {component == 2 ? t.k1y : t.k0y, nullptr, nullptr},
                 {e, nullptr, nullptr}};
    hipLaunchKernelGGL(k_f_cols, dim3((m->nx + kFieldCols - 1) / kFieldCols, 1), dim3(kFieldThreads),
                       f_cols_lds(m), m->stream, f, ca, -1.0);
    return true;
}

void gl::m2d_rho_from_e(gempic_maxwell2d *m, double *const *e, double *rho)
{
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    mass2(m, true, false, e[0], m->buf(2));
    mass2(m, false, true, e[1], m->buf(3));
    hipLaunchKernelGGL(k_div, dim3(gb), dim3(256), 0, m->stream, m->f(), m->buf(2), m->buf(3), rho);
}

namespace {

#define H2D(dst, src) \
    (void)hipMemcpyAsync((dst), (src), sizeof(double) * nn2(m), hipMemcpyHostToDevice, m->stream)
#define D2H(dst, src) \
    (void)hipMemcpyAsync((dst), (src), sizeof(double) * nn2(m), hipMemcpyDeviceToHost, m->stream)

}  // namespace

extern "C" {

int gempic_maxwell2d_create(int64_t nx, double xmin, double xmax, int64_t ny, double ymin,
                            double ymax, int degree, gempic_maxwell2d **out)
{
    if (!out) return abi_fail(GEMPIC_ERR_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (degree < 1 || degree > 3)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: TwoDMaxwell degree must be 1, 2 or 3");
    if (nx < 2 * degree + 1 || ny < 2 * degree + 1 || nx > kNxMax || ny > kNxMax)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: nx, ny must be in [2*degree+1, 256]");
    if (!(xmax > xmin) || !(ymax > ymin))
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: empty domain");
    int rc = abi_check_device();
    if (rc) return rc;
    gempic_maxwell2d *m = new gempic_maxwell2d();
    m->nx = (int)nx;
    m->ny = (int)ny;
    m->degree = degree;
    m->dx = (xmax - xmin) / (double)nx;     // TwoDGrid (mesh.jl:38-39)
    m->dy = (ymax - ymin) / (double)ny;
    std::vector<double> T((size_t)NT2 * kNxMax, 0.0);
    auto slot = [&](int s) { return T.data() + (size_t)s * kNxMax; };
    host_twiddles(nx, slot(T2_CX), slot(T2_SX));
    host_twiddles(ny, slot(T2_CY), slot(T2_SY));
    double l0[8], l1[8];
    host_mass_line(degree, l0);
    host_mass_line(degree - 1, l1);
    for (int k = 0; k <= degree; ++k) {
        slot(T2_L0X)[k] = l0[k] * m->dx;
        slot(T2_L0Y)[k] = l0[k] * m->dy;
    }
    for (int k = 0; k < degree; ++k) {
        slot(T2_L1X)[k] = l1[k] * m->dx;
        slot(T2_L1Y)[k] = l1[k] * m->dy;
    }
    host_mass_eig(nx, degree, slot(T2_L0X), slot(T2_E0X));
    host_mass_eig(ny, degree, slot(T2_L0Y), slot(T2_E0Y));
    host_mass_eig(nx, degree - 1, slot(T2_L1X), slot(T2_E1X));
    host_mass_eig(ny, degree - 1, slot(T2_L1Y), slot(T2_E1Y));
    host_inv_mass_kernel(nx, slot(T2_E0X), slot(T2_K0X));
    host_inv_mass_kernel(nx, slot(T2_E1X), slot(T2_K1X));
    host_inv_mass_kernel(ny, slot(T2_E0Y), slot(T2_K0Y));
    host_inv_mass_kernel(ny, slot(T2_E1Y), slot(T2_K1Y));
    // TwoDPoisson (poisson_2d_fem.jl:48-64)
    const double twopi = 6.283185307179586;
    for (int64_t j = 2; j <= nx; ++j) {
        const double angle = twopi * (double)(j - 1) / (double)nx;
        slot(T2_D1R)[j - 1] = (1.0 - std::cos(angle)) / m->dx;
        slot(T2_D1I)[j - 1] = std::sin(angle) / m->dx;
        slot(T2_DTMX)[j - 1] = 2.0 / (m->dx * m->dx) * (1.0 - std::cos(angle)) * slot(T2_E1X)[j - 1];
    }
    for (int64_t j = 2; j <= ny; ++j) {
        const double angle = twopi * (double)(j - 1) / (double)ny;
        slot(T2_D2R)[j - 1] = (1.0 - std::cos(angle)) / m->dy;
        slot(T2_D2I)[j - 1] = std::sin(angle) / m->dy;
        slot(T2_DTMY)[j - 1] = 2.0 / (m->dy * m->dy) * (1.0 - std::cos(angle)) * slot(T2_E1Y)[j - 1];
    }
    hipError_t e = hipStreamCreate(&m->stream);
    if (e == hipSuccess) e = dalloc2(&m->d_tab, T.size());
    if (e == hipSuccess) e = dalloc2(&m->d_buf, (size_t)16 * nx * ny);
    if (e == hipSuccess) e = dalloc2(&m->d_scalar, 8);
    if (e == hipSuccess)
        e = hipMemcpy(m->d_tab, T.data(), sizeof(double) * T.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) m->fused = fused_setup(m);
    if (e != hipSuccess) {
        gempic_maxwell2d_destroy(m);
        return abi_fail(GEMPIC_ERR_DEVICE, (std::string("TwoDMaxwell allocation: ") + hipGetErrorString(e)).c_str());
    }
    *out = m;
    return GEMPIC_OK;
}

int gempic_maxwell2d_destroy(gempic_maxwell2d *m)
{
    if (!m) return GEMPIC_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    if (m->d_tab) (void)hipFree(m->d_tab);
    if (m->d_buf) (void)hipFree(m->d_buf);
    if (m->d_scalar) (void)hipFree(m->d_scalar);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return GEMPIC_OK;
}

int gempic_maxwell2d_tables(gempic_maxwell2d *m, double *eig0x, double *eig0y, double *eig1x,
                            double *eig1y)
{
    if (!m) return abi_fail(GEMPIC_ERR_ARGUMENT, "null handle");
    const int slots[4] = {T2_E0X, T2_E0Y, T2_E1X, T2_E1Y};
    double *outs[4] = {eig0x, eig0y, eig1x, eig1y};
    for (int s = 0; s < 4; ++s)
        if (outs[s])
            (void)hipMemcpyAsync(outs[s], m->tab(slots[s]), sizeof(double) * (s % 2 ? m->ny : m->nx),
                                 hipMemcpyDeviceToHost, m->stream);
    return finish2(m);
}

int gempic_maxwell2d_solve(gempic_maxwell2d *m, int form, int component, const double *rhs,
                           double *out)
{
    if (!m || !rhs || !out) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL argument");
    if ((form != 1 && form != 2) || component < 1 || component > 3)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: form must be 1 or 2, component 1..3");
    H2D(m->buf(0), rhs);
    inv_mass(m, form, component, m->buf(0), m->buf(1));
    D2H(out, m->buf(1));
    return finish2(m);
}

int gempic_maxwell2d_multiply_mass(gempic_maxwell2d *m, int form, int component, const double *in,
                                   double *out)
{
    if (!m || !in || !out) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL argument");
    if (form < 0 || form > 3 || component < 1 || component > 3)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: form 0..3, component 1..3");
    bool ox, oy;
    form_lines(form, component, ox, oy);
    H2D(m->buf(0), in);
    mass2(m, ox, oy, m->buf(0), m->buf(1));
    D2H(out, m->buf(1));
    return finish2(m);
}

int gempic_maxwell2d_inner_product(gempic_maxwell2d *m, const double *c1, const double *c2,
                                   int component, int form, double *out)
{
    if (!m || !c1 || !c2 || !out) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL argument");
    if (form < 0 || form > 3 || component < 1 || component > 3)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: form 0..3, component 1..3");
    bool ox, oy;
    form_lines(form, component, ox, oy);
    H2D(m->buf(0), c1);
    H2D(m->buf(1), c2);
    mass2(m, ox, oy, m->buf(1), m->buf(2));
    hipLaunchKernelGGL(k_dot, dim3(1), dim3(1024), 0, m->stream, m->f(), m->buf(0), m->buf(2),
                       m->d_scalar);
    (void)hipMemcpyAsync(out, m->d_scalar, sizeof(double), hipMemcpyDeviceToHost, m->stream);
    return finish2(m);
}

int gempic_maxwell2d_compute_e_from_b(gempic_maxwell2d *m, double dt, const double *b1,
                                      const double *b2, const double *b3, double *e1, double *e2,
                                      double *e3)
{
    if (!m || !b1 || !b2 || !b3 || !e1 || !e2 || !e3)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL argument");
    const Field2 f = m->f();
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    H2D(m->buf(0), b1); H2D(m->buf(1), b2); H2D(m->buf(2), b3);
    H2D(m->buf(3), e1); H2D(m->buf(4), e2); H2D(m->buf(5), e3);
    mass2(m, false, true, m->buf(0), m->buf(6));
    mass2(m, true, false, m->buf(1), m->buf(7));
    mass2(m, true, true, m->buf(2), m->buf(8));
    hipLaunchKernelGGL(k_curl_b, dim3(gb), dim3(256), 0, m->stream, f, m->buf(6), m->buf(7),
                       m->buf(8), m->buf(9), m->buf(10), m->buf(11));
    for (int c = 0; c < 3; ++c) {
        inv_mass(m, 1, c + 1, m->buf(9 + c), m->buf(12));
        hipLaunchKernelGGL(k_axpy, dim3(gb), dim3(256), 0, m->stream, f, m->buf(3 + c), m->buf(12), dt);
    }
    D2H(e1, m->buf(3)); D2H(e2, m->buf(4)); D2H(e3, m->buf(5));
    return finish2(m);
}

int gempic_maxwell2d_compute_b_from_e(gempic_maxwell2d *m, double dt, const double *e1,
                                      const double *e2, const double *e3, double *b1, double *b2,
                                      double *b3)
{
    if (!m || !b1 || !b2 || !b3 || !e1 || !e2 || !e3)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL argument");
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    H2D(m->buf(0), e1); H2D(m->buf(1), e2); H2D(m->buf(2), e3);
    H2D(m->buf(3), b1); H2D(m->buf(4), b2); H2D(m->buf(5), b3);
    hipLaunchKernelGGL(k_b_from_e, dim3(gb), dim3(256), 0, m->stream, m->f(), dt, m->buf(0),
                       m->buf(1), m->buf(2), m->buf(3), m->buf(4), m->buf(5));
    D2H(b1, m->buf(3)); D2H(b2, m->buf(4)); D2H(b3, m->buf(5));
    return finish2(m);
}

int gempic_maxwell2d_compute_e_from_j(gempic_maxwell2d *m, const double *current, int component,
                                      double *e)
{
    if (!m || !current || !e) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL argument");
    if (component < 1 || component > 3) return abi_fail(GEMPIC_ERR_ARGUMENT, "component 1..3");
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    H2D(m->buf(0), current);
    H2D(m->buf(1), e);
    inv_mass(m, 1, component, m->buf(0), m->buf(2));
    hipLaunchKernelGGL(k_sub, dim3(gb), dim3(256), 0, m->stream, m->f(), m->buf(1), m->buf(2));
    D2H(e, m->buf(1));
    return finish2(m);
}

int gempic_maxwell2d_compute_rho_from_e(gempic_maxwell2d *m, const double *e1, const double *e2,
                                        const double *e3, double *rho)
{
    if (!m || !e1 || !e2 || !rho) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL argument");
    (void)e3;   // the third component's mass product is not used (maxwell_2d_fem.jl:476-478)
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    H2D(m->buf(0), e1); H2D(m->buf(1), e2);
    mass2(m, true, false, m->buf(0), m->buf(2));
    mass2(m, false, true, m->buf(1), m->buf(3));
    hipLaunchKernelGGL(k_div, dim3(gb), dim3(256), 0, m->stream, m->f(), m->buf(2), m->buf(3),
                       m->buf(4));
    D2H(rho, m->buf(4));
    return finish2(m);
}

int gempic_maxwell2d_compute_e_from_rho(gempic_maxwell2d *m, const double *rho, double *e1,
                                        double *e2)
{
    if (!m || !rho || !e1 || !e2) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL argument");
    const Field2 f = m->f();
    const unsigned gb = grid_blocks((int64_t)m->nx * m->ny);
    double *sr = m->buf(0), *si = m->buf(1);
    H2D(m->buf(2), rho);
    hipLaunchKernelGGL(k_load_real, dim3(gb), dim3(256), 0, m->stream, f, m->buf(2), sr, si);
    // fft2d! (poisson_2d_fem.jl:211-235): x then y
    hipLaunchKernelGGL(k_dft<true>, dim3(m->ny), dim3(256), 0, m->stream, sr, si, m->nx, m->ny,
                       m->tab(T2_CX), m->tab(T2_SX), 0);
    hipLaunchKernelGGL(k_dft<false>, dim3(m->nx), dim3(256), 0, m->stream, sr, si, m->nx, m->ny,
                       m->tab(T2_CY), m->tab(T2_SY), 0);
    PoissonTabs t{m->tab(T2_E0X), m->tab(T2_E0Y), m->tab(T2_DTMX), m->tab(T2_DTMY),
                  m->tab(T2_D1R), m->tab(T2_D1I), m->tab(T2_D2R), m->tab(T2_D2I)};
    hipLaunchKernelGGL(k_poisson_modes, dim3(gb), dim3(256), 0, m->stream, f, t, sr, si,
                       m->buf(3), m->buf(4), m->buf(5), m->buf(6));
    // ifft2d! (:184-209): y then x, real part
    for (int c = 0; c < 2; ++c) {
        double *r = m->buf(3 + 2 * c), *i = m->buf(4 + 2 * c);
        hipLaunchKernelGGL(k_dft<false>, dim3(m->nx), dim3(256), 0, m->stream, r, i, m->nx, m->ny,
                           m->tab(T2_CY), m->tab(T2_SY), 1);
        hipLaunchKernelGGL(k_dft<true>, dim3(m->ny), dim3(256), 0, m->stream, r, i, m->nx, m->ny,
                           m->tab(T2_CX), m->tab(T2_SX), 1);
    }
    D2H(e1, m->buf(3));
    D2H(e2, m->buf(5));
    return finish2(m);
}

// ---------------------------------------------------------------------------
// ParticleMeshCoupling2D verbs: device-resident forms and the batched host
// forms built on them
// ---------------------------------------------------------------------------
}  // extern "C"

namespace {

enum Pmc2Op { P2_CHARGE = 0, P2_EVAL, P2_EVAL2, P2_CHARGE_PP, P2_EVAL_PP };

// validated geometry of a ParticleMeshCoupling2D (constructor :22-45 on
// TwoDGrid, mesh.jl:31-43)
int pmc2_geometry(int op, int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                  double ymax, int degree, int smoothing, int64_t n, Pmc2Args &a)
{
    if (degree < 1 || degree > 3)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: 2D coupling degree must be 1, 2 or 3");
    if (nx < degree + 1 || ny < degree + 1 || nx > kNxMax || ny > kNxMax)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: nx, ny must be in [degree+1, 256]");
    if ((op == P2_CHARGE || op == P2_CHARGE_PP) && nx * ny > kGrid2Max)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: the device deposit holds nx*ny <= 4096");
    if (smoothing != GEMPIC_GALERKIN && smoothing != GEMPIC_COLLOCATION)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: smoothing type not implemented");
    if (!(xmax > xmin) || !(ymax > ymin)) return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: empty domain");
    if (n < 0) return abi_fail(GEMPIC_ERR_ARGUMENT, "n must be >= 0");
    a = Pmc2Args{};
    a.n = n;
    a.nx = (int)nx;
    a.ny = (int)ny;
    a.xmin = xmin;
    a.ymin = ymin;
    a.dx = (xmax - xmin) / (double)nx;
    a.dy = (ymax - ymin) / (double)ny;
    a.rdx = 1.0 / a.dx;
    a.rdy = 1.0 / a.dy;
    a.scaling = smoothing == GEMPIC_COLLOCATION ? 1.0 / (a.dx * a.dy) : 1.0;
    double pf[36];
    gh::splinepp_tables(degree, a.pc, pf);      // SplinePP(degree, n) (splinepp.jl:36-233)
    return GEMPIC_OK;
}

// Per-device workspace of the device-resident verbs: the per-workgroup
// deposit grids and the error flag. Grown on demand (hipMalloc only then);
// calls that overlap on one device share it and race.
struct Ws2 {
    double *part = nullptr;
    size_t cap = 0;          // doubles
    int *err = nullptr;
};
std::mutex g_ws2_mu;
Ws2 g_ws2[64];

int ws2_get(size_t part_doubles, Ws2 *&w)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return abi_fail(GEMPIC_ERR_DEVICE, "hipGetDevice failed");
    std::lock_guard<std::mutex> lock(g_ws2_mu);
    w = &g_ws2[dev];
    if (!w->err) {
        hipError_t e = hipMalloc((void **)&w->err, sizeof(int));
        if (e == hipSuccess) e = hipMemset(w->err, 0, sizeof(int));
        if (e != hipSuccess) {
            w->err = nullptr;
            return abi_fail(GEMPIC_ERR_DEVICE, (std::string("2D coupling workspace: ") + hipGetErrorString(e)).c_str());
        }
    }
    if (part_doubles > w->cap) {
        if (w->part) (void)hipFree(w->part);     // synchronizes with kernels still using it
        w->part = nullptr;
        w->cap = 0;
        hipError_t e = hipMalloc((void **)&w->part, sizeof(double) * part_doubles);
        if (e != hipSuccess)
            return abi_fail(GEMPIC_ERR_DEVICE, (std::string("2D coupling workspace: ") + hipGetErrorString(e)).c_str());
        w->cap = part_doubles;
    }
    return GEMPIC_OK;
}

template <int P, bool DET, bool PP>
const void *dep_fn() { return (const void *)k_pmc2d_deposit<P, DET, PP>; }

template <bool PP>
const void *dep_kernel_pp(int degree, bool det)
{
    if (det) return degree == 1 ? dep_fn<1, true, PP>() : degree == 2 ? dep_fn<2, true, PP>() : dep_fn<3, true, PP>();
    return degree == 1 ? dep_fn<1, false, PP>() : degree == 2 ? dep_fn<2, false, PP>() : dep_fn<3, false, PP>();
}

const void *dep_kernel(int degree, bool det, bool pp)
{
    return pp ? dep_kernel_pp<true>(degree, det) : dep_kernel_pp<false>(degree, det);
}

// LDS of k_pmc2d_deposit: grids (one per wave when deterministic), 16
// staging doubles per wave, 256 hash slots per wave (deterministic)
size_t dep_shmem(bool det, int nw, size_t nn)
{
    return sizeof(double) * ((det ? (size_t)nw : 1) * nn + (size_t)nw * 16) +
           (det ? sizeof(int) * 256 * (size_t)nw : 0);
}

struct DepPlan {
    int nw = 1, blocks = 1;
    size_t sh = 0;
};

// waves per workgroup (4 unless the deterministic per-wave grids do not fit
// in the workgroup's LDS) and workgroups (what the CUs hold at once)
int dep_plan(int degree, bool det, bool pp, int64_t n, size_t nn, DepPlan &pl)
{
    int dev = 0, maxsh = 65536, cus = abi_device_cus();
    if (hipGetDevice(&dev) == hipSuccess)
        (void)hipDeviceGetAttribute(&maxsh, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    const void *fn = dep_kernel(degree, det, pp);
    int nw = det ? GEMPIC_DEP_WAVES : GEMPIC_DEP_WAVES_FAST;
    while (nw > 1 && dep_shmem(det, nw, nn) > (size_t)maxsh) nw /= 2;
    size_t sh = dep_shmem(det, nw, nn);
    if (sh > 65536 &&
        hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh) != hipSuccess) {
        (void)hipGetLastError();
        while (nw > 1 && dep_shmem(det, nw, nn) > 65536) nw /= 2;
        sh = dep_shmem(det, nw, nn);
    }
    if (sh > (size_t)std::max(maxsh, 65536))
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: the 2D grid does not fit in the deposit's LDS");
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, nw * kWave, sh) != hipSuccess || occ < 1) {
        (void)hipGetLastError();
        occ = 1;
    }
    if (GEMPIC_DEP_BLOCKS_PER_CU > 0) occ = std::min(occ, GEMPIC_DEP_BLOCKS_PER_CU);
    const int64_t want = (n + (int64_t)nw * kWave - 1) / ((int64_t)nw * kWave);
    pl.nw = nw;
    pl.sh = sh;
    pl.blocks = (int)std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * occ, want));
    return GEMPIC_OK;
}

// Enqueue one verb on `s` over device arrays (a.x, a.y, ... set by the caller).
// With `fixed` the deposit runs in that workspace as it is (never grown): its
// workgroups are capped so that their grids and the split sums fit.
int pmc2_enqueue(int op, int degree, bool deterministic, Pmc2Args a, double *rho, hipStream_t s,
                 const Ws2 *fixed = nullptr)
{
    const int cus = abi_device_cus();
    const size_t nn = (size_t)a.nx * a.ny;
    int nblk = 0, splits = 0;
    DepPlan pl;
    const bool charge = op == P2_CHARGE || op == P2_CHARGE_PP, pp = op == P2_CHARGE_PP;
    if (charge) {
        int rc = dep_plan(degree, deterministic, pp, a.n, nn, pl);
        if (rc) return rc;
        if (fixed) {
            const int64_t grids = (int64_t)(fixed->cap / nn);      // workgroup grids + split sums
            if (grids < 2)
                return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: the 2D grid does not fit in the fixed deposit scratch");
            const int64_t bmax = grids >= 128 ? grids - 64 : grids / 2;
            pl.blocks = (int)std::min<int64_t>(pl.blocks, bmax);
        }
        const int64_t waves = (int64_t)pl.blocks * pl.nw;
        a.chunk = ((a.n + waves - 1) / waves + kWave - 1) / kWave * kWave;      // per wave
        nblk = (int)((a.n + (int64_t)pl.nw * a.chunk - 1) / ((int64_t)pl.nw * a.chunk));
        splits = std::min(nblk, 64);
    }
    const Ws2 *w = fixed;
    if (!w) {
        Ws2 *g = nullptr;
        int rc = ws2_get((size_t)(nblk + splits) * nn, g);
        if (rc) return rc;
        w = g;
    }
    a.err = w->err;
    a.part = w->part;
    if (charge) {
        const void *fn = dep_kernel(degree, deterministic, pp);
        void *args[] = {&a};
        hipError_t e = hipLaunchKernel(fn, dim3(nblk), dim3(pl.nw * kWave), args, pl.sh, s);
        if (e != hipSuccess)
            return abi_fail(GEMPIC_ERR_DEVICE, (std::string("2D deposit launch: ") + hipGetErrorString(e)).c_str());
        double *part2 = w->part + (size_t)nblk * nn;
        hipLaunchKernelGGL(k_pmc2d_reduce1, dim3((unsigned)((nn + 255) / 256), (unsigned)splits), dim3(256),
                           0, s, (const double *)w->part, nblk, (int)nn, part2);
        hipLaunchKernelGGL(k_pmc2d_reduce2, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s,
                           (const double *)part2, splits, (int)nn, rho);
    } else if (op == P2_EVAL_PP) {
        const unsigned gb = (unsigned)std::max<int64_t>(1, std::min<int64_t>((int64_t)cus * 8, (a.n + 255) / 256));
        if (degree == 1) hipLaunchKernelGGL(k_pmc2d_eval_pp<1>, dim3(gb), dim3(256), 0, s, a);
        else if (degree == 2) hipLaunchKernelGGL(k_pmc2d_eval_pp<2>, dim3(gb), dim3(256), 0, s, a);
        else hipLaunchKernelGGL(k_pmc2d_eval_pp<3>, dim3(gb), dim3(256), 0, s, a);
    } else {
        const bool multi = op == P2_EVAL2;
        const size_t ns = multi ? 2 * nn : (size_t)(a.nx + degree) * (a.ny + degree);
        const bool stage = ns <= GEMPIC_EVAL_STAGE_MAX;                 // <= 64 KiB of LDS
        const size_t sh = stage ? sizeof(double) * ns : 0;
        const int nt = GEMPIC_EVAL_THREADS;
        const unsigned gb = (unsigned)std::max<int64_t>(
            1, std::min<int64_t>((int64_t)cus * GEMPIC_EVAL_BLOCKS_PER_CU, (a.n + nt - 1) / nt));
#define EV(P, M, S) hipLaunchKernelGGL((k_pmc2d_eval<P, M, S>), dim3(gb), dim3(nt), sh, s, a)
#define EVD(M, S) do { if (degree == 1) EV(1, M, S); else if (degree == 2) EV(2, M, S); else EV(3, M, S); } while (0)
        if (multi) { if (stage) EVD(true, true); else EVD(true, false); }
        else { if (stage) EVD(false, true); else EVD(false, false); }
#undef EVD
#undef EV
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return abi_fail(GEMPIC_ERR_DEVICE, (std::string("2D coupling launch: ") + hipGetErrorString(e)).c_str());
    return GEMPIC_OK;
}

// wait for `s`, then report and clear the device's error flag
int pmc2_finish(hipStream_t s)
{
    Ws2 *w = nullptr;
    int rc = ws2_get(0, w);
    if (rc) return rc;
    int herr = 0;
    hipError_t e = hipMemcpyAsync(&herr, w->err, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && herr) e = hipMemsetAsync(w->err, 0, sizeof(int), s);
    if (e == hipSuccess && herr) e = hipStreamSynchronize(s);
    if (e != hipSuccess)
        return abi_fail(GEMPIC_ERR_DEVICE, (std::string("2D coupling: ") + hipGetErrorString(e)).c_str());
    if (herr & E_INEXACT)
        return abi_fail(GEMPIC_ERR_INEXACT, "InexactError: ceil(Int, x) of a non-finite or out-of-range position");
    if (herr & E_BOUNDS)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "BoundsError: a position lies outside the pp table");
    return GEMPIC_OK;
}

hipStream_t pmc2_stream()
{
    static thread_local hipStream_t s = nullptr;
    if (!s) (void)hipStreamCreate(&s);
    return s;
}

// the batched host-array forms: copy in, run deterministically, copy out
int pmc2_host(int op, int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
              double ymax, int degree, int smoothing, int64_t n, const double *x,
              const double *y, const double *w, const double *f1, const double *f2,
              double *out1, double *out2, double *rho)
{
    Pmc2Args a;
    int rc = pmc2_geometry(op, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, a);
    if (rc) return rc;
    rc = abi_check_device();
    if (rc) return rc;
    if (n == 0) return GEMPIC_OK;
    hipStream_t s = pmc2_stream();
    const size_t nn = (size_t)nx * ny;
    double *dx_ = nullptr, *dy_ = nullptr, *dw = nullptr, *df1 = nullptr, *df2 = nullptr;
    double *do1 = nullptr, *do2 = nullptr, *drho = nullptr;
    hipError_t e = hipSuccess;
    auto ok = [&](hipError_t r) { if (r != hipSuccess && e == hipSuccess) e = r; };
    ok(dalloc2(&dx_, n));
    ok(dalloc2(&dy_, n));
    const bool charge = op == P2_CHARGE || op == P2_CHARGE_PP;
    const size_t nf1 = op == P2_EVAL_PP ? nn * (size_t)(degree + 1) * (degree + 1) : nn;
    if (charge) { ok(dalloc2(&dw, n)); ok(dalloc2(&drho, nn)); }
    else { ok(dalloc2(&df1, nf1)); ok(dalloc2(&do1, n)); if (op == P2_EVAL2) { ok(dalloc2(&df2, nn)); ok(dalloc2(&do2, n)); } }
    rc = GEMPIC_OK;
    if (e == hipSuccess) {
        ok(hipMemcpyAsync(dx_, x, sizeof(double) * n, hipMemcpyHostToDevice, s));
        ok(hipMemcpyAsync(dy_, y, sizeof(double) * n, hipMemcpyHostToDevice, s));
        a.x = dx_;
        a.y = dy_;
        if (charge) {
            ok(hipMemcpyAsync(dw, w, sizeof(double) * n, hipMemcpyHostToDevice, s));
            ok(hipMemcpyAsync(drho, rho, sizeof(double) * nn, hipMemcpyHostToDevice, s));
            a.w = dw;
        } else {
            ok(hipMemcpyAsync(df1, f1, sizeof(double) * nf1, hipMemcpyHostToDevice, s));
            a.f1 = df1;
            a.pp = df1;
            a.out1 = do1;
            if (op == P2_EVAL2) {
                ok(hipMemcpyAsync(df2, f2, sizeof(double) * nn, hipMemcpyHostToDevice, s));
                a.f2 = df2;
                a.out2 = do2;
            }
        }
        if (e == hipSuccess) rc = pmc2_enqueue(op, degree, true, a, drho, s);
        if (e == hipSuccess && rc == GEMPIC_OK) {
            if (charge) ok(hipMemcpyAsync(rho, drho, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
            else {
                ok(hipMemcpyAsync(out1, do1, sizeof(double) * n, hipMemcpyDeviceToHost, s));
                if (op == P2_EVAL2) ok(hipMemcpyAsync(out2, do2, sizeof(double) * n, hipMemcpyDeviceToHost, s));
            }
        }
    }
    int frc = GEMPIC_OK;
    if (e == hipSuccess && rc == GEMPIC_OK) frc = pmc2_finish(s);
    else (void)hipStreamSynchronize(s);
    void *bufs[] = {dx_, dy_, dw, df1, df2, do1, do2, drho};
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    if (e != hipSuccess) return abi_fail(GEMPIC_ERR_DEVICE, (std::string("2D coupling: ") + hipGetErrorString(e)).c_str());
    if (rc) return rc;
    return frc;
}

// the device-resident forms: enqueue on the caller's stream; synchronous
// unless GEMPIC_ASYNC (then errors surface at the next synchronous call or
// gempic_pmc2d_sync)
int pmc2_dev(int op, int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
             double ymax, int degree, int smoothing, int64_t n, Pmc2Args io, double *rho,
             int flags, void *stream)
{
    Pmc2Args a;
    int rc = pmc2_geometry(op, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, a);
    if (rc) return rc;
    rc = abi_check_device();
    if (rc) return rc;
    if (n == 0) return GEMPIC_OK;
    a.x = io.x; a.y = io.y; a.w = io.w;
    a.f1 = io.f1; a.f2 = io.f2; a.out1 = io.out1; a.out2 = io.out2;
    hipStream_t s = (hipStream_t)stream;
    rc = pmc2_enqueue(op, degree, (flags & GEMPIC_DETERMINISTIC) != 0, a, rho, s);
    if (rc) return rc;
    if (flags & GEMPIC_ASYNC) return GEMPIC_OK;
    return pmc2_finish(s);
}

// ---------------------------------------------------------------------------
// The allocation-free deposit (gempic_pmc2d_add_charge_dev_noalloc): the
// error flag and the deposit's workgroup grids and split sums live in one
// file-scope device array, `scratch`, instead of the grown workspace. Each
// GPU holds its own copy of the array (GEMPIC_NOALLOC_SCRATCH_BYTES) until
// the process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
#ifndef GEMPIC_NOALLOC_SCRATCH_BYTES
#define GEMPIC_NOALLOC_SCRATCH_BYTES (32u << 20)
#endif
constexpr size_t kNoallocHeader = 256;      // bytes: the int error flag, then padding
static_assert(GEMPIC_NOALLOC_SCRATCH_BYTES % 256 == 0 && GEMPIC_NOALLOC_SCRATCH_BYTES > kNoallocHeader,
              "scratch size");

}  // namespace

static __device__ __align__(256) double scratch[GEMPIC_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {

__global__ void k_clear_flag(int *flag) { *flag = 0; }

std::mutex g_noalloc_mu;
Ws2 g_noalloc[64];                          // per device: views into that device's `scratch`

// the current device's view of `scratch`; resolved (and its flag cleared on
// `s`) at the first call on each device
int noalloc_ws(hipStream_t s, const Ws2 *&w)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64)
        return abi_fail(GEMPIC_ERR_DEVICE, "hipGetDevice failed");
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    Ws2 &v = g_noalloc[dev];
    if (!v.err) {
        void *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return abi_fail(GEMPIC_ERR_DEVICE, (std::string("deposit scratch: hipGetSymbolAddress: ") +
                                                hipGetErrorString(e)).c_str());
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, (int *)ptr);
        e = hipGetLastError();
        if (e != hipSuccess)
            return abi_fail(GEMPIC_ERR_DEVICE, (std::string("deposit scratch: ") + hipGetErrorString(e)).c_str());
        v.err = (int *)ptr;
        v.part = (double *)((char *)ptr + kNoallocHeader);
        v.cap = (GEMPIC_NOALLOC_SCRATCH_BYTES - kNoallocHeader) / sizeof(double);
    }
    w = &v;
    return GEMPIC_OK;
}

// wait for `s`, then report the scratch's error flag and clear it
int noalloc_finish(hipStream_t s, const Ws2 *w)
{
    int herr = 0;
    hipError_t e = hipMemcpyAsync(&herr, w->err, sizeof(int), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess && herr) {
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, s, w->err);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s);
    }
    if (e != hipSuccess)
        return abi_fail(GEMPIC_ERR_DEVICE, (std::string("2D deposit: ") + hipGetErrorString(e)).c_str());
    if (herr & E_INEXACT)
        return abi_fail(GEMPIC_ERR_INEXACT, "InexactError: ceil(Int, x) of a non-finite or out-of-range position");
    return GEMPIC_OK;
}

}  // namespace

extern "C" {

int gempic_pmc2d_add_charge_dev_noalloc(int64_t nx, int64_t ny, double xmin, double xmax,
                                        double ymin, double ymax, int degree, int smoothing,
                                        int64_t n, const double *x, const double *y,
                                        const double *wp, double *rho, int flags, void *stream)
{
    if ((n > 0 && (!x || !y || !wp)) || !rho) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    Pmc2Args a;
    int rc = pmc2_geometry(P2_CHARGE, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, a);
    if (rc) return rc;
    rc = abi_check_device();
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const Ws2 *w = nullptr;
    rc = noalloc_ws(s, w);
    if (rc) return rc;
    if (n > 0) {
        a.x = x; a.y = y; a.w = wp;
        rc = pmc2_enqueue(P2_CHARGE, degree, (flags & GEMPIC_DETERMINISTIC) != 0, a, rho, s, w);
        if (rc) return rc;
    }
    if (flags & GEMPIC_ASYNC) return GEMPIC_OK;
    return noalloc_finish(s, w);
}

int gempic_pmc2d_add_charge(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                            double ymax, int degree, int smoothing, int64_t n, const double *x,
                            const double *y, const double *wp, double *rho)
{
    if ((n > 0 && (!x || !y || !wp)) || !rho) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    return pmc2_host(P2_CHARGE, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, x, y, wp,
                     nullptr, nullptr, nullptr, nullptr, rho);
}

int gempic_pmc2d_evaluate(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                          double ymax, int degree, int smoothing, int64_t n, const double *x,
                          const double *y, const double *field, double *out)
{
    if ((n > 0 && (!x || !y || !out)) || !field) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    return pmc2_host(P2_EVAL, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, x, y, nullptr,
                     field, nullptr, out, nullptr, nullptr);
}

int gempic_pmc2d_evaluate_multiple(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                                   double ymax, int degree, int smoothing, int64_t n,
                                   const double *x, const double *y, const double *field1,
                                   const double *field2, double *out1, double *out2)
{
    if ((n > 0 && (!x || !y || !out1 || !out2)) || !field1 || !field2)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    return pmc2_host(P2_EVAL2, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, x, y,
                     nullptr, field1, field2, out1, out2, nullptr);
}

int gempic_pmc2d_add_charge_dev(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                                double ymax, int degree, int smoothing, int64_t n,
                                const double *x, const double *y, const double *wp, double *rho,
                                int flags, void *stream)
{
    if ((n > 0 && (!x || !y || !wp)) || !rho) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    Pmc2Args io{};
    io.x = x; io.y = y; io.w = wp;
    return pmc2_dev(P2_CHARGE, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, io, rho,
                    flags, stream);
}

int gempic_pmc2d_evaluate_dev(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                              double ymax, int degree, int smoothing, int64_t n, const double *x,
                              const double *y, const double *field, double *out, int flags,
                              void *stream)
{
    if ((n > 0 && (!x || !y || !out)) || !field) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    Pmc2Args io{};
    io.x = x; io.y = y; io.f1 = field; io.out1 = out;
    return pmc2_dev(P2_EVAL, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, io, nullptr,
                    flags, stream);
}

int gempic_pmc2d_evaluate_multiple_dev(int64_t nx, int64_t ny, double xmin, double xmax,
                                       double ymin, double ymax, int degree, int smoothing,
                                       int64_t n, const double *x, const double *y,
                                       const double *field1, const double *field2, double *out1,
                                       double *out2, int flags, void *stream)
{
    if ((n > 0 && (!x || !y || !out1 || !out2)) || !field1 || !field2)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    Pmc2Args io{};
    io.x = x; io.y = y; io.f1 = field1; io.f2 = field2; io.out1 = out1; io.out2 = out2;
    return pmc2_dev(P2_EVAL2, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, io, nullptr,
                    flags, stream);
}

int gempic_pmc2d_add_charge_pp(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                               double ymax, int degree, int smoothing, int64_t n, const double *x,
                               const double *y, const double *wp, double *rho)
{
    if ((n > 0 && (!x || !y || !wp)) || !rho) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    return pmc2_host(P2_CHARGE_PP, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, x, y, wp,
                     nullptr, nullptr, nullptr, nullptr, rho);
}

int gempic_pmc2d_evaluate_pp(int64_t nx, int64_t ny, double xmin, double xmax, double ymin,
                             double ymax, int degree, int smoothing, int64_t n, const double *x,
                             const double *y, const double *pp, double *out)
{
    if ((n > 0 && (!x || !y || !out)) || !pp) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    return pmc2_host(P2_EVAL_PP, nx, ny, xmin, xmax, ymin, ymax, degree, smoothing, n, x, y,
                     nullptr, pp, nullptr, out, nullptr, nullptr);
}

int gempic_b_to_pp_2d(int degree, int64_t nx, int64_t ny, const double *b, double *pp)
{
    if (!b || !pp) return abi_fail(GEMPIC_ERR_ARGUMENT, "NULL array");
    if (degree < 1 || degree > 3)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: SplinePP degree must be 1, 2 or 3");
    if (nx < degree + 1 || ny < degree + 1 || nx * ny > (int64_t)1 << 24)
        return abi_fail(GEMPIC_ERR_ARGUMENT, "ArgumentError: nx, ny must be >= degree+1, nx*ny <= 2^24");
    int rc = abi_check_device();
    if (rc) return rc;
    Pmc2Args a{};
    double pf[36];
    gh::splinepp_tables(degree, a.pc, pf);
    const size_t nn = (size_t)nx * ny, ld = (size_t)(degree + 1) * (degree + 1);
    hipStream_t s = pmc2_stream();
    double *db = nullptr, *dpp = nullptr;
    hipError_t e = dalloc2(&db, nn);
    if (e == hipSuccess) e = dalloc2(&dpp, nn * ld);
    if (e == hipSuccess) e = hipMemcpyAsync(db, b, sizeof(double) * nn, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        const dim3 g((unsigned)((nn + 255) / 256));
        if (degree == 1) hipLaunchKernelGGL(k_b_to_pp_2d<1>, g, dim3(256), 0, s, db, (int)nx, (int)ny, a, dpp);
        else if (degree == 2) hipLaunchKernelGGL(k_b_to_pp_2d<2>, g, dim3(256), 0, s, db, (int)nx, (int)ny, a, dpp);
        else hipLaunchKernelGGL(k_b_to_pp_2d<3>, g, dim3(256), 0, s, db, (int)nx, (int)ny, a, dpp);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(pp, dpp, sizeof(double) * nn * ld, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    else (void)hipStreamSynchronize(s);
    if (db) (void)hipFree(db);
    if (dpp) (void)hipFree(dpp);
    if (e != hipSuccess) return abi_fail(GEMPIC_ERR_DEVICE, (std::string("b_to_pp_2d: ") + hipGetErrorString(e)).c_str());
    return GEMPIC_OK;
}

int gempic_pmc2d_sync(void *stream)
{
    int rc = abi_check_device();
    if (rc) return rc;
    return pmc2_finish((hipStream_t)stream);
}

}  // extern "C"
