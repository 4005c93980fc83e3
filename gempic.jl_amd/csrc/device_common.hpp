// device_common.hpp -- device helpers shared by the 1D particle-field kernels.
//
// Everything here is FP64 and compiled with -ffp-contract=off so that each
// product and sum rounds exactly as in GEMPIC.jl's Julia code (no FMA
// contraction); per-particle arithmetic is then bit-identical to the
// reference's, and the only difference left is the summation order of the
// charge/current deposits.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gk {

constexpr int kWave = 64;

// a / b for a divisor b whose correctly rounded reciprocal rb = RN(1 / b) is
// known, in 3 FP64 instructions against about 10 for an IEEE division
// (Markstein's correction step, IBM J. Res. Develop. 34 (1990) 111-119):
//   q0 = RN(a * rb), at most 1.5 ulp from a / b (rb's half-ulp error, scaled
//        by a, can reach one ulp of the quotient, plus the rounding of q0);
//   r  = a - q0 * b, exact as an fma while a / b is a normal number;
//   q1 = RN(q0 + r * rb). Before its rounding, q0 + r * rb = a / b + (a / b - q0) * (b * rb - 1),
//        an error of at most 1.5 ulp * 2^-53.
// q1 is therefore the correctly rounded quotient unless a / b lies within
// 1.5 * 2^-53 ulp of a midpoint between two doubles. Such quotients exist
// only near a midpoint a / b cannot reach exactly (it lies at least 2^-54 ulp
// from every midpoint), and none has been found: the quotient tests compare
// this form with IEEE division on 32 million generic quotients and on the
// region where the bound is loosest (b's significand near 2, a / b's
// significand near 2), and the parity tests compare positions and cells
// with the oracle's IEEE division bit for bit.
__device__ __forceinline__ double div_by(double a, double b, double rb)
{
    const double q0 = a * rb;
    const double r = __builtin_fma(-q0, b, a);
    return __builtin_fma(r, rb, q0);
}

// Base.mod(x::Float64, y::Float64) of Julia: r = rem(x, y) (exact);
// zero takes y's sign; a nonzero r of the other sign gets y added.
// Used at hamiltonian_splitting_1d2v.jl:79 and hamiltonian_splitting_1d1v.jl:91.
// For y > 0 and |x| < 2y, rem(x, y) is x or x - sign(x)*y, and the latter is
// exact (Sterbenz: y <= |x| < 2y), so the library fmod is only needed for
// positions more than one period away.
__device__ __forceinline__ double jl_mod(double x, double y)
{
    double r;
    const double ax = fabs(x);
    if (ax < y) r = x;
    else if (ax < 2.0 * y) r = x - copysign(y, x);
    else r = fmod(x, y);
    if (r == 0.0) return copysign(0.0, y);
    if ((r > 0.0) != (y > 0.0)) return r + y;
    return r;
}

// floored integer modulus for n > 0; indices in [0, 2n) take a compare-and-
// subtract path instead of an integer division
__device__ __forceinline__ int imod(int a, int n)
{
    int r = a >= n ? a - n : a;
    if ((unsigned)r >= (unsigned)n) {
        r = a % n;
        r = r < 0 ? r + n : r;
    }
    return r;
}

// uniform_bsplines_eval_basis! (low_level_bsplines.jl:63-80): the P+1
// non-zero uniform B-splines of degree P at offset `off`, in registers.
template <int P>
__device__ __forceinline__ void bspline(double off, double (&b)[P + 1])
{
    b[0] = 1.0;
#pragma unroll
    for (int j = 1; j <= P; ++j) {
        double xx = -off;
        const double jr = (double)j;
        const double inv_j = 1.0 / jr;
        double saved = 0.0;
#pragma unroll
        for (int r = 0; r < j; ++r) {
            xx = xx + 1.0;
            double temp = b[r] * inv_j;
            b[r] = saved + xx * temp;
            saved = (jr - xx) * temp;
        }
        b[j] = saved;
    }
}

// Both bases of a smoother pair at one offset: b0 of degree P0 and b1 of
// degree P1. When P1 = P0 - 1 the recurrence for b0 passes through b1 after
// P1 rounds (the same operations as bspline<P1>), so b1 is copied out of it.
template <int P0, int P1>
__device__ __forceinline__ void bspline_pair(double off, double (&b0)[P0 + 1], double (&b1)[P1 + 1])
{
    if constexpr (P1 == P0 - 1) {
        b0[0] = 1.0;
#pragma unroll
        for (int j = 1; j <= P0; ++j) {
            double xx = -off;
            const double jr = (double)j;
            const double inv_j = 1.0 / jr;
            double saved = 0.0;
#pragma unroll
            for (int r = 0; r < j; ++r) {
                xx = xx + 1.0;
                double temp = b0[r] * inv_j;
                b0[r] = saved + xx * temp;
                saved = (jr - xx) * temp;
            }
            b0[j] = saved;
            if (j == P1) {
#pragma unroll
                for (int r = 0; r <= P1; ++r) b1[r] = b0[r];
            }
        }
        if constexpr (P1 == 0) b1[0] = 1.0;
    } else {
        bspline<P1>(off, b1);
        bspline<P0>(off, b0);
    }
}

// ---------------------------------------------------------------------------
// Fused arithmetic (the opt-in fast mode of the particle passes): the same
// quantities as the reference's recurrences, from closed polynomial forms
// evaluated with fma. They agree with the recurrences to a few ulp, not
// bitwise.
// ---------------------------------------------------------------------------

// The P+1 uniform B-splines of degree P <= 3 at offset t in [0, 1], in the
// order of uniform_bsplines_eval_basis! (b[0] is the spline ending at t = 1):
//   P = 1: 1 - t, t
//   P = 2: s^2/2, 1/2 + t s, t^2/2                        (s = 1 - t)
//   P = 3: s^3/6, 2/3 + t^2 (t/2 - 1), 2/3 + s^2 (s/2 - 1), t^3/6
template <int P>
__device__ __forceinline__ void bspline_fast(double t, double (&b)[P + 1])
{
    static_assert(P >= 0 && P <= 3, "closed forms for degrees 0 to 3");
    if constexpr (P == 0) {
        b[0] = 1.0;
    } else if constexpr (P == 1) {
        b[0] = 1.0 - t;
        b[1] = t;
    } else if constexpr (P == 2) {
        const double s = 1.0 - t;
        b[0] = 0.5 * (s * s);
        b[1] = __builtin_fma(t, s, 0.5);
        b[2] = 0.5 * (t * t);
    } else {
        const double s = 1.0 - t, t2 = t * t, s2 = s * s;
        constexpr double c6 = 1.0 / 6.0, c23 = 2.0 / 3.0;
        b[0] = (s2 * s) * c6;
        b[1] = __builtin_fma(t2, __builtin_fma(t, 0.5, -1.0), c23);
        b[2] = __builtin_fma(s2, __builtin_fma(s, 0.5, -1.0), c23);
        b[3] = (t2 * t) * c6;
    }
}

// Both bases of a smoother pair from one offset (closed forms; the squares
// are shared when P1 = P0 - 1)
template <int P0, int P1>
__device__ __forceinline__ void bspline_pair_fast(double t, double (&b0)[P0 + 1], double (&b1)[P1 + 1])
{
    if constexpr (P0 == 3 && P1 == 2) {
        const double s = 1.0 - t, t2 = t * t, s2 = s * s;
        constexpr double c6 = 1.0 / 6.0, c23 = 2.0 / 3.0;
        b0[0] = (s2 * s) * c6;
        b0[1] = __builtin_fma(t2, __builtin_fma(t, 0.5, -1.0), c23);
        b0[2] = __builtin_fma(s2, __builtin_fma(s, 0.5, -1.0), c23);
        b0[3] = (t2 * t) * c6;
        b1[0] = 0.5 * s2;
        b1[1] = __builtin_fma(t, s, 0.5);
        b1[2] = 0.5 * t2;
    } else {
        bspline_fast<P0>(t, b0);
        bspline_fast<P1>(t, b1);
    }
}

// Second derivatives of those splines (zero below degree 2)
template <int P>
__device__ __forceinline__ void bspline_d2_fast(double t, double (&d)[P + 1])
{
    if constexpr (P < 2) {
#pragma unroll
        for (int k = 0; k <= P; ++k) d[k] = 0.0;
    } else if constexpr (P == 2) {
        d[0] = 1.0;
        d[1] = -2.0;
        d[2] = 1.0;
    } else {
        d[0] = 1.0 - t;
        d[1] = __builtin_fma(3.0, t, -2.0);
        d[2] = __builtin_fma(-3.0, t, 1.0);
        d[3] = t;
    }
}

// Node index of the first of the P+1 nodes touched from cell `cell`
// (reference: index - p .. index, each taken mod n).
__device__ __forceinline__ int first_node(int cell, int P, int nx)
{
    // cell is in [0, nx)
    int k = cell - P;
    while (k < 0) k += nx;
    return k;
}

__device__ __forceinline__ int next_node(int k, int nx)
{
    k += 1;
    return k >= nx ? k - nx : k;
}

// Sum over the 64 lanes of a wave by an xor butterfly: every lane ends with
// the bitwise-identical total (each stage adds the same two operands in
// either order), so the result is deterministic.
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ __forceinline__ int lane_id() { return (int)(threadIdx.x & (kWave - 1)); }

// A double moved across lanes by a DPP pattern (two 32-bit DPP moves); lanes
// whose source is outside the row, or whose row is not in ROW_MASK, get 0.
template <int CTRL, int ROW_MASK = 0xf>
__device__ __forceinline__ double dpp_f64(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffll), CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Sum over the 64 lanes (all active), complete in lane 63, in a fixed order:
// running sums within each row of 16 lanes (row_shr 1, 2, 4, 8), then the
// row totals combined with row_bcast 15 and row_bcast 31. Register-only: no
// LDS round trips.
__device__ __forceinline__ double wave_sum_lane63(double v)
{
    v += dpp_f64<0x111>(v);          // row_shr:1
    v += dpp_f64<0x112>(v);          // row_shr:2
    v += dpp_f64<0x114>(v);          // row_shr:4
    v += dpp_f64<0x118>(v);          // row_shr:8
    v += dpp_f64<0x142, 0xa>(v);     // row_bcast:15 into rows 1 and 3
    v += dpp_f64<0x143, 0xc>(v);     // row_bcast:31 into rows 2 and 3
    return v;
}

// ---------------------------------------------------------------------------
// Deterministic deposit accumulation.
//
// Each lane keeps two cached cells (A, B) with the W = P+1 node sums of the
// particles it has seen in each. Particles are kept sorted by position, so a
// lane's particles mostly fall in one cell (entry A, in registers) and the
// odd one across a cell boundary lands in entry B, which lives in the lane's
// private LDS slots (B is touched rarely, and keeping it out of registers
// lets more waves share a SIMD). When a lane meets a third cell, its
// least-recently-used entry is flushed: the wave runs a segmented reduction
// over all lanes flushing the same cell (DPP sums, fixed order) and lane 63
// adds the W sums into the wave's private LDS row. The order of every
// addition is a fixed function of the particle order, so the deposit is
// bit-reproducible from run to run.
// ---------------------------------------------------------------------------
template <int WN>
struct Cache2 {
    int ca, cb;
    bool mru_b;
    double a[WN];
    double *bs;           // entry B: value w at bs[w * kWave] (this lane's LDS slots)

    // slots: the wave's WN * kWave doubles of LDS for this cache
    __device__ __forceinline__ void init(double *slots)
    {
        bs = slots + lane_id();
        reset();
    }
    __device__ __forceinline__ void reset()
    {
        ca = -1;
        cb = -1;
        mru_b = false;
#pragma unroll
        for (int w = 0; w < WN; ++w) {
            a[w] = 0.0;
            bs[w * kWave] = 0.0;
        }
    }
};

// Segmented flush: lanes with `want` set contribute v[w] to node
// first_node(cell, WN-1) + w of `row`; called by all 64 lanes of the wave.
template <int WN>
__device__ __forceinline__ void seg_flush(bool want, int cell, const double (&v)[WN], double *row,
                                          int nx)
{
    unsigned long long pending = __ballot(want);
    const int lane = lane_id();
    while (pending) {
        int leader = __ffsll((long long)pending) - 1;
        int c = __shfl(cell, leader, kWave);
        bool mine = want && (cell == c);
        unsigned long long m = __ballot(mine);
        int k = first_node(c, WN - 1, nx);
#pragma unroll
        for (int w = 0; w < WN; ++w) {
            double s = wave_sum_lane63(mine ? v[w] : 0.0);
            if (lane == kWave - 1) row[k] += s;
            k = next_node(k, nx);
        }
        pending &= ~m;
    }
}

template <int WN>
__device__ __forceinline__ void cache_add(Cache2<WN> &C, bool valid, int cell,
                                          const double (&v)[WN], double *row, int nx)
{
    bool hit_a = valid && cell == C.ca;
    bool hit_b = valid && !hit_a && cell == C.cb;
    bool miss = valid && !hit_a && !hit_b;
    if (__ballot(miss)) {
        bool use_a = miss && (C.ca < 0 || (C.cb >= 0 && C.mru_b));
        bool use_b = miss && !use_a;
        bool ev_a = use_a && C.ca >= 0;
        bool ev_b = use_b && C.cb >= 0;
        int ecell = ev_a ? C.ca : C.cb;
        double ev[WN];
#pragma unroll
        for (int w = 0; w < WN; ++w) ev[w] = ev_a ? C.a[w] : C.bs[w * kWave];
        seg_flush<WN>(ev_a || ev_b, ecell, ev, row, nx);
        if (use_a) {
            C.ca = cell;
#pragma unroll
            for (int w = 0; w < WN; ++w) C.a[w] = 0.0;
            hit_a = true;
        }
        if (use_b) {
            C.cb = cell;
#pragma unroll
            for (int w = 0; w < WN; ++w) C.bs[w * kWave] = 0.0;
            hit_b = true;
        }
    }
    // branches rather than selects: when every lane hits entry A (sorted
    // particles) the B update is skipped for the whole wave
    if (hit_a) {
#pragma unroll
        for (int w = 0; w < WN; ++w) C.a[w] += v[w];
        C.mru_b = false;
    }
    if (hit_b) {
#pragma unroll
        for (int w = 0; w < WN; ++w) C.bs[w * kWave] += v[w];
        C.mru_b = true;
    }
}

template <int WN>
__device__ __forceinline__ void cache_drain(Cache2<WN> &C, double *row, int nx)
{
    seg_flush<WN>(C.ca >= 0, C.ca, C.a, row, nx);
    double bv[WN];
#pragma unroll
    for (int w = 0; w < WN; ++w) bv[w] = C.bs[w * kWave];
    seg_flush<WN>(C.cb >= 0, C.cb, bv, row, nx);
    C.reset();
}

// Two rows cached together: entries of 2*WN values, the first WN for row0's
// nodes first_node(cell, WN-1) + w and the next WN for row1's same nodes
// (two deposits that always land in the same cell). Same caching as
// Cache2/cache_add; a flush writes both rows.
template <int WN>
__device__ __forceinline__ void seg_flush_rows2(bool want, int cell, const double (&v)[2 * WN],
                                                double *row0, double *row1, int nx)
{
    unsigned long long pending = __ballot(want);
    const int lane = lane_id();
    while (pending) {
        int leader = __ffsll((long long)pending) - 1;
        int c = __shfl(cell, leader, kWave);
        bool mine = want && (cell == c);
        unsigned long long m = __ballot(mine);
        const int k0 = first_node(c, WN - 1, nx);
        int k = k0;
#pragma unroll
        for (int w = 0; w < WN; ++w) {
            double s = wave_sum_lane63(mine ? v[w] : 0.0);
            if (lane == kWave - 1) row0[k] += s;
            k = next_node(k, nx);
        }
        k = k0;
#pragma unroll
        for (int w = 0; w < WN; ++w) {
            double s = wave_sum_lane63(mine ? v[WN + w] : 0.0);
            if (lane == kWave - 1) row1[k] += s;
            k = next_node(k, nx);
        }
        pending &= ~m;
    }
}

template <int WN>
__device__ __forceinline__ void cache_add_rows2(Cache2<2 * WN> &C, bool valid, int cell,
                                                const double (&v)[2 * WN], double *row0,
                                                double *row1, int nx)
{
    constexpr int W2 = 2 * WN;
    bool hit_a = valid && cell == C.ca;
    bool hit_b = valid && !hit_a && cell == C.cb;
    bool miss = valid && !hit_a && !hit_b;
    if (__ballot(miss)) {
        bool use_a = miss && (C.ca < 0 || (C.cb >= 0 && C.mru_b));
        bool use_b = miss && !use_a;
        bool ev_a = use_a && C.ca >= 0;
        bool ev_b = use_b && C.cb >= 0;
        int ecell = ev_a ? C.ca : C.cb;
        double ev[W2];
#pragma unroll
        for (int w = 0; w < W2; ++w) ev[w] = ev_a ? C.a[w] : C.bs[w * kWave];
        seg_flush_rows2<WN>(ev_a || ev_b, ecell, ev, row0, row1, nx);
        if (use_a) {
            C.ca = cell;
#pragma unroll
            for (int w = 0; w < W2; ++w) C.a[w] = 0.0;
            hit_a = true;
        }
        if (use_b) {
            C.cb = cell;
#pragma unroll
            for (int w = 0; w < W2; ++w) C.bs[w * kWave] = 0.0;
            hit_b = true;
        }
    }
    if (hit_a) {
#pragma unroll
        for (int w = 0; w < W2; ++w) C.a[w] += v[w];
        C.mru_b = false;
    }
    if (hit_b) {
#pragma unroll
        for (int w = 0; w < W2; ++w) C.bs[w * kWave] += v[w];
        C.mru_b = true;
    }
}

template <int WN>
__device__ __forceinline__ void cache_drain_rows2(Cache2<2 * WN> &C, double *row0, double *row1,
                                                  int nx)
{
    seg_flush_rows2<WN>(C.ca >= 0, C.ca, C.a, row0, row1, nx);
    double bv[2 * WN];
#pragma unroll
    for (int w = 0; w < 2 * WN; ++w) bv[w] = C.bs[w * kWave];
    seg_flush_rows2<WN>(C.cb >= 0, C.cb, bv, row0, row1, nx);
    C.reset();
}

}  // namespace gk
