// fs2d_jacobi.hpp -- device pieces shared by the one-grid sweep kernels
// (fs2d_sweeps.hip) and the batched kernels (fs2d_batch.hip): the diffuse
// division, the relaxation of one cell, and the register-strip Jacobi tile.
#pragma once

#include "fs2d_internal.hpp"

namespace fs2d {
namespace jac {

__device__ __forceinline__ float sgn(bool neg, float x) { return neg ? -x : x; }

// The divisor of a diffuse solve: d = 1 + 4a, y = RN(1/d), y_lo = RN(1/d - y)
// (so y + y_lo is 1/d to about 2^-48), and lo, the smallest nonzero
// |dividend| the FMA-based quotients take (the host's division_floor:
// max(2^-100, d 2^-124), or +inf when d > 2^100; raised where the two-product
// quotient is in use, see two_product_quotient).
struct Divisor {
    float d, y, lo, y_lo, y2, y2_lo;
};

__device__ __forceinline__ Divisor divisor_of(const SweepArgs& A) {
    return Divisor{A.denom, A.inv_denom, A.div_lo, A.inv_denom_lo, A.inv_hi2, A.inv_lo2};
}

// Dividends outside [lo, 2^100] (zero, tiny, huge, inf, NaN) take the IEEE
// division.  The test is wave-uniform: the common step, where no lane needs
// it, passes one untaken branch.
__device__ __forceinline__ float div_outside(float q, float x, Divisor v) {
    const float ax = fabsf(x);
    const bool out = !(ax >= v.lo && ax <= 0x1p100f);
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(out) != 0, 0)) q = out ? x / v.d : q;
    return q;
}

// x / d correctly rounded.  q1 = RN(x y + RN(x y_lo)) is within 2^-47
// relative of x/d before its rounding, so faithful; Markstein's step from a
// faithful quotient, its remainder exact by FMA, rounds correctly (P.
// Markstein, IBM J. R&D 34(1), 1990; Muller et al., Handbook of
// Floating-Point Arithmetic, FMA-based division), for lo <= |x| <= 2^100.
// Pinned against exact rational arithmetic in tests/test_division.py.
__device__ __forceinline__ float div_rn(float x, Divisor v) {
    const float q1 = __builtin_fmaf(x, v.y, x * v.y_lo);
    return div_outside(__builtin_fmaf(__builtin_fmaf(-q1, v.d, x), v.y, q1), x, v);
}

// x / d for a divisor whose two-product quotient RN(x y2 + RN(x y2_lo)) the
// host has checked to be correctly rounded for every finite dividend from lo
// up (Brisebarre, Muller and Raina, "Accelerating correctly rounded
// floating-point division when the divisor is known in advance", IEEE Trans.
// Computers 53(8), 2004; the check is two_product_quotient).  With y2_lo > 0
// an infinite dividend gives the infinity IEEE division gives, so one test,
// |x| >= lo (false for zero, tiny, subnormal and NaN), guards it.
__device__ __forceinline__ float div_rn_checked(float x, Divisor v) {
    float q = __builtin_fmaf(x, v.y2, x * v.y2_lo);
    const bool out = !(fabsf(x) >= v.lo);
    if (__builtin_expect(__builtin_amdgcn_ballot_w64(out) != 0, 0)) q = out ? x / v.d : q;
    return q;
}

// The update of one cell; operand order as in src/fluid_solver_cpu.cpp:107-108
// (diffuse: N + S + W + E) and :199-200 (pressure: d + E + W + S + N).
__device__ __forceinline__ float relax(int kind, float b, float n, float s, float w, float e, float a, Divisor dv) {
    if (kind == SWEEP_DIFFUSE) return div_rn(b + a * (((n + s) + w) + e), dv);
    return ((((b + e) + w) + s) + n) / 4.0f;
}

// ===================== Gauss-Seidel tile hyperplane ======================
// CHECKED: the diffuse quotient is div_rn_checked (a divisor the host checked).
template <int KIND, bool CHECKED = false>
__device__ __forceinline__ float relax_k(float b, float n, float s, float w, float e, float a, Divisor dv) {
    if (KIND == SWEEP_DIFFUSE) {
        const float x = b + a * (((n + s) + w) + e);
        return CHECKED ? div_rn_checked(x, dv) : div_rn(x, dv);
    }
    return ((((b + e) + w) + s) + n) / 4.0f;
}

// The same update with the diffuse quotient taken on the fast path only: the
// wave's lanes whose dividend the fast path does not cover are ORed into bad
// (a ballot, no branch), and the caller redoes its steps with relax_k when bad
// is nonzero.  Pressure updates have no division and leave bad alone.
template <int KIND, bool CHECKED>
__device__ __forceinline__ float relax_fast(float b, float n, float s, float w, float e, float a, Divisor dv,
                                            unsigned long long& bad) {
    if (KIND == SWEEP_DIFFUSE) {
        const float x = b + a * (((n + s) + w) + e);
        if (CHECKED) {
            bad |= __builtin_amdgcn_ballot_w64(!(fabsf(x) >= dv.lo));
            return __builtin_fmaf(x, dv.y2, x * dv.y2_lo);
        }
        const float ax = fabsf(x);
        bad |= __builtin_amdgcn_ballot_w64(!(ax >= dv.lo && ax <= 0x1p100f));
        const float q1 = __builtin_fmaf(x, dv.y, x * dv.y_lo);
        return __builtin_fmaf(__builtin_fmaf(-q1, dv.d, x), dv.y, q1);
    }
    return ((((b + e) + w) + s) + n) / 4.0f;
}

// ===================== Jacobi, register strips ===========================
// One wave per tile and no LDS: lane l holds column gj0 + l of a 64-column
// region, its TY + 2M rows in registers (iterate and right-hand side).  A
// sweep walks the rows top to bottom: west and east come from lanes l -+ 1
// by a full-wave DPP shift, north is the row's previous value (kept before it
// is overwritten), south the next row's.  Each sweep leaves one more row and
// lane at each side stale, so after M sweeps rows M .. M+TY-1 and lanes
// M .. 63-M hold sweep M; those are written.  Region cells outside the grid
// hold clamped values that only ever reach other outside cells or ring cells,
// and ring cells inside the region are rebuilt from the new interior after
// every sweep, so no test guards the interior update.
constexpr int JW_WAVES = 4;  // independent tiles per block

__device__ __forceinline__ float lane_west(float x) {  // lane l receives lane l - 1
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_east(float x) {  // lane l receives lane l + 1
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, false));
}

template <int M, int TY, int KIND, bool EDGE>
__device__ __forceinline__ void jacobi_strip(const SweepArgs& A, const float* __restrict__ src,
                                             float* __restrict__ dst, bool zero_src, int gi0, int gj0, int lane) {
    constexpr int H = TY + 2 * M;
    const int R = A.R, P = A.P;
    const int gj = gj0 + lane;
    float X[H], B[H];
    // ---- all loads in flight before the first use
    if (EDGE) {
        const int cj = min(max(gj, 0), R - 1), bj = min(max(gj, 1), R - 2);
        if (zero_src) {
#pragma unroll
            for (int q = 0; q < H; ++q) X[q] = 0.f;
        } else {
#pragma unroll
            for (int q = 0; q < H; ++q) X[q] = src[(size_t)min(max(gi0 + q, 0), R - 1) * P + cj];
        }
#pragma unroll
        for (int q = 1; q < H - 1; ++q) B[q] = A.b[(size_t)min(max(gi0 + q, 1), R - 2) * P + bj];
    } else {
        const size_t o = (size_t)gi0 * P + gj;
        if (zero_src) {
#pragma unroll
            for (int q = 0; q < H; ++q) X[q] = 0.f;
        } else {
#pragma unroll
            for (int q = 0; q < H; ++q) X[q] = src[o + (size_t)q * P];
        }
#pragma unroll
        for (int q = 1; q < H - 1; ++q) B[q] = A.b[o + (size_t)q * P];
    }
    B[0] = B[H - 1] = 0.f;

    const float a = A.a;
    const Divisor dv = divisor_of(A);
    const Ring rg = A.ring;
    const bool left = gj == 0, right = gj == R - 1;
    const bool cols_ring = EDGE && (gj0 <= 0 || gj0 + 63 >= R - 1);  // wave-uniform
#pragma unroll
    for (int s = 1; s <= M; ++s) {
        float north = X[s - 1];
#pragma unroll
        for (int q = s; q < H - s; ++q) {
            const float c = X[q];
            float v = relax_k<KIND>(B[q], north, X[q + 1], lane_west(c), lane_east(c), a, dv);
            north = c;
            if (EDGE && s < M) {
                // ring of this sweep: the bottom ring row from the new row above,
                // the top ring row (row q - 1) from the new first interior row,
                // the side columns from the new neighbouring lane
                const int gi = gi0 + q;
                if (gi == R - 1) v = sgn(rg.neg_tb, X[q - 1]);
                if (cols_ring) {
                    const float e = sgn(rg.neg_lr, lane_east(v)), w = sgn(rg.neg_lr, lane_west(v));
                    v = left ? e : (right ? w : v);
                }
                if (gi == 1) X[q - 1] = sgn(rg.neg_tb, v);
            }
            X[q] = v;
        }
    }

    // ---- write rows M .. M+TY-1, lanes M .. 63-M: interior cells and the ring
    // next to them
    if (lane < M || lane >= 64 - M) return;
    if (!EDGE) {
        float* o = dst + (size_t)gi0 * P + gj;
#pragma unroll
        for (int q = M; q < M + TY; ++q) o[(size_t)q * P] = X[q];
        return;
    }
    if (gj < 1 || gj > R - 2) return;
    const int n = R - 2;
#pragma unroll
    for (int q = M; q < M + TY; ++q) {
        const int gi = gi0 + q;
        if (gi < 1 || gi > n) continue;
        const float v = X[q];
        const size_t g = (size_t)gi * P + gj;
        dst[g] = v;
        const float tb = sgn(rg.neg_tb, v), lr = sgn(rg.neg_lr, v);
        const float cn = 0.5f * (tb + lr);
        if (gi == 1) dst[g - P] = tb;
        if (gi == n) dst[g + P] = tb;
        if (gj == 1) dst[g - 1] = lr;
        if (gj == n) dst[g + 1] = lr;
        if (gi == 1 && gj == 1) dst[g - P - 1] = cn;
        if (gi == 1 && gj == n) dst[g - P + 1] = cn;
        if (gi == n && gj == 1) dst[g + P - 1] = cn;
        if (gi == n && gj == n) dst[g + P + 1] = cn;
    }
}

// Blocks are dealt round-robin to the 8 XCDs; the first 8 * floor(nb / 8)
// are renumbered so that each XCD walks a contiguous run of tiles (rows of
// tiles in order), whose shared halo rows and columns then meet in its L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
    const int full = nb - nb % 8;
    return b < full ? (b % 8) * (full / 8) + b / 8 : b;
}

}  // namespace jac
}  // namespace fs2d
