// fs2d_sweeps.hip -- the diffuse and pressure relaxation sweeps on gfx950.
//
// Two sweep orders solve (1 + 4a) x - a * sum(nbrs) = x0 (diffuse,
// src/fluid_solver_cpu.cpp:95-114) and 4 p - sum(nbrs) = d (pressure,
// :196-204) with a boundary ring refreshed after every sweep:
//
//  * Gauss-Seidel, lexicographic, in place -- the reference's order.  Cell
//    (i, j) of sweep k reads (i-1, j) and (i, j-1) of sweep k and (i+1, j),
//    (i, j+1) of sweep k-1, and the ring of sweep k-1.  Every such input lies
//    on the hyperplane t - 1 of t = i + j + 2k, so cells are evaluated
//    hyperplane by hyperplane with results identical to the serial loop.  The
//    grid is cut into tiles of 64 rows x TW columns; a tile-sweep (I, J, k)
//    depends only on (I-1, J, k), (I, J-1, k), (I+1, J, k-1), (I, J+1, k-1),
//    all on level L - 1 of L = I + J + 2k, and every reader of a value this
//    tile-sweep overwrites also sits on level L - 1.  One launch per level
//    therefore runs all of its tile-sweeps in parallel, in place.  Inside a
//    tile one wave sweeps with lane = row, skewed by one column per lane (an
//    anti-diagonal per step): the west neighbour is the lane's own previous
//    result, the north neighbour the previous result of lane - 1, passed with
//    a full-wave DPP shift; south, east and the right-hand side come from LDS.
//    The ring of sweep k-1 (k >= 1) is the sign-adjusted adjacent interior
//    cell of sweep k-1, so it is never stored until the last sweep writes it;
//    sweep 0 reads the stored (stale) ring as the reference does.  The
//    default kernel, k_gs_unit, blocks M sweeps per tile in skewed
//    coordinates (one wave per sweep, lagged; see its comment).
//
//  * Jacobi -- every neighbour from the previous iterate.  M sweeps are fused
//    per launch (temporal blocking).  The default kernel, k_jacobi_wave, keeps
//    a tile's rows in registers, one wave per tile (fs2d_jacobi.hpp); the
//    alternative, k_jacobi_block, has a 256-thread block load its tile plus an
//    M-cell halo into LDS once, relax M times in LDS on a shrinking region
//    (recomputing the ring in LDS after each sweep), and write only the tile.
#include <type_traits>

#include "fs2d_internal.hpp"
#include "fs2d_jacobi.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <unordered_map>
#include <vector>

namespace fs2d {
namespace {

using namespace jac;

constexpr int GS_ROWS = 64;  // tile rows = wave lanes

// Value of iterate cell (gi, gj) as sweep k sees it (gi, gj already clamped
// into the grid): sweep 0 reads the stored ring, later sweeps the
// sign-adjusted adjacent interior cell of sweep k-1.
__device__ __forceinline__ const float* ring_source(const float* x, int gi, int gj, int R, int P, bool virtual_ring,
                                                    Ring rg, bool& neg, bool& ring_i, bool& ring_j) {
    ring_i = (gi == 0) | (gi == R - 1);
    ring_j = (gj == 0) | (gj == R - 1);
    int si = gi, sj = gj;
    neg = false;
    if (virtual_ring) {
        if (ring_i) {
            si = (gi == 0) ? 1 : R - 2;
            neg = rg.neg_tb;
        } else if (ring_j) {
            sj = (gj == 0) ? 1 : R - 2;
            neg = rg.neg_lr;
        }
    }
    return x + (size_t)si * P + sj;
}

template <int TW, int KIND>
__global__ __launch_bounds__(64) void k_gs_tile(SweepArgs A, const int2* __restrict__ work) {
    // T holds tile rows -1..64 and columns -1..TW (row pitch TP = TW + 2 is
    // even, so the skewed access (r, s - r) of 64 lanes hits 64 distinct
    // banks); Bm holds the right-hand side of the tile.
    constexpr int TP = TW + 2;
    constexpr int NQ = (TW + 2 + 63) / 64;  // halo columns per lane
    constexpr int NB = TW / 64;             // right-hand-side columns per lane
    constexpr int CH = 22;                  // rows per batch of in-flight loads (66 = 3 * 22)
    __shared__ float T[(GS_ROWS + 2) * TP];
    __shared__ float Bm[GS_ROWS * TW];

    const int2 wk = work[blockIdx.x];
    const int I = wk.x >> 16, J = wk.x & 0xffff, k = wk.y;
    const int R = A.R, P = A.P;
    float* __restrict__ x = A.x;
    const float* __restrict__ b = A.b;
    const int gi0 = 1 + GS_ROWS * I, gj0 = 1 + TW * J;
    const int rows = min(GS_ROWS, R - 1 - gi0), cols = min(TW, R - 1 - gj0);
    const int lane = threadIdx.x;
    const bool first = (k == 0);
    const bool zero_first = first && A.zero_start;
    const Ring rg = A.ring;

    // ---- load the sweep k-1 iterate with its halo.  A batch of loads is in
    // flight before any is consumed.  Tiles whose halo stays off the ring use
    // plain addressing; the others select each source address branch-free
    // (clamped into the grid; cells outside the tile and the halo corners
    // receive values that are never read).
    const bool plain = !zero_first && gi0 >= 2 && gi0 + GS_ROWS <= R - 2 && gj0 >= 2 && gj0 + TW <= R - 2;
    if (plain) {
        const float* base = x + (size_t)(gi0 - 1) * P + (gj0 - 1) + lane;
#pragma unroll
        for (int r0 = 0; r0 < GS_ROWS + 2; r0 += CH) {
            float vals[CH][NQ];
#pragma unroll
            for (int q = 0; q < CH; ++q)
#pragma unroll
                for (int t = 0; t < NQ; ++t)
                    vals[q][t] = (lane + 64 * t <= TW + 1) ? base[(size_t)(r0 + q) * P + 64 * t] : 0.f;
#pragma unroll
            for (int q = 0; q < CH; ++q)
#pragma unroll
                for (int t = 0; t < NQ; ++t)
                    if (lane + 64 * t <= TW + 1) T[(r0 + q) * TP + lane + 64 * t] = vals[q][t];
        }
    } else {
#pragma unroll
        for (int r0 = -1; r0 <= GS_ROWS; r0 += CH) {
            float vals[CH][NQ];
#pragma unroll
            for (int q = 0; q < CH; ++q) {
#pragma unroll
                for (int t = 0; t < NQ; ++t) {
                    const int r = r0 + q, c = lane - 1 + 64 * t;
                    const int gi = min(gi0 + r, R - 1), gj = min(gj0 + c, R - 1);
                    bool neg, ring_i, ring_j;
                    const float* src = ring_source(x, gi, gj, R, P, !first, rg, neg, ring_i, ring_j);
                    float v = *src;
                    v = neg ? -v : v;
                    // iterate starting at 0: only the north / west halos
                    // already hold sweep 0 of the neighbouring tiles
                    if (zero_first && !((r == -1 && !ring_i) || (c == -1 && !ring_j))) v = 0.f;
                    vals[q][t] = v;
                }
            }
#pragma unroll
            for (int q = 0; q < CH; ++q)
#pragma unroll
                for (int t = 0; t < NQ; ++t) {
                    const int r = r0 + q, c = lane - 1 + 64 * t;
                    if (c <= TW) T[(r + 1) * TP + c + 1] = vals[q][t];
                }
        }
    }
#pragma unroll
    for (int r0 = 0; r0 < GS_ROWS; r0 += 32) {
        float vals[32][NB];
#pragma unroll
        for (int q = 0; q < 32; ++q)
#pragma unroll
            for (int t = 0; t < NB; ++t) {
                const int r = min(r0 + q, rows - 1), c = min(lane + 64 * t, cols - 1);
                vals[q][t] = b[(size_t)(gi0 + r) * P + gj0 + c];
            }
#pragma unroll
        for (int q = 0; q < 32; ++q)
#pragma unroll
            for (int t = 0; t < NB; ++t) Bm[(r0 + q) * TW + lane + 64 * t] = vals[q][t];
    }
    __syncthreads();

    // ---- skewed sweep: at step s lane r relaxes column s - r.  The LDS
    // operands of step s + 1 are read during step s (they are old values:
    // lane r + 1 overwrites cell (r + 1, c + 1) only at step s + 2), so only
    // xprev and the DPP shift carry the dependency from step to step.
    // Inactive lanes compute on clamped operands and discard the result.
    const float a = A.a;
    const Divisor dv = divisor_of(A);
    const int steps = rows + cols - 1;
    const bool lane_ok = lane < rows;
    const float w_halo = T[(lane + 1) * TP];
    float* const trow = T + (lane + 1) * TP;
    float xprev = 0.f;
    int cc = min(max(-lane, 0), TW - 1);
    float so_n = trow[TP + cc + 1], e_n = trow[cc + 2], b_n = Bm[lane * TW + cc], nh_n = T[cc + 1];
    for (int s = 0; s < steps; ++s) {
        const int c = s - lane;
        const bool active = lane_ok && c >= 0 && c < cols;
        const int cw = cc;
        const float so = so_n, e = e_n, bb = b_n, n_halo = nh_n;
        cc = min(max(c + 1, 0), TW - 1);
        so_n = trow[TP + cc + 1];
        e_n = trow[cc + 2];
        b_n = Bm[lane * TW + cc];
        nh_n = T[cc + 1];
        // previous result of lane - 1 = cell (lane - 1, c) of this sweep
        const float from_up = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(xprev), 0x138, 0xf, 0xf, false));
        const float n = (lane == 0) ? n_halo : from_up;
        const float w = (c == 0) ? w_halo : xprev;
        const float v = relax_k<KIND>(bb, n, so, w, e, a, dv);
        if (active) trow[cw + 1] = v;
        xprev = active ? v : xprev;
    }
    __syncthreads();

    // ---- store the tile; the last sweep also writes the ring it borders.
#pragma unroll 8
    for (int r = 0; r < GS_ROWS; ++r)
#pragma unroll
        for (int t = 0; t < NB; ++t) {
            const int c = lane + 64 * t;
            if (r < rows && c < cols) x[(size_t)(gi0 + r) * P + gj0 + c] = T[(r + 1) * TP + c + 1];
        }
    if (k == A.K - 1) {
        const bool top = (gi0 == 1), bottom = (gi0 + rows == R - 1);
        const bool left = (gj0 == 1), right = (gj0 + cols == R - 1);
        for (int c = lane; c < cols; c += 64) {
            if (top) x[gj0 + c] = sgn(rg.neg_tb, T[TP + c + 1]);
            if (bottom) x[(size_t)(R - 1) * P + gj0 + c] = sgn(rg.neg_tb, T[rows * TP + c + 1]);
        }
        if (lane < rows) {
            if (left) x[(size_t)(gi0 + lane) * P] = sgn(rg.neg_lr, T[(lane + 1) * TP + 1]);
            if (right) x[(size_t)(gi0 + lane) * P + R - 1] = sgn(rg.neg_lr, T[(lane + 1) * TP + cols]);
        }
        if (lane == 0) {
            auto corner = [&](float v) { return 0.5f * (sgn(rg.neg_tb, v) + sgn(rg.neg_lr, v)); };
            if (top && left) x[0] = corner(T[TP + 1]);
            if (top && right) x[R - 1] = corner(T[TP + cols]);
            if (bottom && left) x[(size_t)(R - 1) * P] = corner(T[rows * TP + 1]);
            if (bottom && right) x[(size_t)(R - 1) * P + R - 1] = corner(T[rows * TP + cols]);
        }
    }
}

// ============ Gauss-Seidel, temporally blocked (M sweeps per unit) ==========
// In skewed coordinates i' = i + k, j' = j + k every Gauss-Seidel dependence
// points to smaller-or-equal (i', j', k), so rectangular tiles of
// 64 x TW cells x M sweeps are legal.  A unit (I', J', kappa) covers
// i' in 64 I' + [1, 64], j' in TW J' + [1, TW], k in kappa M + [0, M).  Wave w
// relaxes sweep kappa M + w; lane r owns skewed row i' = 64 I' + 1 + r and at
// step s relaxes skewed column j' = TW J' + 1 + s - r.  In these coordinates all
// M sweeps of a cell's neighbourhood sit on the same anti-diagonal: a cell's
// sweep-k value is written at step t, read by sweep k + 1 at step t + 1 and
// overwritten at step t + 2, so the M waves relax in place in LDS, each
// trailing the previous sweep's wave (the lagged schedule below).  Units with
// equal I' + J' + kappa touch disjoint cells and are launched together.
// Unit (I', J', kappa) of block b of a level launch: from the plan's work
// list, or, with no list, the b-th unit of level S.level in (kappa, I') order
// over the whole unit grid.  The implicit order includes units that hold no
// interior cell; such a unit relaxes and writes nothing.
__device__ __forceinline__ int2 unit_of(const SweepSet& S, const int2* __restrict__ work, int b) {
    if (work) return work[b];
    const int L = S.level;
    const int kap_lo = max(0, L - (S.gi - 1) - (S.gj - 1)), kap_hi = min(S.gk - 1, L);
    for (int kap = kap_lo; kap <= kap_hi; ++kap) {
        const int m = L - kap;
        const int lo = max(0, m - (S.gj - 1)), hi = min(S.gi - 1, m);
        if (b <= hi - lo) return make_int2(((lo + b) << 16) | (m - lo - b), kap);
        b -= hi - lo + 1;
    }
    return make_int2(0, 0);  // not reached: a level launches exactly its unit count
}

// LDS of a unit: the iterate region X and the right-hand side region Bs.
template <int TW, int M>
struct UnitLds {
    static constexpr int XH = GS_ROWS + 1 + M, XW = TW + 1 + M;  // iterate region
    static constexpr int XP = XW + (XW & 1);                      // even pitch: skew is conflict-free
    static constexpr int BH = GS_ROWS - 1 + M, BW = TW - 1 + M;   // right-hand side region
    static constexpr int BP = BW + (BW & 1);
    // Guard space after each array absorbs the operand reads of inactive lanes
    // (their skewed columns run up to GS_ROWS + TW + M past the cell); lanes
    // that relax nothing in a step store nothing.
    static constexpr int GUARD = GS_ROWS + TW + 2 * M + 8;
    static constexpr int NXS = XH * XP + GUARD, NBS = BH * BP + GUARD;
};

// CHECKED (diffuse only): every field's divisor passed two_product_quotient.
template <int TW, int M, int L, int KIND, bool CHECKED>
__device__ __forceinline__ void gs_unit_run(const SweepArgs& A, const int2 wk, float* __restrict__ X,
                                            float* __restrict__ Bs);

template <int TW, int M, int L, int KIND, bool CHECKED>
__global__ __launch_bounds__(64 * M) void k_gs_unit(SweepSet S, const int2* __restrict__ work) {
    // blocks [f * units, (f + 1) * units) relax field f of the set, unit by unit
    using U = UnitLds<TW, M>;
    __shared__ float X[U::NXS];
    __shared__ float Bs[U::NBS];
    const int f = blockIdx.x / S.units_per_field;
    gs_unit_run<TW, M, L, KIND, CHECKED>(S.a[f], unit_of(S, work, blockIdx.x - f * S.units_per_field), X, Bs);
}

// One level launch of two solves pipelined together: blocks
// [0, (count - 1) n_a) relax the diffuse fields 0 .. count - 2 of the set on the
// units of work_a (n_a per field), the rest relax the pressure field count - 1
// on the units of work_b.
template <int TW, int M, int L, bool CHECKED>
__global__ __launch_bounds__(64 * M) void k_gs_unit_mixed(SweepSet S, const int2* __restrict__ work_a, int n_a,
                                                          const int2* __restrict__ work_b) {
    using U = UnitLds<TW, M>;
    __shared__ float X[U::NXS];
    __shared__ float Bs[U::NBS];
    const int b = blockIdx.x, nd = (S.count - 1) * n_a;
    if (b < nd) {
        const int f = b / n_a;
        gs_unit_run<TW, M, L, SWEEP_DIFFUSE, CHECKED>(S.a[f], work_a[b - f * n_a], X, Bs);
    } else {
        gs_unit_run<TW, M, L, SWEEP_PRESSURE, false>(S.a[S.count - 1], work_b[b - nd], X, Bs);
    }
}

// One cell update in the plain units' interval loop: exact (relax_k, the IEEE
// division where the fast quotient does not cover the dividend) or fast
// (relax_fast, recording such lanes in bad).
template <int KIND, bool CHECKED, bool EXACT>
__device__ __forceinline__ float relax_step(std::integral_constant<bool, EXACT>, float b, float n, float s, float w,
                                            float e, float a, Divisor dv, unsigned long long& bad) {
    if constexpr (EXACT) return relax_k<KIND, CHECKED>(b, n, s, w, e, a, dv);
    else return relax_fast<KIND, CHECKED>(b, n, s, w, e, a, dv, bad);
}

template <int TW, int M, int L, int KIND, bool CHECKED>
__device__ __forceinline__ void gs_unit_run(const SweepArgs& A, const int2 wk, float* __restrict__ X,
                                            float* __restrict__ Bs) {
    using U = UnitLds<TW, M>;
    constexpr int NT = 64 * M;
    constexpr int XH = U::XH, XW = U::XW, XP = U::XP, BH = U::BH, BW = U::BW, BP = U::BP;
    constexpr int NX = (XH * XW + NT - 1) / NT, NBX = (BH * BW + NT - 1) / NT;
    const int I = wk.x >> 16, J = wk.x & 0xffff, kap = wk.y;
    const int R = A.R, P = A.P, n = R - 2, K = A.K;
    const int k_lo = kap * M, k_hi = min(k_lo + M, K) - 1, k_top = k_lo + M - 1;
    const int gi_lo = GS_ROWS * I - k_top, gj_lo = TW * J - k_top;  // global of X(0, 0)
    const int bi_lo = gi_lo + 1, bj_lo = gj_lo + 1;                 // global of Bs(0, 0)
    const int tid = threadIdx.x;
    float* __restrict__ x = A.x;
    const Ring rg = A.ring;

    // ---- load: iterate region and right-hand side in one batch, all loads in
    // flight before any LDS store (one memory round trip per unit).
    if constexpr (KIND == SWEEP_PRESSURE) {
        // By rows: the first TW columns of each region go as wave wv's rows
        // wv, wv + M, ..., lane l's columns l, l + 64, ..., so the row (and its
        // clamp into the grid) is wave-uniform and each load is a row pointer
        // plus the lane's clamped column.  The narrow tail columns (M + 1 of
        // the iterate region, M - 1 of the right-hand side) go cell by cell
        // over all threads.
        constexpr int CW = TW / 64;                                  // 64-column chunks of the first TW columns
        constexpr int RX = (XH + M - 1) / M, RB = (BH + M - 1) / M;  // rows per wave
        constexpr int XT = XW - TW, BT = BW - TW;                    // tail widths
        constexpr int TX = (XH * XT + NT - 1) / NT, TB = (BH * BT + NT - 1) / NT;
        static_assert(XT > 0 && BT > 0 && XT < 64 && BT < 64, "tails narrower than a chunk");
        const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
        int jx[CW], jb[CW];
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            jx[c] = min(max(gj_lo + 64 * c + lane, 0), R - 1);
            jb[c] = min(max(bj_lo + 64 * c + lane, 1), R - 2);
        }
        float v[RX][CW], vt[TX], vb[RB][CW], vbt[TB];
#pragma unroll
        for (int q = 0; q < RX; ++q) {
            // rows past the region (last pass of some waves) repeat its last row, unstored
            const int gi = min(max(gi_lo + min(wv + q * M, XH - 1), 0), R - 1);
            const float* __restrict__ xr = x + (size_t)gi * P;
#pragma unroll
            for (int c = 0; c < CW; ++c) v[q][c] = xr[jx[c]];
        }
#pragma unroll
        for (int q = 0; q < TX; ++q) {
            const int t = min(tid + q * NT, XH * XT - 1);
            const int gi = min(max(gi_lo + t / XT, 0), R - 1), gj = min(max(gj_lo + TW + t % XT, 0), R - 1);
            vt[q] = x[(size_t)gi * P + gj];
        }
        if (KIND == SWEEP_PRESSURE && A.du) {
            // the divergence of (du, dv), src/fluid_solver_cpu.cpp:188-193
            auto div = [&](const float* __restrict__ ur, const float* __restrict__ vn, const float* __restrict__ vs,
                           int j) { return A.neg_half_h * (ur[j + 1] - ur[j - 1] + vs[j] - vn[j]); };
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const size_t o = (size_t)min(max(bi_lo + min(wv + q * M, BH - 1), 1), R - 2) * P;
#pragma unroll
                for (int c = 0; c < CW; ++c) vb[q][c] = div(A.du + o, A.dv + o - P, A.dv + o + P, jb[c]);
            }
#pragma unroll
            for (int q = 0; q < TB; ++q) {
                const int t = min(tid + q * NT, BH * BT - 1);
                const size_t o = (size_t)min(max(bi_lo + t / BT, 1), R - 2) * P;
                vbt[q] = div(A.du + o, A.dv + o - P, A.dv + o + P, min(max(bj_lo + TW + t % BT, 1), R - 2));
            }
        } else {
#pragma unroll
            for (int q = 0; q < RB; ++q) {
                const float* __restrict__ br = A.b + (size_t)min(max(bi_lo + min(wv + q * M, BH - 1), 1), R - 2) * P;
#pragma unroll
                for (int c = 0; c < CW; ++c) vb[q][c] = br[jb[c]];
            }
#pragma unroll
            for (int q = 0; q < TB; ++q) {
                const int t = min(tid + q * NT, BH * BT - 1);
                const int gi = min(max(bi_lo + t / BT, 1), R - 2), gj = min(max(bj_lo + TW + t % BT, 1), R - 2);
                vbt[q] = A.b[(size_t)gi * P + gj];
            }
        }
#pragma unroll
        for (int q = 0; q < RX; ++q) {
            const int row = wv + q * M;
            if (row >= XH) break;
#pragma unroll
            for (int c = 0; c < CW; ++c) X[row * XP + 64 * c + lane] = v[q][c];
        }
#pragma unroll
        for (int q = 0; q < TX; ++q) {
            const int t = tid + q * NT;
            if (t < XH * XT) X[(t / XT) * XP + TW + t % XT] = vt[q];
        }
#pragma unroll
        for (int q = 0; q < RB; ++q) {
            const int row = wv + q * M;
            if (row >= BH) break;
#pragma unroll
            for (int c = 0; c < CW; ++c) Bs[row * BP + 64 * c + lane] = vb[q][c];
        }
#pragma unroll
        for (int q = 0; q < TB; ++q) {
            const int t = tid + q * NT;
            if (t < BH * BT) Bs[(t / BT) * BP + TW + t % BT] = vbt[q];
        }
    } else {
        // Cell by cell over the flattened regions (measured faster for the
        // diffuse kernels than the row form above, slower for the pressure ones).
        float v[NX], vb[NBX];
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int idx = min(tid + q * NT, XH * XW - 1);
            const int gi = min(max(gi_lo + idx / XW, 0), R - 1), gj = min(max(gj_lo + idx % XW, 0), R - 1);
            v[q] = x[(size_t)gi * P + gj];
        }
#pragma unroll
        for (int q = 0; q < NBX; ++q) {
            const int idx = min(tid + q * NT, BH * BW - 1);
            const int gi = min(max(bi_lo + idx / BW, 1), R - 2), gj = min(max(bj_lo + idx % BW, 1), R - 2);
            vb[q] = A.b[(size_t)gi * P + gj];
        }
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int idx = tid + q * NT;
            if (idx < XH * XW) X[(idx / XW) * XP + idx % XW] = v[q];
        }
#pragma unroll
        for (int q = 0; q < NBX; ++q) {
            const int idx = tid + q * NT;
            if (idx < BH * BW) Bs[(idx / BW) * BP + idx % BW] = vb[q];
        }
    }

    // ---- the M sweeps
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), r = tid & 63;
    const int k = k_lo + w;
    const int i = GS_ROWS * I + 1 + r - k;
    const bool wave_ok = (k <= k_hi);
    const int jbase = TW * J + 1 - k;
    const int ofs = M - w;  // X column of skewed column 0 is 1 + (M - 1 - w)
    const float a = A.a;
    const Divisor dv = divisor_of(A);
    float xprev = 0.f;
    // A unit whose cells all lie in rows and columns 2..n-1 never meets the
    // ring: its sweep needs no clamping (the LDS arrays carry guard space for
    // the reads of inactive lanes, whose results are discarded).  The unit
    // kernel expects a zero-start iterate to be zero-filled by the caller.
    const bool plain = GS_ROWS * I + 1 - k_hi >= 2 && GS_ROWS * I + GS_ROWS - k_lo <= n - 1 &&
                       TW * J + 1 - k_hi >= 2 && TW * J + TW - k_lo <= n - 1;
    // Lagged waves: wave w runs its local step s at tick s + L w, and a barrier
    // opens every interval of L ticks.  An operand of step s from the previous
    // sweep was written by wave w - 1 at its step s - 1, i.e. at tick
    // s + L w - L - 1, before the interval's barrier, and no other wave writes
    // a cell this wave reads within the interval; so each wave reads all
    // operands of its L steps right after the barrier and then relaxes them
    // with only the DPP / register chain between steps.
    constexpr int STEPS = GS_ROWS + TW - 1;
    constexpr int NINT = (STEPS + L * (M - 1) + L - 1) / L;
    // The operand reads of an interval are pinned ahead of its relaxations
    // (sched_barrier), so they are never sunk into per-tick branches.
    const int me0 = (1 + r + (M - 1 - w)) * XP + ofs - r;  // X index of this lane's cell at step s: me0 + s
    const int bme0 = (r + (M - 1 - w)) * BP + ofs - 1 - r;  // Bs index of its right-hand side: bme0 + s
    if (plain) {
        // Step phases (lane r relaxes column c = s - r): while s < 64, lane
        // r's first cell (c = 0, whose west neighbour is the tile halo) is at
        // step r and lanes r <= s are active; while 64 <= s < TW every lane
        // is inside the tile (c in [1, TW - 1]); from s = TW on lanes
        // r >= s - (TW - 1) are active.  Intervals start at multiples of L,
        // so an interval lies in one phase.  The north halo matters only to
        // lane 0, which is active while s < TW; it enters as the DPP shift's
        // old value (lane 0 has no source lane).
        static_assert(GS_ROWS == 64 && TW % 64 == 0 && 64 % L == 0, "the step phases assume 64-row tiles");
        // The west neighbour of a lane's first cell (step r) lies in the tile's
        // west halo column.  Waves w' <= w never write that cell (they would
        // reach it at a negative tile column); waves w' > w write it at tick
        // r - 1 + (2 + L) w' - 2 w, after this wave's step r (tick r + L w).
        // So its value at the unit's start is the one step r reads: read it
        // once, into xprev, which a lane keeps until its first step (r), so
        // that step takes it as W like every later step takes the lane's
        // previous cell.
        __syncthreads();
        xprev = X[me0 + r - 1];
        const int guard_r = XH * XP + r;  // lane r's word in the guard space after X
        for (int it = 0; it < NINT; ++it) {
            __syncthreads();
            const int s0 = it * L - L * w;
            // a wave with no step of this interval in [0, STEPS) only joins the barrier
            if (!wave_ok || s0 + L <= 0 || s0 >= STEPS) continue;
            const int x0 = me0 + s0, b0 = bme0 + s0;
            float so[L], e[L], bb[L];
            // An interval's steps take the fast quotient; if any lane's dividend
            // needed the IEEE division, they are redone exactly from the
            // interval's start (no other wave reads this interval's stores
            // before the next barrier, so the first pass's stores are simply
            // overwritten).  One branch per interval instead of one per step.
            unsigned long long bad = 0;
            auto run_interval = [&](auto steps) {
                const float xp0 = xprev;
                steps(std::false_type{});
                if constexpr (KIND == SWEEP_DIFFUSE) {
                    if (__builtin_expect(bad != 0, 0)) {
                        xprev = xp0;
                        steps(std::true_type{});
                    }
                }
            };
            if (s0 < 64) {
                float nh[L];
#pragma unroll
                for (int q = 0; q < L; ++q) {
                    so[q] = X[x0 + q + XP];
                    e[q] = X[x0 + q + 1];
                    bb[q] = Bs[b0 + q];
                    nh[q] = X[x0 + q - XP];
                }
                __builtin_amdgcn_sched_barrier(0);
                auto steps = [&](auto exact_t) {
#pragma unroll
                    for (int q = 0; q < L; ++q) {
                        const int s = s0 + q;
                        const float nv = __int_as_float(__builtin_amdgcn_update_dpp(
                            __float_as_int(nh[q]), __float_as_int(xprev), 0x138, 0xf, 0xf, false));
                        const float v = relax_step<KIND, CHECKED>(exact_t, bb[q], nv, so[q], xprev, e[q], a, dv, bad);
                        const bool act = (r <= s);
                        // an idle lane stores to its own guard word (no exec mask change)
                        X[(act ? x0 : guard_r - q) + q] = v;
                        xprev = act ? v : xprev;
                    }
                };
                run_interval(steps);
            } else if (s0 < TW) {
                float nh[L];
#pragma unroll
                for (int q = 0; q < L; ++q) {
                    so[q] = X[x0 + q + XP];
                    e[q] = X[x0 + q + 1];
                    bb[q] = Bs[b0 + q];
                    nh[q] = X[x0 + q - XP];
                }
                __builtin_amdgcn_sched_barrier(0);
                auto steps = [&](auto exact_t) {
#pragma unroll
                    for (int q = 0; q < L; ++q) {
                        const float nv = __int_as_float(__builtin_amdgcn_update_dpp(
                            __float_as_int(nh[q]), __float_as_int(xprev), 0x138, 0xf, 0xf, false));
                        const float v = relax_step<KIND, CHECKED>(exact_t, bb[q], nv, so[q], xprev, e[q], a, dv, bad);
                        X[x0 + q] = v;
                        xprev = v;
                    }
                };
                run_interval(steps);
            } else {
#pragma unroll
                for (int q = 0; q < L; ++q) {
                    so[q] = X[x0 + q + XP];
                    e[q] = X[x0 + q + 1];
                    bb[q] = Bs[b0 + q];
                }
                __builtin_amdgcn_sched_barrier(0);
                auto steps = [&](auto exact_t) {
#pragma unroll
                    for (int q = 0; q < L; ++q) {
                        const int s = s0 + q;
                        const float nv =
                            __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(xprev), 0x138, 0xf, 0xf, false));
                        const float v = relax_step<KIND, CHECKED>(exact_t, bb[q], nv, so[q], xprev, e[q], a, dv, bad);
                        X[(r >= s - (TW - 1) ? x0 : guard_r - q) + q] = v;
                        xprev = v;
                    }
                };
                run_interval(steps);
            }
        }
    } else {
        // Units that meet the ring: a neighbour on the ring is, for sweep
        // k > 0, the sign-adjusted previous-sweep value of the cell itself
        // (read with the other operands; the cell is overwritten only after
        // its own relaxation), and for sweep 0 the stored ring, which the
        // unit loaded.  Write-back stores the ring of every edge cell, so the
        // next unit finds the ring of the latest sweep there.  The loop is
        // specialised by the edges the unit meets (top/bottom rows, left/right
        // columns); the tests it leaves out cannot hold in such a unit.
        const bool row_ok = wave_ok && i >= 1 && i <= n;
        const bool kpos = (k > 0);
        const bool top = row_ok && i == 1, bot = row_ok && i == n;
        const int c_left = 1 - jbase, c_right = n - jbase;  // skewed columns of j = 1 and j = n
        const int c_lo = row_ok ? max(0, c_left) : TW, c_hi = min(TW - 1, c_right);
        const unsigned tb_sign = rg.neg_tb ? 0x80000000u : 0u, lr_sign = rg.neg_lr ? 0x80000000u : 0u;
        const bool n_virtual = top && kpos, s_virtual = bot && kpos;
        // W at the lane's first column: the tile halo, or on the left edge
        // the ring (virtual for k > 0); wave-uniform
        const bool w_virtual = kpos && c_left >= 0;
        const bool rows_edge = GS_ROWS * I + 1 - k_hi <= 1 || GS_ROWS * I + GS_ROWS - k_lo >= n;
        const bool cols_edge = TW * J + 1 - k_hi <= 1 || TW * J + TW - k_lo >= n;
        // West of the lane's first cell (c_lo, at step c_lo + r): the tile's west
        // halo, written only by later waves after that step, or the ring, which
        // no sweep writes -- read once (as in the plain units).
        __syncthreads();
        const float w_first = X[me0 + c_lo + r - 1];
        const int guard_r = XH * XP + r;  // lane r's word in the guard space after X
        auto sweeps = [&](auto rows_t, auto cols_t) {
            constexpr bool RE = decltype(rows_t)::value, CE = decltype(cols_t)::value;
            for (int it = 0; it < NINT; ++it) {
                __syncthreads();
                const int s0 = it * L - L * w;
                if (!wave_ok || s0 + L <= 0 || s0 >= STEPS) continue;
                const int x0 = me0 + s0, b0 = bme0 + s0;
                float so[L], e[L], bb[L], nh[L], self[L];
#pragma unroll
                for (int q = 0; q < L; ++q) {
                    so[q] = X[x0 + q + XP];
                    e[q] = X[x0 + q + 1];
                    bb[q] = Bs[b0 + q];
                    nh[q] = X[x0 + q - XP];
                    self[q] = X[x0 + q];
                }
                __builtin_amdgcn_sched_barrier(0);
                // fast quotient for the interval, redone exactly if needed (as in the plain units)
                unsigned long long bad = 0;
                auto steps = [&](auto exact_t) {
#pragma unroll
                    for (int q = 0; q < L; ++q) {
                        const int c = s0 + q - r;
                        // lane 0 receives the tile halo as the shift's old value
                        const float up = __int_as_float(__builtin_amdgcn_update_dpp(
                            __float_as_int(nh[q]), __float_as_int(xprev), 0x138, 0xf, 0xf, false));
                        float nv = up, sv = so[q], ev = e[q], wv;
                        if constexpr (RE) {
                            const float tbv = __uint_as_float(__float_as_uint(self[q]) ^ tb_sign);
                            nv = n_virtual ? tbv : (top ? nh[q] : up);
                            sv = s_virtual ? tbv : so[q];
                        }
                        if constexpr (CE) {
                            const float lrv = __uint_as_float(__float_as_uint(self[q]) ^ lr_sign);
                            wv = (c == c_lo) ? (w_virtual ? lrv : w_first) : xprev;
                            ev = (kpos && c == c_right) ? lrv : e[q];
                        } else {
                            wv = (c == c_lo) ? w_first : xprev;
                        }
                        const float v = relax_step<KIND, CHECKED>(exact_t, bb[q], nv, sv, wv, ev, a, dv, bad);
                        if constexpr (KIND == SWEEP_PRESSURE) {
                            X[(c >= c_lo && c <= c_hi ? x0 : guard_r - q) + q] = v;
                        } else {
                            if (c >= c_lo && c <= c_hi) X[x0 + q] = v;
                        }
                        xprev = v;
                    }
                };
                const float xp0 = xprev;
                steps(std::false_type{});
                if constexpr (KIND == SWEEP_DIFFUSE) {
                    if (__builtin_expect(bad != 0, 0)) {
                        xprev = xp0;
                        steps(std::true_type{});
                    }
                }
            }
        };
        using yes = std::integral_constant<bool, true>;
        using no = std::integral_constant<bool, false>;
        if (rows_edge && cols_edge)
            sweeps(yes{}, yes{});
        else if (rows_edge)
            sweeps(yes{}, no{});
        else
            sweeps(no{}, yes{});
    }
    __syncthreads();

    // ---- write back every cell this unit relaxed (its latest sweep).  A unit
    // that meets the ring then writes the ring beside each of its edge cells
    // (after the solve's last sweep that is the final ring), from LDS.  The
    // stores are non-temporal: the next level reads them from beyond this
    // XCD's L2 anyway, and fewer dirty L2 lines shorten the launch boundary.
    auto relaxed = [&](int gi, int gj) {
        const int lo = max(k_lo, max(GS_ROWS * I + 1 - gi, TW * J + 1 - gj));
        const int hi = min(k_hi, min(GS_ROWS * I + GS_ROWS - gi, TW * J + TW - gj));
        return gi >= 1 && gi <= n && gj >= 1 && gj <= n && lo <= hi;
    };
    {
        // Region row li holds row gi of the sweeps k in [ka, kb] that relax
        // it; those relax columns TW J + 1 - kb .. TW J + TW - ka (relaxed()
        // for one row).  Rows are wave-uniform, as in the load.
        constexpr int CX = (XW + 63) / 64;
        const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
        for (int li = wv; li < XH; li += M) {
            const int gi = gi_lo + li;
            const int ka = max(k_lo, GS_ROWS * I + 1 - gi), kb = min(k_hi, GS_ROWS * I + GS_ROWS - gi);
            if (gi < 1 || gi > n || ka > kb) continue;
            const int j_lo = max(TW * J + 1 - kb, 1), j_hi = min(TW * J + TW - ka, n);
            float* __restrict__ xr = x + (size_t)gi * P;
#pragma unroll
            for (int c = 0; c < CX; ++c) {
                const int lj = 64 * c + lane, gj = gj_lo + lj;
                if (lj < XW && gj >= j_lo && gj <= j_hi) __builtin_nontemporal_store(X[li * XP + lj], xr + gj);
            }
        }
    }
    if (!plain) {
        // rows 1 and n, then columns 1 and n, of this unit's region
        for (int e = tid; e < 2 * XW + 2 * XH; e += NT) {
            const bool row = e < 2 * XW;
            const int t = row ? e % XW : (e - 2 * XW) % XH;
            const bool far = row ? e >= XW : e - 2 * XW >= XH;
            const int gi = row ? (far ? n : 1) : gi_lo + t;
            const int gj = row ? gj_lo + t : (far ? n : 1);
            if (!relaxed(gi, gj)) continue;
            const float v = X[(gi - gi_lo) * XP + (gj - gj_lo)];
            const size_t o = (size_t)gi * P + gj;
            if (row) {
                const float tb = sgn(rg.neg_tb, v);
                const size_t ring = far ? o + P : o - P;
                __builtin_nontemporal_store(tb, x + ring);
                const float cn = 0.5f * (tb + sgn(rg.neg_lr, v));  // the corners beside it
                if (gj == 1) __builtin_nontemporal_store(cn, x + ring - 1);
                if (gj == n) __builtin_nontemporal_store(cn, x + ring + 1);
            } else {
                __builtin_nontemporal_store(sgn(rg.neg_lr, v), x + (far ? o + 1 : o - 1));
            }
        }
    }
}

// ===================== Jacobi, M sweeps per launch =======================
constexpr int JX = 64, JY = 32, JT = 256;

template <int M>
__global__ __launch_bounds__(JT) void k_jacobi_block(SweepArgs A, const float* __restrict__ src,
                                                    float* __restrict__ dst, int zero_src) {
    constexpr int W = JX + 2 * M, H = JY + 2 * M;                // iterate region (tile + M halo)
    constexpr int BW = JX + 2 * (M - 1), BH = JY + 2 * (M - 1);  // right-hand side region
    constexpr int NX = (H * W + JT - 1) / JT, NBX = (BH * BW + JT - 1) / JT;
    __shared__ float buf[2][H * W];
    __shared__ float rhs[BH * BW];

    const int R = A.R, P = A.P;
    const int gi0 = 1 + JY * blockIdx.y - M, gj0 = 1 + JX * blockIdx.x - M;  // global of buf(0, 0)
    const int tid = threadIdx.x;
    const Ring rg = A.ring;

    // all loads of the region are issued before any LDS store
    {
        float vals[NX];
#pragma unroll
        for (int q = 0; q < NX; ++q) {
            const int idx = tid + q * JT;
            const int r = idx / W, c = idx % W;
            const int gi = min(max(gi0 + r, 0), R - 1), gj = min(max(gj0 + c, 0), R - 1);
            vals[q] = zero_src ? 0.f : src[(size_t)gi * P + gj];
        }
#pragma unroll
        for (int q = 0; q < NX; ++q)
            if (tid + q * JT < H * W) buf[0][tid + q * JT] = vals[q];
    }
    {
        float vals[NBX];
#pragma unroll
        for (int q = 0; q < NBX; ++q) {
            const int idx = tid + q * JT;
            const int r = idx / BW, c = idx % BW;
            const int gi = min(max(gi0 + 1 + r, 1), R - 2), gj = min(max(gj0 + 1 + c, 1), R - 2);
            vals[q] = A.b[(size_t)gi * P + gj];
        }
#pragma unroll
        for (int q = 0; q < NBX; ++q)
            if (tid + q * JT < BH * BW) rhs[tid + q * JT] = vals[q];
    }
    __syncthreads();

    const int kind = A.kind;
    const float a = A.a;
    const Divisor dv = divisor_of(A);
#pragma unroll
    for (int s = 1; s <= M; ++s) {
        const int e = M - s;  // the region relaxed by sweep s is the tile grown by e
        const int rw = JX + 2 * e, rh = JY + 2 * e;
        const float* cur = buf[(s - 1) & 1];
        float* nxt = buf[s & 1];
        const bool last = (s == M);
        for (int idx = tid; idx < rh * rw; idx += JT) {
            const int r = M - e + idx / rw, c = M - e + idx % rw;  // buf coordinates
            const int gi = gi0 + r, gj = gj0 + c;
            if (gi < 1 || gi > R - 2 || gj < 1 || gj > R - 2) continue;
            const int o = r * W + c;
            const float v = relax(kind, rhs[(r - 1) * BW + c - 1], cur[o - W], cur[o + W], cur[o - 1], cur[o + 1], a,
                                  dv);
            const float tb = sgn(rg.neg_tb, v), lr = sgn(rg.neg_lr, v);
            const float cn = 0.5f * (tb + lr);
            if (!last) {
                nxt[o] = v;
                // ring cells of this sweep (their inner neighbour is in the region)
                if (gi == 1) nxt[o - W] = tb;
                if (gi == R - 2) nxt[o + W] = tb;
                if (gj == 1) nxt[o - 1] = lr;
                if (gj == R - 2) nxt[o + 1] = lr;
                if (gi == 1 && gj == 1) nxt[o - W - 1] = cn;
                if (gi == 1 && gj == R - 2) nxt[o - W + 1] = cn;
                if (gi == R - 2 && gj == 1) nxt[o + W - 1] = cn;
                if (gi == R - 2 && gj == R - 2) nxt[o + W + 1] = cn;
            } else {
                const size_t g = (size_t)gi * P + gj;
                dst[g] = v;
                if (gi == 1) dst[g - P] = tb;
                if (gi == R - 2) dst[g + P] = tb;
                if (gj == 1) dst[g - 1] = lr;
                if (gj == R - 2) dst[g + 1] = lr;
                if (gi == 1 && gj == 1) dst[g - P - 1] = cn;
                if (gi == 1 && gj == R - 2) dst[g - P + 1] = cn;
                if (gi == R - 2 && gj == 1) dst[g + P - 1] = cn;
                if (gi == R - 2 && gj == R - 2) dst[g + P + 1] = cn;
            }
        }
        __syncthreads();
    }
}

template <int M, int TY, int KIND>
__global__ __launch_bounds__(64 * JW_WAVES) void k_jacobi_wave(SweepArgs A, const float* __restrict__ src,
                                                              float* __restrict__ dst, int zero_src, int tiles_x,
                                                              int tiles) {
    const int t = xcd_block(blockIdx.x, gridDim.x) * JW_WAVES + threadIdx.x / 64;
    if (t >= tiles) return;  // whole wave
    const int ty = t / tiles_x, tx = t - ty * tiles_x;
    constexpr int OX = 64 - 2 * M;
    const int gi0 = 1 + TY * ty - M, gj0 = 1 + OX * tx - M;
    const int R = A.R;
    const bool plain = gi0 >= 1 && gi0 + TY + 2 * M - 1 <= R - 2 && gj0 >= 1 && gj0 + 63 <= R - 2;
    const int lane = threadIdx.x & 63;
    if (plain)
        jacobi_strip<M, TY, KIND, false>(A, src, dst, zero_src != 0, gi0, gj0, lane);
    else
        jacobi_strip<M, TY, KIND, true>(A, src, dst, zero_src != 0, gi0, gj0, lane);
}

constexpr int JW_TY = 32;  // rows per tile

template <int M>
hipError_t launch_jw(hipStream_t s, const SweepArgs& a, const float* src, float* dst, int zero_src) {
    constexpr int OX = 64 - 2 * M;
    const int n = a.R - 2;
    const int tiles_x = (n + OX - 1) / OX, tiles_y = (n + JW_TY - 1) / JW_TY;
    const int tiles = tiles_x * tiles_y;
    const dim3 grid((tiles + JW_WAVES - 1) / JW_WAVES);
    if (a.kind == SWEEP_DIFFUSE)
        hipLaunchKernelGGL((k_jacobi_wave<M, JW_TY, SWEEP_DIFFUSE>), grid, dim3(64 * JW_WAVES), 0, s, a, src, dst,
                           zero_src, tiles_x, tiles);
    else
        hipLaunchKernelGGL((k_jacobi_wave<M, JW_TY, SWEEP_PRESSURE>), grid, dim3(64 * JW_WAVES), 0, s, a, src, dst,
                           zero_src, tiles_x, tiles);
    return hipGetLastError();
}

hipError_t launch_jacobi_wave(hipStream_t s, const SweepArgs& a, const float* src, float* dst, int zero_src,
                              int m) {
    switch (m) {
        case 1: return launch_jw<1>(s, a, src, dst, zero_src);
        case 2: return launch_jw<2>(s, a, src, dst, zero_src);
        case 3: return launch_jw<3>(s, a, src, dst, zero_src);
        case 4: return launch_jw<4>(s, a, src, dst, zero_src);
        case 5: return launch_jw<5>(s, a, src, dst, zero_src);
        case 6: return launch_jw<6>(s, a, src, dst, zero_src);
        case 7: return launch_jw<7>(s, a, src, dst, zero_src);
        case 8: return launch_jw<8>(s, a, src, dst, zero_src);
        default: return hipErrorInvalidValue;
    }
}

template <int M>
hipError_t launch_jb(hipStream_t s, const SweepArgs& a, const float* src, float* dst, int zero_src) {
    const dim3 grid((a.R - 2 + JX - 1) / JX, (a.R - 2 + JY - 1) / JY);
    hipLaunchKernelGGL(k_jacobi_block<M>, grid, dim3(JT), 0, s, a, src, dst, zero_src);
    return hipGetLastError();
}

hipError_t launch_jacobi_block(hipStream_t s, const SweepArgs& a, const float* src, float* dst, int zero_src,
                               int m) {
    switch (m) {
        case 1: return launch_jb<1>(s, a, src, dst, zero_src);
        case 2: return launch_jb<2>(s, a, src, dst, zero_src);
        case 3: return launch_jb<3>(s, a, src, dst, zero_src);
        case 4: return launch_jb<4>(s, a, src, dst, zero_src);
        case 5: return launch_jb<5>(s, a, src, dst, zero_src);
        case 6: return launch_jb<6>(s, a, src, dst, zero_src);
        case 7: return launch_jb<7>(s, a, src, dst, zero_src);
        case 8: return launch_jb<8>(s, a, src, dst, zero_src);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

// The pressure-only unit launches, built in a translation unit of their own
// (fs2d_sweeps_p.hip, with LLVM's ILP-first machine scheduler, which times
// faster for these kernels and slower for the diffuse and joint ones).
bool launch_pressure_units(hipStream_t s, int TW, int M, int L, dim3 grid, const SweepSet& level,
                           const int2* w)
#ifdef FS2D_PRESSURE_UNITS_ONLY
{
#define FS2D_P(tw, m, l)                                                                         \
    if (TW == tw && M == m && L == l) {                                                          \
        hipLaunchKernelGGL((k_gs_unit<tw, m, l, SWEEP_PRESSURE, false>), grid, dim3(64 * m), 0, s, \
                           level, w);                                                            \
        return true;                                                                             \
    }
#define FS2D_PL(tw, m) FS2D_P(tw, m, 1) FS2D_P(tw, m, 4) FS2D_P(tw, m, 8)
    FS2D_PL(64, 2) FS2D_PL(64, 4) FS2D_PL(64, 8) FS2D_PL(128, 4) FS2D_PL(128, 8)
#undef FS2D_PL
#undef FS2D_P
    return false;
}
#else
    ;
#endif

#ifndef FS2D_PRESSURE_UNITS_ONLY

// ---- Gauss-Seidel plan ----------------------------------------------------
hipError_t gs_plan_create(GsPlan* plan, int R, int K, int TW) {
    plan->TW = TW;
    plan->M = 0;
    plan->GI = (R - 2 + GS_ROWS - 1) / GS_ROWS;
    plan->GJ = (R - 2 + TW - 1) / TW;
    plan->K = K;
    const int raw_levels = (K > 0) ? (plan->GI - 1) + (plan->GJ - 1) + 2 * (K - 1) + 1 : 0;
    plan->levels = 0;
    plan->level_of = nullptr;  // tile plans are never pipelined
    plan->host_work = nullptr;
    plan->work = nullptr;
    plan->host_offsets = nullptr;
    if (plan->GI >= (1 << 15) || plan->GJ >= (1 << 16)) return hipErrorInvalidValue;
    // Tile-sweeps per hyperplane level; empty levels (e.g. odd levels of a
    // one-tile grid) are dropped, which keeps the order of the others.
    std::vector<int> count(raw_levels, 0);
    for (int k = 0; k < K; ++k)
        for (int I = 0; I < plan->GI; ++I)
            for (int J = 0; J < plan->GJ; ++J) ++count[I + J + 2 * k];
    std::vector<int> slot(raw_levels, -1);
    std::vector<int> offsets;
    int acc = 0;
    for (int L = 0; L < raw_levels; ++L) {
        if (count[L] == 0) continue;
        slot[L] = static_cast<int>(offsets.size());
        offsets.push_back(acc);
        acc += count[L];
    }
    plan->levels = static_cast<int>(offsets.size());
    offsets.push_back(acc);
    plan->host_offsets = new int[offsets.size()];
    std::copy(offsets.begin(), offsets.end(), plan->host_offsets);
    if (acc == 0) return hipSuccess;
    std::vector<int2> work(acc);
    std::vector<int> fill(offsets.begin(), offsets.end() - 1);
    for (int k = 0; k < K; ++k)
        for (int I = 0; I < plan->GI; ++I)
            for (int J = 0; J < plan->GJ; ++J) work[fill[slot[I + J + 2 * k]]++] = make_int2((I << 16) | J, k);
    hipError_t e = hipMalloc(&plan->work, sizeof(int2) * acc);
    if (e != hipSuccess) return e;
    return hipMemcpy(plan->work, work.data(), sizeof(int2) * acc, hipMemcpyHostToDevice);
}

hipError_t gs_unit_plan_create(GsPlan* plan, int R, int K, int TW, int M, int lag) {
    return gs_unit_plan_build(plan, R, K, TW, M, lag, true);
}

hipError_t gs_unit_plan_build(GsPlan* plan, int R, int K, int TW, int M, int lag, bool upload) {
    const int n = R - 2;
    plan->TW = TW;
    plan->M = M;
    plan->lag = lag;
    plan->GI = (n + K - 1 + GS_ROWS - 1) / GS_ROWS;
    plan->GJ = (n + K - 1 + TW - 1) / TW;
    plan->K = K;
    plan->levels = 0;
    plan->work = nullptr;
    plan->host_offsets = nullptr;
    plan->level_of = nullptr;
    plan->host_work = nullptr;
    if (plan->GI >= (1 << 15) || plan->GJ >= (1 << 16)) return hipErrorInvalidValue;
    const int GK = (K + M - 1) / M;
    // units (I', J', kappa) holding at least one interior cell, by level I' + J' + kappa
    std::vector<std::vector<int2>> by_level(K > 0 ? plan->GI + plan->GJ + GK : 0);
    for (int kap = 0; kap < GK; ++kap) {
        const int k_lo = kap * M, k_hi = std::min(k_lo + M, K) - 1;
        for (int I = 0; I < plan->GI; ++I)
            for (int J = 0; J < plan->GJ; ++J) {
                bool any = false;
                for (int k = k_lo; k <= k_hi && !any; ++k)
                    any = GS_ROWS * I + 1 <= n + k && GS_ROWS * I + GS_ROWS >= 1 + k && TW * J + 1 <= n + k &&
                          TW * J + TW >= 1 + k;
                if (any) by_level[I + J + kap].push_back(make_int2((I << 16) | J, kap));
            }
    }
    // Within a level, units that meet the ring first: they are the slower ones
    // (k_gs_unit's edge loop), so in a level holding more units than the CUs run
    // at once they go in the first round rather than stretch the last one.
    auto meets_ring = [&](const int2& u) {
        const int I = u.x >> 16, J = u.x & 0xffff, k_lo = u.y * M, k_hi = std::min(k_lo + M, K) - 1;
        return !(GS_ROWS * I + 1 - k_hi >= 2 && GS_ROWS * I + GS_ROWS - k_lo <= n - 1 && TW * J + 1 - k_hi >= 2 &&
                 TW * J + TW - k_lo <= n - 1);
    };
    for (auto& units : by_level) std::stable_partition(units.begin(), units.end(), meets_ring);
    std::vector<int> offsets, level_of;
    std::vector<int2> work;
    for (size_t lv = 0; lv < by_level.size(); ++lv) {
        if (by_level[lv].empty()) continue;
        offsets.push_back(static_cast<int>(work.size()));
        level_of.push_back(static_cast<int>(lv));
        work.insert(work.end(), by_level[lv].begin(), by_level[lv].end());
    }
    plan->levels = static_cast<int>(offsets.size());
    offsets.push_back(static_cast<int>(work.size()));
    plan->host_offsets = new int[offsets.size()];
    std::copy(offsets.begin(), offsets.end(), plan->host_offsets);
    plan->level_of = new int[level_of.size() + 1];
    std::copy(level_of.begin(), level_of.end(), plan->level_of);
    plan->host_work = new int2[work.size() + 1];
    std::copy(work.begin(), work.end(), plan->host_work);
    if (work.empty() || !upload) return hipSuccess;
    hipError_t e = hipMalloc(&plan->work, sizeof(int2) * work.size());
    if (e != hipSuccess) return e;
    return hipMemcpy(plan->work, work.data(), sizeof(int2) * work.size(), hipMemcpyHostToDevice);
}

void gs_plan_destroy(GsPlan* plan) {
    if (plan->work) (void)hipFree(plan->work);
    delete[] plan->host_offsets;
    delete[] plan->level_of;
    delete[] plan->host_work;
    plan->work = nullptr;
    plan->host_offsets = nullptr;
    plan->level_of = nullptr;
    plan->host_work = nullptr;
}

// ---- the two-product quotient check ----------------------------------------
namespace {
// The significands m of [1, 2) for which x = 2^E m gives
// RN(x y + RN(x y_lo)) != RN(x / d) (every operation a float one).  Two
// copies of one loop: with the FMA instruction, and through the C library.
__attribute__((target("fma"))) long two_product_misses_hw(float d, float y, float yl, int E) {
    const float scale = std::ldexp(1.0f, E);
    long misses = 0;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        const uint32_t bits = 0x3f800000u | m;
        float x;
        std::memcpy(&x, &bits, 4);
        x *= scale;
        if (__builtin_fmaf(x, y, x * yl) != x / d) ++misses;
    }
    return misses;
}

long two_product_misses_soft(float d, float y, float yl, int E) {
    const float scale = std::ldexp(1.0f, E);
    long misses = 0;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        const uint32_t bits = 0x3f800000u | m;
        float x;
        std::memcpy(&x, &bits, 4);
        x *= scale;
        if (std::fma(x, y, x * yl) != x / d) ++misses;
    }
    return misses;
}
}  // namespace

void two_product_quotient(SweepArgs& A) {
    A.div_checked = 0;
    if (A.kind != SWEEP_DIFFUSE) return;
    const float d = A.denom;
    if (!(d > 1.f && d <= 0x1p100f)) return;
    // y2: RN(1/d), or the float below it when that is above 1/d; 1 - d y2 is
    // exact in double (d y2 has at most 48 significant bits) and positive
    // unless 1/d is a float (then the plain quotient takes it).
    float y = A.inv_denom;
    if (1.0 - static_cast<double>(d) * y < 0.0) y = std::nextafter(y, 0.f);
    const double rem = 1.0 - static_cast<double>(d) * y;
    if (!(rem > 0.0)) return;
    const float yl = static_cast<float>(rem / d);
    if (!(yl > 0.f)) return;
    // From lo up, RN(x y2_lo) and the quotient are normal, so every operation
    // scales exactly with the dividend's power of two, and one binade,
    // [2^E, 2^(E + 1)) with 2^E above lo's own floor, decides all of them.
    const double t = std::ldexp(1.0, -126) / static_cast<double>(yl);
    if (!(t <= 0x1p99)) return;
    float tf = static_cast<float>(t);
    if (static_cast<double>(tf) < t) tf = std::nextafter(tf, 0x1p127f);
    const float lo = std::max(A.div_lo, tf);
    const int E = std::max(0, std::ilogb(tf) + 1);
    static std::mutex mu;
    static std::unordered_map<uint64_t, bool> exact;
    uint32_t db, lb;
    std::memcpy(&db, &d, 4);
    std::memcpy(&lb, &yl, 4);
    const uint64_t key = (static_cast<uint64_t>(db) << 32) | lb;
    bool ok;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = exact.find(key);
        if (it != exact.end()) {
            ok = it->second;
        } else {
            const long misses = __builtin_cpu_supports("fma") ? two_product_misses_hw(d, y, yl, E)
                                                              : two_product_misses_soft(d, y, yl, E);
            ok = misses == 0;
            exact.emplace(key, ok);
        }
    }
    if (ok) {
        A.div_checked = 1;
        A.div_lo = lo;
        A.inv_hi2 = y;
        A.inv_lo2 = yl;
    }
}

// Whether fields [0, count) of a set all take the checked two-product quotient.
static bool all_checked(const SweepSet& set, int count) {
    for (int f = 0; f < count; ++f)
        if (!set.a[f].div_checked) return false;
    return count > 0;
}

// A copy of a set whose diffuse fields are marked by two_product_quotient.
static SweepSet with_checked_divisions(const SweepSet& set) {
    SweepSet out = set;
    for (int f = 0; f < std::min(out.count, SweepSet::kMax); ++f)
        if (out.a[f].kind == SWEEP_DIFFUSE) two_product_quotient(out.a[f]);
    return out;
}

template <int TW, int M, int L>
void launch_unit(hipStream_t s, int n, const int2* w, const SweepSet& set) {
    SweepSet level = set;
    level.units_per_field = n;
    const dim3 grid(n * set.count);
    if (set.a[0].kind != SWEEP_DIFFUSE)
        launch_pressure_units(s, TW, M, L, grid, level, w);
    else if (all_checked(set, set.count))
        hipLaunchKernelGGL((k_gs_unit<TW, M, L, SWEEP_DIFFUSE, true>), grid, dim3(64 * M), 0, s, level, w);
    else
        hipLaunchKernelGGL((k_gs_unit<TW, M, L, SWEEP_DIFFUSE, false>), grid, dim3(64 * M), 0, s, level, w);
}

template <int TW, int M>
bool launch_unit_lag(hipStream_t s, int lag, int n, const int2* w, const SweepSet& set) {
    switch (lag) {
        case 1: launch_unit<TW, M, 1>(s, n, w, set); return true;
        case 4: launch_unit<TW, M, 4>(s, n, w, set); return true;
        case 8: launch_unit<TW, M, 8>(s, n, w, set); return true;
        default: return false;
    }
}

template <int TW, int M, int L>
void launch_mixed(hipStream_t s, int n_a, const int2* wa, int n_b, const int2* wb, const SweepSet& set) {
    const dim3 grid((set.count - 1) * n_a + n_b);
    if (all_checked(set, set.count - 1))
        hipLaunchKernelGGL((k_gs_unit_mixed<TW, M, L, true>), grid, dim3(64 * M), 0, s, set, wa, n_a, wb);
    else
        hipLaunchKernelGGL((k_gs_unit_mixed<TW, M, L, false>), grid, dim3(64 * M), 0, s, set, wa, n_a, wb);
}

template <int TW, int M>
bool launch_mixed_lag(hipStream_t s, int lag, int n_a, const int2* wa, int n_b, const int2* wb, const SweepSet& set) {
    switch (lag) {
        case 1: launch_mixed<TW, M, 1>(s, n_a, wa, n_b, wb, set); return true;
        case 4: launch_mixed<TW, M, 4>(s, n_a, wa, n_b, wb, set); return true;
        case 8: launch_mixed<TW, M, 8>(s, n_a, wa, n_b, wb, set); return true;
        default: return false;
    }
}

bool launch_mixed_any(hipStream_t s, int TW, int M, int lag, int n_a, const int2* wa, int n_b, const int2* wb,
                      const SweepSet& set) {
    if (TW == 64) {
        switch (M) {
            case 2: return launch_mixed_lag<64, 2>(s, lag, n_a, wa, n_b, wb, set);
            case 4: return launch_mixed_lag<64, 4>(s, lag, n_a, wa, n_b, wb, set);
            case 8: return launch_mixed_lag<64, 8>(s, lag, n_a, wa, n_b, wb, set);
            default: return false;
        }
    }
    if (TW == 128) {
        switch (M) {
            case 4: return launch_mixed_lag<128, 4>(s, lag, n_a, wa, n_b, wb, set);
            case 8: return launch_mixed_lag<128, 8>(s, lag, n_a, wa, n_b, wb, set);
            default: return false;
        }
    }
    return false;
}

bool launch_unit_any(hipStream_t s, int TW, int M, int lag, int n, const int2* w, const SweepSet& set) {
    if (TW == 64) {
        switch (M) {
            case 2: return launch_unit_lag<64, 2>(s, lag, n, w, set);
            case 4: return launch_unit_lag<64, 4>(s, lag, n, w, set);
            case 8: return launch_unit_lag<64, 8>(s, lag, n, w, set);
            default: return false;
        }
    }
    if (TW == 128) {
        switch (M) {
            case 4: return launch_unit_lag<128, 4>(s, lag, n, w, set);
            case 8: return launch_unit_lag<128, 8>(s, lag, n, w, set);
            default: return false;
        }
    }
    return false;
}

hipError_t launch_gs_solve_set(hipStream_t s, const GsPlan& plan, const SweepSet& set_in) {
    if (set_in.count < 1 || set_in.count > SweepSet::kMax) return hipErrorInvalidValue;
    const SweepSet set = with_checked_divisions(set_in);
    for (int f = 0; f < set.count; ++f)
        if (plan.K != set.a[f].K || set.a[f].kind != set.a[0].kind || set.a[f].R != set.a[0].R)
            return hipErrorInvalidValue;
    if (plan.M == 0 && set.count != 1) return hipErrorInvalidValue;
    const SweepArgs& a = set.a[0];
    for (int L = 0; L < plan.levels; ++L) {
        const int n = plan.host_offsets[L + 1] - plan.host_offsets[L];
        if (n <= 0) continue;
        const int2* w = plan.work + plan.host_offsets[L];
        const bool dif = a.kind == SWEEP_DIFFUSE;
        if (plan.M == 0) {
            if (plan.TW == 64 && dif)
                hipLaunchKernelGGL((k_gs_tile<64, SWEEP_DIFFUSE>), dim3(n), dim3(64), 0, s, a, w);
            else if (plan.TW == 64)
                hipLaunchKernelGGL((k_gs_tile<64, SWEEP_PRESSURE>), dim3(n), dim3(64), 0, s, a, w);
            else if (plan.TW == 128 && dif)
                hipLaunchKernelGGL((k_gs_tile<128, SWEEP_DIFFUSE>), dim3(n), dim3(64), 0, s, a, w);
            else if (plan.TW == 128)
                hipLaunchKernelGGL((k_gs_tile<128, SWEEP_PRESSURE>), dim3(n), dim3(64), 0, s, a, w);
            else
                return hipErrorInvalidValue;
        } else {
            bool ok = false;
            if (plan.TW == 64) {
                switch (plan.M) {
                    case 2: ok = launch_unit_lag<64, 2>(s, plan.lag, n, w, set); break;
                    case 4: ok = launch_unit_lag<64, 4>(s, plan.lag, n, w, set); break;
                    case 8: ok = launch_unit_lag<64, 8>(s, plan.lag, n, w, set); break;
                    default: break;
                }
            } else if (plan.TW == 128) {
                switch (plan.M) {
                    case 4: ok = launch_unit_lag<128, 4>(s, plan.lag, n, w, set); break;
                    case 8: ok = launch_unit_lag<128, 8>(s, plan.lag, n, w, set); break;
                    default: break;
                }
            }
            if (!ok) return hipErrorInvalidValue;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// Units of level L of the implicit enumeration (unit_of without a list).
static int implicit_level_units(int GI, int GJ, int GK, int L) {
    int count = 0;
    for (int kap = std::max(0, L - (GI - 1) - (GJ - 1)); kap <= std::min(GK - 1, L); ++kap) {
        const int m = L - kap;
        count += std::min(GI - 1, m) - std::max(0, m - (GJ - 1)) + 1;
    }
    return count;
}

hipError_t launch_gs_solve_implicit(hipStream_t s, int R, int K, int TW, int M, int lag, const SweepArgs& a) {
    if (K <= 0) return hipSuccess;
    if (K > kImplicitMaxSweeps || (TW != 64 && TW != 128) || (M != 2 && M != 4 && M != 8) ||
        (TW == 128 && M == 2))
        return hipErrorInvalidValue;
    const int n = R - 2;
    SweepSet set{};
    set.a[0] = a;
    set.count = 1;
    set = with_checked_divisions(set);
    set.gi = (n + K - 1 + GS_ROWS - 1) / GS_ROWS;
    set.gj = (n + K - 1 + TW - 1) / TW;
    set.gk = (K + M - 1) / M;
    for (int L = 0; L <= set.gi + set.gj + set.gk - 3; ++L) {
        const int units = implicit_level_units(set.gi, set.gj, set.gk, L);
        if (units <= 0) continue;
        set.level = L;
        bool ok = false;
        if (TW == 64) {
            switch (M) {
                case 2: ok = launch_unit_lag<64, 2>(s, lag, units, nullptr, set); break;
                case 4: ok = launch_unit_lag<64, 4>(s, lag, units, nullptr, set); break;
                default: ok = launch_unit_lag<64, 8>(s, lag, units, nullptr, set); break;
            }
        } else {
            ok = (M == 4) ? launch_unit_lag<128, 4>(s, lag, units, nullptr, set)
                          : launch_unit_lag<128, 8>(s, lag, units, nullptr, set);
        }
        if (!ok) return hipErrorInvalidValue;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

int gs_pipeline_offset(const GsPlan& pp, int R, int Kd, int TWd, int Md) {
    // The diffuse's last writer of cell (i, j) -- interior, or a ring cell
    // through its adjacent interior cell -- is the unit relaxing sweep Kd - 1
    // there: skewed (i + Kd - 1, j + Kd - 1), so level
    // (i + Kd - 2) / 64 + (j + Kd - 2) / TWd + (Kd - 1) / Md, non-decreasing in
    // i and j.  A pressure unit reads the velocity one cell beyond the
    // divergence region it loads (rows up to 64 I' - k_lo + 63, columns up to
    // TW J' - k_lo + TW - 1, clamped into the interior); its last writer is at
    // that region's far corner.
    const int n = R - 2;
    int off = 0;
    for (int L = 0; L < pp.levels; ++L)
        for (int e = pp.host_offsets[L]; e < pp.host_offsets[L + 1]; ++e) {
            const int2 w = pp.host_work[e];
            const int I = w.x >> 16, J = w.x & 0xffff, k_lo = w.y * pp.M;
            const int ib = std::min(GS_ROWS * I - k_lo + GS_ROWS - 1, n) + 1;
            const int jb = std::min(pp.TW * J - k_lo + pp.TW - 1, n) + 1;
            const int ic = std::min(std::max(ib, 1), n), jc = std::min(std::max(jb, 1), n);
            const int last_writer = (ic + Kd - 2) / GS_ROWS + (jc + Kd - 2) / TWd + (Kd - 1) / Md;
            off = std::max(off, last_writer + 1 - pp.level_of[L]);
        }
    return off;
}

hipError_t launch_gs_pipelined(hipStream_t s, const GsPlan& pd, const SweepSet& diffuse_in, const GsPlan& pp,
                               const SweepArgs& p, int offset) {
    const SweepSet diffuse = with_checked_divisions(diffuse_in);
    if (pd.M == 0 || pp.M == 0 || pd.TW != pp.TW || pd.M != pp.M || pd.lag != pp.lag || !pd.level_of ||
        !pp.level_of || offset < 0 || pd.levels < 1 || pp.levels < 1)
        return hipErrorInvalidValue;
    if (diffuse.count < 1 || diffuse.count > SweepSet::kMax - 1 || p.kind != SWEEP_PRESSURE || p.K != pp.K)
        return hipErrorInvalidValue;
    for (int f = 0; f < diffuse.count; ++f)
        if (diffuse.a[f].kind != SWEEP_DIFFUSE || diffuse.a[f].K != pd.K || diffuse.a[f].R != p.R)
            return hipErrorInvalidValue;
    SweepSet mixed = diffuse;
    mixed.a[diffuse.count] = p;
    mixed.count = diffuse.count + 1;
    SweepSet pset{};
    pset.a[0] = p;
    pset.count = 1;
    const int last = std::max(pd.level_of[pd.levels - 1], pp.level_of[pp.levels - 1] + offset);
    int ld = 0, lp = 0;
    for (int t = 0; t <= last; ++t) {
        int na = 0, nb = 0;
        const int2 *wa = nullptr, *wb = nullptr;
        if (ld < pd.levels && pd.level_of[ld] == t) {
            na = pd.host_offsets[ld + 1] - pd.host_offsets[ld];
            wa = pd.work + pd.host_offsets[ld];
            ++ld;
        }
        if (lp < pp.levels && pp.level_of[lp] + offset == t) {
            nb = pp.host_offsets[lp + 1] - pp.host_offsets[lp];
            wb = pp.work + pp.host_offsets[lp];
            ++lp;
        }
        bool ok = true;
        if (na > 0 && nb > 0)
            ok = launch_mixed_any(s, pd.TW, pd.M, pd.lag, na, wa, nb, wb, mixed);
        else if (na > 0)
            ok = launch_unit_any(s, pd.TW, pd.M, pd.lag, na, wa, diffuse);
        else if (nb > 0)
            ok = launch_unit_any(s, pp.TW, pp.M, pp.lag, nb, wb, pset);
        if (!ok) return hipErrorInvalidValue;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return (ld == pd.levels && lp == pp.levels) ? hipSuccess : hipErrorInvalidValue;
}

hipError_t launch_gs_solve(hipStream_t s, const GsPlan& plan, const SweepArgs& a) {
    SweepSet set{};
    set.a[0] = a;
    set.count = 1;
    return launch_gs_solve_set(s, plan, set);
}

// ---- Jacobi ------------------------------------------------------------------
int jacobi_launches(int K, int block) { return (K + block - 1) / block; }

hipError_t launch_jacobi_pass(hipStream_t s, const SweepArgs& a, const float* src, float* dst, int zero_src, int m,
                              int kernel) {
    return kernel == JACOBI_LDS ? launch_jacobi_block(s, a, src, dst, zero_src, m)
                                : launch_jacobi_wave(s, a, src, dst, zero_src, m);
}

hipError_t launch_jacobi_solve(hipStream_t s, const SweepArgs& a, float* scratch, int block, int kernel) {
    // Sweep order: x -> scratch -> x -> ...; an odd launch count ends in
    // scratch, and the caller swaps the two buffers.
    const float* src = a.x;
    float* dst = scratch;
    int done = 0;
    bool zero = a.zero_start != 0;
    while (done < a.K) {
        const int m = std::min(block, a.K - done);
        hipError_t e = launch_jacobi_pass(s, a, src, dst, zero ? 1 : 0, m, kernel);
        if (e != hipSuccess) return e;
        zero = false;
        done += m;
        src = dst;
        dst = (dst == scratch) ? a.x : scratch;
    }
    return hipSuccess;
}

#endif  // FS2D_PRESSURE_UNITS_ONLY

}  // namespace fs2d
