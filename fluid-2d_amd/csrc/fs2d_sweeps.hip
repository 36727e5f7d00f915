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
//    sweep 0 reads the stored (stale) ring as the reference does.
//
//  * Jacobi -- every neighbour from the previous iterate.  M sweeps are fused
//    per launch (temporal blocking): a 256-thread block loads its tile plus an
//    M-cell halo into LDS once, relaxes M times in LDS on a shrinking region
//    (recomputing the ring in LDS after each sweep), and writes only the tile.
#include "fs2d_internal.hpp"

#include <algorithm>
#include <vector>

namespace fs2d {
namespace {

__device__ __forceinline__ float sgn(bool neg, float x) { return neg ? -x : x; }

// The update of one cell; operand order as in src/fluid_solver_cpu.cpp:107-108
// (diffuse: N + S + W + E) and :199-200 (pressure: d + E + W + S + N).
__device__ __forceinline__ float relax(int kind, float b, float n, float s, float w, float e, float a, float denom) {
    if (kind == SWEEP_DIFFUSE) return (b + a * (((n + s) + w) + e)) / denom;
    return ((((b + e) + w) + s) + n) / 4.0f;
}

// ===================== Gauss-Seidel tile hyperplane ======================
constexpr int GS_ROWS = 64;  // tile rows = wave lanes

template <int TW>
__global__ __launch_bounds__(64) void k_gs_tile(SweepArgs A, const int2* __restrict__ work) {
    // T: rows -1..64, cols -1..TW of the tile (row pitch TW + 2 is even, so the
    // skewed access (r, s - r) hits 64 distinct banks).  Bm: right-hand side.
    __shared__ float T[GS_ROWS + 2][TW + 2];
    __shared__ float Bm[GS_ROWS][TW];

    const int2 wk = work[blockIdx.x];
    const int I = wk.x >> 16, J = wk.x & 0xffff, k = wk.y;
    const int R = A.R, P = A.P;
    float* __restrict__ x = A.x;
    const int gi0 = 1 + GS_ROWS * I, gj0 = 1 + TW * J;
    const int rows = min(GS_ROWS, R - 1 - gi0), cols = min(TW, R - 1 - gj0);
    const int lane = threadIdx.x;
    const bool first = (k == 0);
    const bool zero_first = first && A.zero_start;

    // ---- load the sweep k-1 iterate with its halo (corners of the halo are
    // never read and are skipped).
    for (int r = -1; r <= rows; ++r) {
        const int gi = gi0 + r;
        const bool ring_i = (gi == 0 || gi == R - 1);
        for (int c = lane - 1; c <= cols; c += 64) {
            if ((r == -1 || r == rows) && (c == -1 || c == cols)) continue;
            const int gj = gj0 + c;
            const bool ring_j = (gj == 0 || gj == R - 1);
            float val;
            if (zero_first) {
                // Iterate starts at 0 (ring included); only the north and west
                // halos already hold sweep 0 of the neighbouring tiles.
                val = ((r == -1 && !ring_i) || (c == -1 && !ring_j)) ? x[(size_t)gi * P + gj] : 0.f;
            } else if (!first && ring_i) {
                val = sgn(A.ring.neg_tb, x[(size_t)(gi == 0 ? 1 : R - 2) * P + gj]);
            } else if (!first && ring_j) {
                val = sgn(A.ring.neg_lr, x[(size_t)gi * P + (gj == 0 ? 1 : R - 2)]);
            } else {
                val = x[(size_t)gi * P + gj];
            }
            T[r + 1][c + 1] = val;
        }
    }
    for (int r = 0; r < rows; ++r)
        for (int c = lane; c < cols; c += 64) Bm[r][c] = A.b[(size_t)(gi0 + r) * P + gj0 + c];
    __syncthreads();

    // ---- skewed sweep: at step s lane r relaxes column s - r.
    const int kind = A.kind;
    const float a = A.a, denom = A.denom;
    float xprev = 0.f;
    const int steps = rows + cols - 1;
    for (int s = 0; s < steps; ++s) {
        const int c = s - lane;
        // previous result of lane - 1 = cell (lane - 1, c) of this sweep
        const float from_up = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(xprev), 0x138, 0xf, 0xf, false));
        if (lane < rows && c >= 0 && c < cols) {
            const float n = (lane == 0) ? T[0][c + 1] : from_up;
            const float w = (c == 0) ? T[lane + 1][0] : xprev;
            const float so = T[lane + 2][c + 1];
            const float e = T[lane + 1][c + 2];
            const float v = relax(kind, Bm[lane][c], n, so, w, e, a, denom);
            T[lane + 1][c + 1] = v;
            xprev = v;
        }
    }
    __syncthreads();

    // ---- store the tile; the last sweep also writes the ring it borders.
    for (int r = 0; r < rows; ++r)
        for (int c = lane; c < cols; c += 64) x[(size_t)(gi0 + r) * P + gj0 + c] = T[r + 1][c + 1];
    if (k == A.K - 1) {
        const bool top = (gi0 == 1), bottom = (gi0 + rows == R - 1);
        const bool left = (gj0 == 1), right = (gj0 + cols == R - 1);
        const Ring rg = A.ring;
        if (top)
            for (int c = lane; c < cols; c += 64) x[gj0 + c] = sgn(rg.neg_tb, T[1][c + 1]);
        if (bottom)
            for (int c = lane; c < cols; c += 64) x[(size_t)(R - 1) * P + gj0 + c] = sgn(rg.neg_tb, T[rows][c + 1]);
        if (left)
            for (int r = lane; r < rows; r += 64) x[(size_t)(gi0 + r) * P] = sgn(rg.neg_lr, T[r + 1][1]);
        if (right)
            for (int r = lane; r < rows; r += 64) x[(size_t)(gi0 + r) * P + R - 1] = sgn(rg.neg_lr, T[r + 1][cols]);
        if (lane == 0) {
            auto corner = [&](float v) { return 0.5f * (sgn(rg.neg_tb, v) + sgn(rg.neg_lr, v)); };
            if (top && left) x[0] = corner(T[1][1]);
            if (top && right) x[R - 1] = corner(T[1][cols]);
            if (bottom && left) x[(size_t)(R - 1) * P] = corner(T[rows][1]);
            if (bottom && right) x[(size_t)(R - 1) * P + R - 1] = corner(T[rows][cols]);
        }
    }
}

// ===================== Jacobi, M sweeps per launch =======================
constexpr int JX = 64, JY = 32, JT = 256;

template <int M>
__global__ __launch_bounds__(JT) void k_jacobi_block(SweepArgs A, const float* __restrict__ src,
                                                    float* __restrict__ dst, int zero_src) {
    constexpr int W = JX + 2 * M, H = JY + 2 * M;          // iterate region (tile + M halo)
    constexpr int BW = JX + 2 * (M - 1), BH = JY + 2 * (M - 1);  // right-hand side region
    __shared__ float buf[2][H][W];
    __shared__ float rhs[BH][BW];

    const int R = A.R, P = A.P;
    const int gi0 = 1 + JY * blockIdx.y - M, gj0 = 1 + JX * blockIdx.x - M;  // global of buf(0, 0)
    const int tid = threadIdx.x;
    const Ring rg = A.ring;

    for (int idx = tid; idx < H * W; idx += JT) {
        const int r = idx / W, c = idx % W;
        const int gi = gi0 + r, gj = gj0 + c;
        float val = 0.f;
        if (!zero_src && gi >= 0 && gi < R && gj >= 0 && gj < R) val = src[(size_t)gi * P + gj];
        buf[0][r][c] = val;
    }
    for (int idx = tid; idx < BH * BW; idx += JT) {
        const int r = idx / BW, c = idx % BW;
        const int gi = gi0 + 1 + r, gj = gj0 + 1 + c;
        rhs[r][c] = (gi >= 1 && gi <= R - 2 && gj >= 1 && gj <= R - 2) ? A.b[(size_t)gi * P + gj] : 0.f;
    }
    __syncthreads();

    const int kind = A.kind;
    const float a = A.a, denom = A.denom;
#pragma unroll 1
    for (int s = 1; s <= M; ++s) {
        const int e = M - s;  // the region relaxed by sweep s is the tile grown by e
        const int rw = JX + 2 * e, rh = JY + 2 * e;
        const float(*cur)[W] = buf[(s - 1) & 1];
        float(*nxt)[W] = buf[s & 1];
        const bool last = (s == M);
        for (int idx = tid; idx < rh * rw; idx += JT) {
            const int r = M - e + idx / rw, c = M - e + idx % rw;  // buf coordinates
            const int gi = gi0 + r, gj = gj0 + c;
            if (gi < 1 || gi > R - 2 || gj < 1 || gj > R - 2) continue;
            const float v = relax(kind, rhs[r - 1][c - 1], cur[r - 1][c], cur[r + 1][c], cur[r][c - 1],
                                  cur[r][c + 1], a, denom);
            const float tb = sgn(rg.neg_tb, v), lr = sgn(rg.neg_lr, v);
            const float cn = 0.5f * (tb + lr);
            if (!last) {
                nxt[r][c] = v;
                // ring cells of this sweep (their inner neighbour is in the region)
                if (gi == 1) nxt[r - 1][c] = tb;
                if (gi == R - 2) nxt[r + 1][c] = tb;
                if (gj == 1) nxt[r][c - 1] = lr;
                if (gj == R - 2) nxt[r][c + 1] = lr;
                if (gi == 1 && gj == 1) nxt[r - 1][c - 1] = cn;
                if (gi == 1 && gj == R - 2) nxt[r - 1][c + 1] = cn;
                if (gi == R - 2 && gj == 1) nxt[r + 1][c - 1] = cn;
                if (gi == R - 2 && gj == R - 2) nxt[r + 1][c + 1] = cn;
            } else {
                const size_t o = (size_t)gi * P + gj;
                dst[o] = v;
                if (gi == 1) dst[o - P] = tb;
                if (gi == R - 2) dst[o + P] = tb;
                if (gj == 1) dst[o - 1] = lr;
                if (gj == R - 2) dst[o + 1] = lr;
                if (gi == 1 && gj == 1) dst[o - P - 1] = cn;
                if (gi == 1 && gj == R - 2) dst[o - P + 1] = cn;
                if (gi == R - 2 && gj == 1) dst[o + P - 1] = cn;
                if (gi == R - 2 && gj == R - 2) dst[o + P + 1] = cn;
            }
        }
        __syncthreads();
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

// ---- Gauss-Seidel plan ----------------------------------------------------
hipError_t gs_plan_create(GsPlan* plan, int R, int K, int TW) {
    plan->TW = TW;
    plan->GI = (R - 2 + GS_ROWS - 1) / GS_ROWS;
    plan->GJ = (R - 2 + TW - 1) / TW;
    plan->K = K;
    plan->levels = (K > 0) ? (plan->GI - 1) + (plan->GJ - 1) + 2 * (K - 1) + 1 : 0;
    plan->work = nullptr;
    plan->host_offsets = nullptr;
    if (plan->GI >= (1 << 15) || plan->GJ >= (1 << 16)) return hipErrorInvalidValue;
    std::vector<int> count(plan->levels + 1, 0);
    for (int k = 0; k < K; ++k)
        for (int I = 0; I < plan->GI; ++I)
            for (int J = 0; J < plan->GJ; ++J) ++count[I + J + 2 * k];
    plan->host_offsets = new int[plan->levels + 1];
    int acc = 0;
    for (int L = 0; L < plan->levels; ++L) {
        plan->host_offsets[L] = acc;
        acc += count[L];
    }
    plan->host_offsets[plan->levels] = acc;
    if (acc == 0) return hipSuccess;
    std::vector<int2> work(acc);
    std::vector<int> fill(plan->host_offsets, plan->host_offsets + plan->levels);
    for (int k = 0; k < K; ++k)
        for (int I = 0; I < plan->GI; ++I)
            for (int J = 0; J < plan->GJ; ++J) work[fill[I + J + 2 * k]++] = make_int2((I << 16) | J, k);
    hipError_t e = hipMalloc(&plan->work, sizeof(int2) * acc);
    if (e != hipSuccess) return e;
    return hipMemcpy(plan->work, work.data(), sizeof(int2) * acc, hipMemcpyHostToDevice);
}

void gs_plan_destroy(GsPlan* plan) {
    if (plan->work) (void)hipFree(plan->work);
    delete[] plan->host_offsets;
    plan->work = nullptr;
    plan->host_offsets = nullptr;
}

hipError_t launch_gs_solve(hipStream_t s, const GsPlan& plan, const SweepArgs& a) {
    if (plan.K != a.K) return hipErrorInvalidValue;
    for (int L = 0; L < plan.levels; ++L) {
        const int n = plan.host_offsets[L + 1] - plan.host_offsets[L];
        const int2* w = plan.work + plan.host_offsets[L];
        if (plan.TW == 64)
            hipLaunchKernelGGL(k_gs_tile<64>, dim3(n), dim3(64), 0, s, a, w);
        else if (plan.TW == 128)
            hipLaunchKernelGGL(k_gs_tile<128>, dim3(n), dim3(64), 0, s, a, w);
        else
            return hipErrorInvalidValue;
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// ---- Jacobi ------------------------------------------------------------------
int jacobi_launches(int K, int block) { return (K + block - 1) / block; }

hipError_t launch_jacobi_solve(hipStream_t s, const SweepArgs& a, float* scratch, int block) {
    // Sweep order: x -> scratch -> x -> ...; an odd launch count ends in
    // scratch, and the caller swaps the two buffers (see the API layer).
    const float* src = a.x;
    float* dst = scratch;
    int done = 0;
    bool zero = a.zero_start != 0;
    while (done < a.K) {
        const int m = std::min(block, a.K - done);
        hipError_t e = launch_jacobi_block(s, a, src, dst, zero ? 1 : 0, m);
        if (e != hipSuccess) return e;
        zero = false;
        done += m;
        src = dst;
        dst = (dst == scratch) ? a.x : scratch;
    }
    return hipSuccess;
}

}  // namespace fs2d
