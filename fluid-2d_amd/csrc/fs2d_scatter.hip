// fs2d_scatter.hip -- density advection, src/fluid_solver_cpu.cpp:127-152.
//
// The reference zero-fills the grid and then, for every interior source cell
// (i, j) in row-major order, adds the four bilinear shares of rho0(i, j) to
// the cells around (i + dt0 v, j + dt0 u); a source whose position leaves
// [0.5, R - 1.5] is skipped.  Every destination therefore receives its shares
// in source row-major order, summed from +0, and that order fixes the float
// result.  Two paths reproduce it with no atomics, chosen on the device from
// the largest speed (the radius W every source stays within, +1):
//
//  * W <= kWindow: k_scatter_window.  A wave streams down a 64-column strip,
//    evaluates each source once (target cell and its four shares), passes it
//    to the destinations within W columns by lane shifts, and each
//    destination adds the shares of a source row in column order to
//    accumulators for the 2W + 1 rows it reaches.
//  * W > kWindow: the sources are sorted by target cell with a stable LSD
//    radix sort of 9-bit digits (k_sort_keys, k_sort_hist, k_sort_scan,
//    k_sort_scatter), so each target's sources stay in row-major order; each
//    destination then merges, in source order, the four target cells whose
//    shares reach it (k_sort_bounds, k_scatter_merge).  Work is O(R^2)
//    whatever the speeds.
//
// Both paths' kernels are launched every step and return at once when the
// other path is active, so a captured step graph serves every velocity.
#include "fs2d_cells.hpp"
#include "fs2d_internal.hpp"

namespace fs2d {
namespace {

constexpr int kWindow = 4;  // largest radius W of the window path
constexpr int NT = 256;      // threads per block, every kernel here

// The radius of the scatter: a source moves at most dt0 * max|vel| + 2^-8
// cells, and its shares land on trunc(pos) and trunc(pos) + 1, so every share
// stays within W = floor(that) + 1 rows and columns of its source.
__device__ __forceinline__ int scatter_radius(const unsigned* max_bits, float dt0, int R) {
    const float reach = dt0 * __uint_as_float(*max_bits) + 0x1p-8f;
    return (reach < (float)R) ? (int)reach + 1 : R;
}

// Source (si, sj): its position (the reference's expressions and order,
// src/fluid_solver_cpu.cpp:129-144); false if it is skipped (or not finite).
__device__ __forceinline__ bool source_lands(int si, int sj, float uu, float vv, float dt0, float hi, int& i0, int& j0,
                                             float& s0, float& s2) {
    const float x = (float)sj + dt0 * uu;
    const float y = (float)si + dt0 * vv;
    if (x < 0.5f || x > hi || y < 0.5f || y > hi) return false;
    if (x != x || y != y) return false;
    j0 = (int)x;
    i0 = (int)y;
    s0 = x - (float)j0;
    s2 = y - (float)i0;
    return true;
}

// ---- window path -------------------------------------------------------------
// A wave owns a strip of 64 - 2W destination columns plus W source columns
// either side (lane = column; strips overlap by 2W columns) and a band of
// destination rows, and walks the band's source rows top to bottom (the
// band's rows plus W either side).  Each lane evaluates its own source of the
// row once -- target cell and its four shares -- and the sources up to W
// columns either side reach it by full-wave DPP shifts.  A destination keeps
// 2W + 1 accumulators, one per row the current source row reaches; a source
// row adds its shares to them in column order, so every destination sums its
// shares in source row-major order, from +0.  Where a source misses a
// destination, +0 is added: an accumulator starts at +0 and a round-to-nearest
// sum is -0 only when both operands are, so it is never -0 and adding +0
// leaves it unchanged.  When the walk passes row d + W, destination row d is
// complete and is written.
constexpr int kStripWaves = 4;  // waves per block (independent strips)

struct Shares {  // one source as the destinations see it
    int t;       // (i0 - si + 128) << 8 | (j0 - sj + 128), or -1 when skipped
    float a, b, c, d;  // shares to (i0, j0), (i0 + 1, j0), (i0, j0 + 1), (i0 + 1, j0 + 1); 0 when skipped
};

__device__ __forceinline__ Shares eval_shares(bool in, int si, int sj, float uu, float vv, float r, float dt0,
                                              float hi) {
    Shares s{-1, 0.f, 0.f, 0.f, 0.f};
    int i0, j0;
    float s0, s2;
    if (in && source_lands(si, sj, uu, vv, dt0, hi, i0, j0, s0, s2)) {
        const float s1 = 1.0f - s0, s3 = 1.0f - s2;  // src/fluid_solver_cpu.cpp:142-150
        s.t = ((i0 - si + 128) << 8) | (j0 - sj + 128);
        s.a = s1 * s3 * r;
        s.b = s1 * s2 * r;
        s.c = s0 * s3 * r;
        s.d = s0 * s2 * r;
    }
    return s;
}

// Lane l takes lane l - 1's value (from_west) or lane l + 1's (from_east); the
// lane without a neighbour takes 0 (a strip's edge lanes are never written).
__device__ __forceinline__ float from_west(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_east(float x) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x130, 0xf, 0xf, false));
}
template <bool WEST>
__device__ __forceinline__ Shares shift_shares(const Shares& x) {
    auto mv = [](float f) { return WEST ? from_west(f) : from_east(f); };
    Shares y;
    y.t = __float_as_int(mv(__int_as_float(x.t)));
    y.a = mv(x.a);
    y.b = mv(x.b);
    y.c = mv(x.c);
    y.d = mv(x.d);
    return y;
}

// Layout of one launch for radius W: strips of 64 - 2W destination columns,
// bands of `band` rows, one wave per (band, strip); `band` is the smallest
// (at least 16) that keeps every wave resident at once (`slots` waves).
struct WindowLayout {
    int strips, band, waves;
};
__host__ __device__ inline WindowLayout window_layout(int R, int W, int slots) {
    const int n = R - 2, sw = 64 - 2 * W;
    const int strips = (n + sw - 1) / sw;
    const int maxb = slots / strips > 1 ? slots / strips : 1;
    int band = (n + maxb - 1) / maxb;
    band = band < 16 ? 16 : band;
    band = band > n ? n : band;
    return WindowLayout{strips, band, strips * ((n + band - 1) / band)};
}

struct Row {
    float u, v, r;
};

// Radius 1: a source's target is (si - 1 or si, sj - 1 or sj), so it reaches
// the destinations of columns sj - 1 .. sj + 1 and rows si - 1 .. si + 1.  The
// source places its shares per destination column and row before the shifts,
// and a destination only adds.
__device__ __forceinline__ void window_band_r1(float* __restrict__ out, const float* __restrict__ rho0,
                                               const float* __restrict__ u, const float* __restrict__ v, int R,
                                               int P, float dt0, int col, int r0, int r1, bool dest) {
    const float hi = (float)R - 1.5f;
    const bool col_in = col >= 1 && col <= R - 2;
    const int s_first = r0 - 1, s_last = r1;  // destination rows [r0, r1)
    auto fetch = [&](int si) {
        Row x{0.f, 0.f, 0.f};
        if (col_in && si >= 1 && si <= R - 2 && si <= s_last) {
            const size_t o = (size_t)si * P + col;
            x.u = u[o];
            x.v = v[o];
            x.r = rho0[o];
        }
        return x;
    };
    Row cur = fetch(s_first), nx1 = fetch(s_first + 1);  // loads run two rows ahead
    float acc0 = 0.f, acc1 = 0.f, acc2 = 0.f;              // rows si - 1, si, si + 1
    for (int si = s_first; si <= s_last; ++si) {
        const Row nx2 = fetch(si + 2);
        const Shares m = eval_shares(col_in && si >= 1 && si <= R - 2, si, col, cur.u, cur.v, cur.r, dt0, hi);
        const bool west = (m.t & 255) < 128;  // j0 = sj - 1 (a skipped source has t = -1: false)
        const bool same = (m.t >> 8) == 128;  // i0 = si
        // per destination column: its (i0-row, i1-row) shares
        const float wl = west ? m.a : 0.f, wu = west ? m.b : 0.f;  // column sj - 1
        const float cl = west ? m.c : m.a, cu = west ? m.d : m.b;  // column sj
        const float el = west ? 0.f : m.c, eu = west ? 0.f : m.d;  // column sj + 1
        // placed on rows si - 1, si, si + 1
        const float w0 = same ? 0.f : wl, w1 = same ? wl : wu, w2 = same ? wu : 0.f;
        const float c0 = same ? 0.f : cl, c1 = same ? cl : cu, c2 = same ? cu : 0.f;
        const float e0 = same ? 0.f : el, e1 = same ? el : eu, e2 = same ? eu : 0.f;
        // column order: the west source (its east shares), this one, the east
        // source (its west shares)
        acc0 = acc0 + from_west(e0);
        acc1 = acc1 + from_west(e1);
        acc2 = acc2 + from_west(e2);
        acc0 = acc0 + c0;
        acc1 = acc1 + c1;
        acc2 = acc2 + c2;
        acc0 = acc0 + from_east(w0);
        acc1 = acc1 + from_east(w1);
        acc2 = acc2 + from_east(w2);
        const int di = si - 1;  // complete now
        if (dest && di >= r0) {
            out[(size_t)di * P + col] = acc0;
            cells::write_ring(out, di, col, acc0, R, P, Ring{false, false});
        }
        acc0 = acc1;
        acc1 = acc2;
        acc2 = 0.f;
        cur = nx1;
        nx1 = nx2;
    }
}

template <int W>
__device__ __forceinline__ void window_band(float* __restrict__ out, const float* __restrict__ rho0,
                                            const float* __restrict__ u, const float* __restrict__ v, int R, int P,
                                            float dt0, int col, int r0, int r1, bool dest) {
    constexpr int NA = 2 * W + 1;
    const float hi = (float)R - 1.5f;
    const bool col_in = col >= 1 && col <= R - 2;
    const int s_first = r0 - W, s_last = r1 - 1 + W;  // destination rows [r0, r1)
    auto fetch = [&](int si) {
        Row x{0.f, 0.f, 0.f};
        if (col_in && si >= 1 && si <= R - 2 && si <= s_last) {
            const size_t o = (size_t)si * P + col;
            x.u = u[o];
            x.v = v[o];
            x.r = rho0[o];
        }
        return x;
    };
    Row cur = fetch(s_first), nx1 = fetch(s_first + 1);  // loads run two rows ahead
    float acc[NA];  // acc[x]: row si - W + x
#pragma unroll
    for (int x = 0; x < NA; ++x) acc[x] = 0.f;

    for (int si = s_first; si <= s_last; ++si) {
        const Row nx2 = fetch(si + 2);
        const Shares m = eval_shares(col_in && si >= 1 && si <= R - 2, si, col, cur.u, cur.v, cur.r, dt0, hi);
        // the source e columns from this lane's (e = -W .. W, in column order)
        auto merge = [&](const Shares& s, int e) {
            const int dI = (s.t >> 8) - 128, dJ = (s.t & 255) - 128;
            const int cj = -e - dJ;  // this column - j0: 0 takes the j0 shares, 1 the j1 shares
            const float lo = cj == 0 ? s.a : (cj == 1 ? s.c : 0.f);
            const float up = cj == 0 ? s.b : (cj == 1 ? s.d : 0.f);
            const int xl = dI + W;
#pragma unroll
            for (int x = 0; x < NA; ++x) acc[x] = acc[x] + (x == xl ? lo : (x == xl + 1 ? up : 0.f));
        };
        Shares west[W];  // west[k - 1]: k columns west
        west[0] = shift_shares<true>(m);
#pragma unroll
        for (int k = 2; k <= W; ++k) west[k - 1] = shift_shares<true>(west[k - 2]);
#pragma unroll
        for (int k = W; k >= 1; --k) merge(west[k - 1], -k);
        merge(m, 0);
        Shares east = m;
#pragma unroll
        for (int k = 1; k <= W; ++k) {
            east = shift_shares<false>(east);
            merge(east, k);
        }
        const int di = si - W;  // complete now
        if (dest && di >= r0 && di <= R - 2) {
            out[(size_t)di * P + col] = acc[0];
            cells::write_ring(out, di, col, acc[0], R, P, Ring{false, false});
        }
#pragma unroll
        for (int x = 0; x + 1 < NA; ++x) acc[x] = acc[x + 1];
        acc[NA - 1] = 0.f;
        cur = nx1;
        nx1 = nx2;
    }
}

// One wave per (band, strip), strip fastest; a block holds kStripWaves
// consecutive waves, and block b runs block-item (b mod 8) * per + b / 8, so
// each XCD (blocks are dealt to the 8 XCDs in turn) walks a contiguous run of
// strips and bands and shares their overlapping lines in its L2.
__global__ __launch_bounds__(NT, 6) void k_scatter_window(float* __restrict__ out, const float* __restrict__ rho0,
                                                          const float* __restrict__ u, const float* __restrict__ v,
                                                          int R, int P, float dt0,
                                                          const unsigned* __restrict__ max_bits, int slots, int per) {
    const int W = scatter_radius(max_bits, dt0, R);
    if (W > kWindow) return;
    const WindowLayout L = window_layout(R, W, slots);
    const int item = ((int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3)) * kStripWaves + (int)(threadIdx.x >> 6);
    if (item >= L.waves) return;  // whole waves leave; no block-wide barrier follows
    const int lane = threadIdx.x & 63;
    const int bi = item / L.strips, strip = item - bi * L.strips;
    const int col = 1 + strip * (64 - 2 * W) - W + lane;
    const bool dest = lane >= W && lane < 64 - W && col <= R - 2;
    const int r0 = 1 + L.band * bi, r1 = min(r0 + L.band, R - 1);
    switch (W) {
        case 1: window_band_r1(out, rho0, u, v, R, P, dt0, col, r0, r1, dest); break;
        case 2: window_band<2>(out, rho0, u, v, R, P, dt0, col, r0, r1, dest); break;
        case 3: window_band<3>(out, rho0, u, v, R, P, dt0, col, r0, r1, dest); break;
        default: window_band<4>(out, rho0, u, v, R, P, dt0, col, r0, r1, dest); break;
    }
}

// Waves of k_scatter_window resident at once on the current device (queried
// once per device).
int window_slots() {
    constexpr int kMaxDev = 64;
    static int slots[kMaxDev] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDev) dev = -1;
    if (dev >= 0 && slots[dev] > 0) return slots[dev];
    int cus = 0, blocks = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev < 0 ? 0 : dev) != hipSuccess || cus < 1)
        cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_scatter_window, NT, 0) != hipSuccess || blocks < 1)
        blocks = 1;
    const int n = cus * blocks * kStripWaves;
    if (dev >= 0) slots[dev] = n;
    return n;
}

// ---- sorted path ---------------------------------------------------------------
// The sorted path's kernels are launched every step and leave at once unless
// the gate says some source may move past the window path's radius: on one
// grid from its largest speed, in a batch from a flag set per step.
struct GridGate {
    const unsigned* max_bits;
    float dt0;
    int R;
    __device__ bool sorted() const { return scatter_radius(max_bits, dt0, R) > kWindow; }
};
struct FlagGate {
    const int* flag;
    __device__ bool sorted() const { return *flag != 0; }
};

// Items are 64-bit: target cell i0 * (R - 1) + j0 in the high word (or
// (R - 1)^2 for a skipped source, sorted last), the source's interior
// row-major index in the low word.  The sort takes 9-bit digits over chunks of
// CH items; every kernel walks its work with a grid of at most kSortBlocks
// blocks, so that while the window path is active each returns at once.
constexpr int IPT = 16, CH = NT * IPT;
constexpr int DB = 9, ND = 1 << DB;  // digit bits, digits
constexpr int kSortBlocks = 512;

// Items of every chunk, the chunk counts of the first digit, and the target
// ranges cleared.
__global__ __launch_bounds__(NT) void k_sort_keys(unsigned long long* __restrict__ items, unsigned* __restrict__ hist,
                                                  int2* __restrict__ bounds, const float* __restrict__ u,
                                                  const float* __restrict__ v, int R, int P, float dt0, int chunks,
                                                  const unsigned* __restrict__ max_bits) {
    if (scatter_radius(max_bits, dt0, R) <= kWindow) return;
    __shared__ unsigned cnt[ND];
    const int n = R - 2;
    const long long N = (long long)n * n, nb = (long long)(R - 1) * (R - 1);
    for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < nb; i += (long long)gridDim.x * NT)
        bounds[i] = make_int2(0, 0);
    const float hi = (float)R - 1.5f;
    for (int c = blockIdx.x; c < chunks; c += gridDim.x) {
        for (int d = threadIdx.x; d < ND; d += NT) cnt[d] = 0u;
        __syncthreads();
        for (int q = 0; q < IPT; ++q) {
            const long long s = (long long)c * CH + q * NT + threadIdx.x;
            if (s >= N) break;
            const int si = 1 + (int)(s / n), sj = 1 + (int)(s % n);
            const size_t o = (size_t)si * P + sj;
            int i0, j0;
            float s0, s2;
            unsigned key = (unsigned)nb;
            if (source_lands(si, sj, u[o], v[o], dt0, hi, i0, j0, s0, s2)) key = (unsigned)i0 * (R - 1) + j0;
            items[s] = ((unsigned long long)key << 32) | (unsigned long long)s;
            atomicAdd(&cnt[key & (ND - 1)], 1u);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < ND; d += NT) hist[d * chunks + c] = cnt[d];
        __syncthreads();
    }
}

// Chunk counts of digit (key >> shift) & (ND - 1): hist[digit * chunks + chunk].
template <class G>
__global__ __launch_bounds__(NT) void k_sort_hist(const unsigned long long* __restrict__ items, long long N, int shift,
                                                  unsigned* __restrict__ hist, int chunks, G gate) {
    if (!gate.sorted()) return;
    __shared__ unsigned cnt[ND];
    for (int c = blockIdx.x; c < chunks; c += gridDim.x) {
        for (int d = threadIdx.x; d < ND; d += NT) cnt[d] = 0u;
        __syncthreads();
        for (int q = 0; q < IPT; ++q) {
            const long long i = (long long)c * CH + q * NT + threadIdx.x;
            if (i < N) atomicAdd(&cnt[(unsigned)(items[i] >> (32 + shift)) & (ND - 1)], 1u);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < ND; d += NT) hist[d * chunks + c] = cnt[d];
        __syncthreads();
    }
}

// Exclusive scan of hist (digit-major) in place, one block.
template <class G>
__global__ __launch_bounds__(1024) void k_sort_scan(unsigned* __restrict__ hist, int total, G gate) {
    if (!gate.sorted()) return;
    __shared__ unsigned part[1024];
    const int t = threadIdx.x;
    const int per = (total + 1023) / 1024;
    const int lo = min(t * per, total), hi = min(lo + per, total);
    unsigned sum = 0;
    for (int k = lo; k < hi; ++k) sum += hist[k];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // inclusive scan of the partial sums
        const unsigned add = t >= off ? part[t - off] : 0u;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    unsigned run = part[t] - sum;
    for (int k = lo; k < hi; ++k) {
        const unsigned c = hist[k];
        hist[k] = run;
        run += c;
    }
}

// Stable scatter of each chunk by digit: items of a chunk keep their order
// within each digit (ranks by pass, wave, then lane).
template <class G>
__global__ __launch_bounds__(NT) void k_sort_scatter(const unsigned long long* __restrict__ in,
                                                     unsigned long long* __restrict__ out, long long N, int shift,
                                                     const unsigned* __restrict__ offs, int chunks, G gate) {
    if (!gate.sorted()) return;
    __shared__ unsigned run[ND];           // items of each digit placed by earlier passes
    __shared__ unsigned wcnt[NT / 64][ND];  // items of each digit in each wave of this pass
    const int t = threadIdx.x, w = t >> 6, lane = t & 63;
    const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));  // lanes below this one
    for (int c = blockIdx.x; c < chunks; c += gridDim.x) {
        for (int d = t; d < ND; d += NT) run[d] = offs[d * chunks + c];
        for (int q = 0; q < IPT; ++q) {
            for (int d = t; d < ND; d += NT)
                for (int k = 0; k < NT / 64; ++k) wcnt[k][d] = 0u;
            __syncthreads();
            const long long i = (long long)c * CH + q * NT + t;
            const bool valid = i < N;
            const unsigned long long item = valid ? in[i] : 0ull;
            const unsigned dg = (unsigned)(item >> (32 + shift)) & (ND - 1);
            // lanes of this wave holding the same digit
            unsigned long long peers = __ballot(valid);
            for (int b = 0; b < DB; ++b) {
                const unsigned long long bal = __ballot((dg >> b) & 1u);
                peers &= ((dg >> b) & 1u) ? bal : ~bal;
            }
            const unsigned below = (unsigned)__popcll(peers & lt);
            if (valid && below == 0) wcnt[w][dg] = (unsigned)__popcll(peers);
            __syncthreads();
            if (valid) {
                unsigned pos = run[dg] + below;
                for (int k = 0; k < w; ++k) pos += wcnt[k][dg];
                if (pos < N) out[pos] = item;
            }
            __syncthreads();
            for (int d = t; d < ND; d += NT) {
                unsigned add = 0;
                for (int k = 0; k < NT / 64; ++k) add += wcnt[k][d];
                run[d] += add;
            }
        }
        __syncthreads();
    }
}

// For each target cell, the range [first, last) of its items after the sort.
template <class G>
__global__ __launch_bounds__(NT) void k_sort_bounds(const unsigned long long* __restrict__ items, long long N,
                                                    int2* __restrict__ bounds, unsigned nb, G gate) {
    if (!gate.sorted()) return;
    for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < N; i += (long long)gridDim.x * NT) {
        const unsigned key = (unsigned)(items[i] >> 32);
        if (key >= nb) continue;
        if (i == 0 || (unsigned)(items[i - 1] >> 32) != key) bounds[key].x = (int)i;
        if (i == N - 1 || (unsigned)(items[i + 1] >> 32) != key) bounds[key].y = (int)i + 1;
    }
}

// Each destination merges, in source order, the sources of the four target
// cells whose shares reach it: (di, dj) takes s1 s3 from target (di, dj),
// s1 s2 from (di - 1, dj), s0 s3 from (di, dj - 1), s0 s2 from (di - 1, dj - 1).
__global__ __launch_bounds__(NT) void k_scatter_merge(float* __restrict__ out, const float* __restrict__ rho0,
                                                      const float* __restrict__ u, const float* __restrict__ v,
                                                      const unsigned long long* __restrict__ items,
                                                      const int2* __restrict__ bounds, int R, int P, float dt0,
                                                      const unsigned* __restrict__ max_bits) {
    if (scatter_radius(max_bits, dt0, R) <= kWindow) return;
    const int n = R - 2;
    const float hi = (float)R - 1.5f;
    for (long long c = (long long)blockIdx.x * NT + threadIdx.x; c < (long long)n * n; c += (long long)gridDim.x * NT) {
        const int di = 1 + (int)(c / n), dj = 1 + (int)(c % n);
        int cur[4], end[4];
        const int ti[4] = {di, di - 1, di, di - 1}, tj[4] = {dj, dj, dj - 1, dj - 1};
        for (int k = 0; k < 4; ++k) {
            const int2 b = bounds[(size_t)ti[k] * (R - 1) + tj[k]];
            cur[k] = b.x;
            end[k] = b.y;
        }
        float acc = 0.f;
        for (;;) {
            int best = -1;
            unsigned bs = 0xffffffffu;
            for (int k = 0; k < 4; ++k) {
                if (cur[k] >= end[k]) continue;
                const unsigned s = (unsigned)items[cur[k]];
                if (s < bs) {
                    bs = s;
                    best = k;
                }
            }
            if (best < 0) break;
            ++cur[best];
            const int si = 1 + (int)(bs / (unsigned)n), sj = 1 + (int)(bs % (unsigned)n);
            const size_t o = (size_t)si * P + sj;
            int i0, j0;
            float s0, s2;
            source_lands(si, sj, u[o], v[o], dt0, hi, i0, j0, s0, s2);
            const float s1 = 1.0f - s0, s3 = 1.0f - s2;
            const float r0 = rho0[o];
            const float sh = (best == 0) ? s1 * s3 * r0 : (best == 1) ? s1 * s2 * r0 : (best == 2) ? s0 * s3 * r0 : s0 * s2 * r0;
            acc = acc + sh;
        }
        out[(size_t)di * P + dj] = acc;
        cells::write_ring(out, di, dj, acc, R, P, Ring{false, false});
    }
}

// ---- a batch of simulations -------------------------------------------------------
// Window path: one wave per (simulation, band, strip) of a fixed work list,
// strips of kBatchStrip destination columns with kWindow source columns
// either side (so every radius up to kWindow fits); a wave whose simulation
// moves farther leaves, and the batched sorted path serves that simulation.
__global__ __launch_bounds__(NT, 6) void k_bscatter_window(float* __restrict__ out, const float* __restrict__ rho0,
                                                           const float* __restrict__ u, const float* __restrict__ v,
                                                           const ScatterSim* __restrict__ sims,
                                                           const int2* __restrict__ windows, int waves,
                                                           const unsigned* __restrict__ max_bits) {
    const int item = (int)blockIdx.x * kStripWaves + (int)(threadIdx.x >> 6);
    if (item >= waves) return;  // whole waves leave; no block-wide barrier follows
    const int2 wk = windows[item];
    const ScatterSim S = sims[wk.x];
    const int W = scatter_radius(max_bits + wk.x, S.dt0, S.R);
    if (W > kWindow) return;
    const int lane = threadIdx.x & 63;
    const int band = wk.y >> 16, strip = wk.y & 0xffff;
    const int col = 1 + strip * kBatchStrip - kWindow + lane;
    const bool dest = lane >= kWindow && lane < kWindow + kBatchStrip && col <= S.R - 2;
    const int r0 = 1 + kBatchBand * band, r1 = min(r0 + kBatchBand, S.R - 1);
    out += S.off;
    rho0 += S.off;
    u += S.off;
    v += S.off;
    switch (W) {
        case 1: window_band_r1(out, rho0, u, v, S.R, S.P, S.dt0, col, r0, r1, dest); break;
        case 2: window_band<2>(out, rho0, u, v, S.R, S.P, S.dt0, col, r0, r1, dest); break;
        case 3: window_band<3>(out, rho0, u, v, S.R, S.P, S.dt0, col, r0, r1, dest); break;
        default: window_band<4>(out, rho0, u, v, S.R, S.P, S.dt0, col, r0, r1, dest); break;
    }
}

// The batch's gate: some simulation moves farther than the window path serves.
__global__ __launch_bounds__(256) void k_bsort_gate(int* __restrict__ flag, const ScatterSim* __restrict__ sims,
                                                    int count, const unsigned* __restrict__ max_bits) {
    int any = 0;
    for (int s = threadIdx.x; s < count; s += 256) any |= scatter_radius(max_bits + s, sims[s].dt0, sims[s].R) > kWindow;
    any = __syncthreads_or(any);
    if (threadIdx.x == 0) *flag = any;
}

// Simulation of batch cell g: the last one whose first cell is <= g.
__device__ __forceinline__ int sim_of_cell(const ScatterSim* __restrict__ sims, int count, unsigned g) {
    int lo = 0, hi = count - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (sims[mid].cells <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Sort items over the whole batch: the sources of simulations on the sorted
// path keyed by their batch target cell, every other source keyed past the
// last target (sorted last, never merged).
__global__ __launch_bounds__(NT) void k_bsort_keys(unsigned long long* __restrict__ items, unsigned* __restrict__ hist,
                                                   int2* __restrict__ bounds, const float* __restrict__ u,
                                                   const float* __restrict__ v, const ScatterSim* __restrict__ sims,
                                                   int count, long long N, unsigned nb, int chunks,
                                                   const unsigned* __restrict__ max_bits, FlagGate gate) {
    if (!gate.sorted()) return;
    __shared__ unsigned cnt[ND];
    for (long long i = (long long)blockIdx.x * NT + threadIdx.x; i < (long long)nb; i += (long long)gridDim.x * NT)
        bounds[i] = make_int2(0, 0);
    for (int c = blockIdx.x; c < chunks; c += gridDim.x) {
        for (int d = threadIdx.x; d < ND; d += NT) cnt[d] = 0u;
        __syncthreads();
        for (int q = 0; q < IPT; ++q) {
            const long long g = (long long)c * CH + q * NT + threadIdx.x;
            if (g >= N) break;
            const int s = sim_of_cell(sims, count, (unsigned)g);
            const ScatterSim S = sims[s];
            const int n = S.R - 2;
            const int local = (int)((unsigned)g - S.cells);
            const int si = 1 + local / n, sj = 1 + local % n;
            unsigned key = nb;
            if (scatter_radius(max_bits + s, S.dt0, S.R) > kWindow) {
                const size_t o = S.off + (size_t)si * S.P + sj;
                int i0, j0;
                float s0, s2;
                if (source_lands(si, sj, u[o], v[o], S.dt0, (float)S.R - 1.5f, i0, j0, s0, s2))
                    key = S.targets + (unsigned)i0 * (unsigned)(S.R - 1) + (unsigned)j0;
            }
            items[g] = ((unsigned long long)key << 32) | (unsigned long long)g;
            atomicAdd(&cnt[key & (ND - 1)], 1u);
        }
        __syncthreads();
        for (int d = threadIdx.x; d < ND; d += NT) hist[d * chunks + c] = cnt[d];
        __syncthreads();
    }
}

// As k_scatter_merge, for the cells of simulations on the sorted path.
__global__ __launch_bounds__(NT) void k_bscatter_merge(float* __restrict__ out, const float* __restrict__ rho0,
                                                       const float* __restrict__ u, const float* __restrict__ v,
                                                       const unsigned long long* __restrict__ items,
                                                       const int2* __restrict__ bounds,
                                                       const ScatterSim* __restrict__ sims, int count, long long N,
                                                       const unsigned* __restrict__ max_bits, FlagGate gate) {
    if (!gate.sorted()) return;
    for (long long g = (long long)blockIdx.x * NT + threadIdx.x; g < N; g += (long long)gridDim.x * NT) {
        const int s = sim_of_cell(sims, count, (unsigned)g);
        const ScatterSim S = sims[s];
        if (scatter_radius(max_bits + s, S.dt0, S.R) <= kWindow) continue;
        const int n = S.R - 2, R = S.R, P = S.P;
        const float hi = (float)R - 1.5f;
        const int local = (int)((unsigned)g - S.cells);
        const int di = 1 + local / n, dj = 1 + local % n;
        int cur[4], end[4];
        const int ti[4] = {di, di - 1, di, di - 1}, tj[4] = {dj, dj, dj - 1, dj - 1};
        for (int k = 0; k < 4; ++k) {
            const int2 b = bounds[(size_t)S.targets + (size_t)ti[k] * (R - 1) + tj[k]];
            cur[k] = b.x;
            end[k] = b.y;
        }
        float acc = 0.f;
        for (;;) {
            int best = -1;
            unsigned bs = 0xffffffffu;
            for (int k = 0; k < 4; ++k) {
                if (cur[k] >= end[k]) continue;
                const unsigned src = (unsigned)items[cur[k]];
                if (src < bs) {
                    bs = src;
                    best = k;
                }
            }
            if (best < 0) break;
            ++cur[best];
            const int sl = (int)(bs - S.cells);
            const int si = 1 + sl / n, sj = 1 + sl % n;
            const size_t o = S.off + (size_t)si * P + sj;
            int i0, j0;
            float s0, s2;
            source_lands(si, sj, u[o], v[o], S.dt0, hi, i0, j0, s0, s2);
            const float s1 = 1.0f - s0, s3 = 1.0f - s2;
            const float r0 = rho0[o];
            const float sh = (best == 0) ? s1 * s3 * r0 : (best == 1) ? s1 * s2 * r0 : (best == 2) ? s0 * s3 * r0 : s0 * s2 * r0;
            acc = acc + sh;
        }
        float* o = out + S.off;
        o[(size_t)di * P + dj] = acc;
        cells::write_ring(o, di, dj, acc, R, P, Ring{false, false});
    }
}

}  // namespace

// Bits of sort keys 0 .. nb (nb itself for skipped sources).
static int key_bits(unsigned long long nb) {
    int bits = 1;
    while ((1ull << bits) <= nb) ++bits;
    return bits;
}
// Targets 0 .. (R - 1)^2 of one grid.
static int sort_key_bits(int R) { return key_bits((unsigned long long)(R - 1) * (unsigned long long)(R - 1)); }

int scatter_launches(int R) {
    const int passes = (sort_key_bits(R) + DB - 1) / DB;
    return 3 + 3 * passes;  // window, keys, (passes - 1) histograms, passes x (scan, scatter), bounds, merge
}

size_t scatter_sort_bytes(int R) {
    const size_t n = static_cast<size_t>(R - 2);
    const size_t N = n * n;
    const size_t chunks = std::max<size_t>((N + CH - 1) / CH, 1);
    return 2 * N * sizeof(unsigned long long) + static_cast<size_t>(R - 1) * (R - 1) * sizeof(int2) +
           ND * chunks * sizeof(unsigned) + 4096;
}

hipError_t launch_scatter_density(hipStream_t s, float* out, const float* rho0, const float* u, const float* v, int R,
                                  int P, float dt0, const unsigned* max_bits, void* sort_scratch) {
    const int n = R - 2;
    const int slots = window_slots();
    int waves = 0;  // the most any radius of the window path takes
    for (int W = 1; W <= kWindow; ++W) waves = std::max(waves, window_layout(R, W, slots).waves);
    const int blocks = (waves + kStripWaves - 1) / kStripWaves, per = (blocks + 7) / 8;
    hipLaunchKernelGGL(k_scatter_window, dim3(8 * per), dim3(NT), 0, s, out, rho0, u, v, R, P, dt0, max_bits, slots,
                       per);
    if (!sort_scratch) return hipGetLastError();
    // sorted path (returns at once while the window path is active)
    const long long N = (long long)n * n;
    const int chunks = static_cast<int>(std::max<long long>((N + CH - 1) / CH, 1));
    auto* items0 = static_cast<unsigned long long*>(sort_scratch);
    auto* items1 = items0 + N;
    auto* bounds = reinterpret_cast<int2*>(items1 + N);
    const long long nb = (long long)(R - 1) * (R - 1);
    auto* hist = reinterpret_cast<unsigned*>(bounds + nb);
    const unsigned chunk_blocks = static_cast<unsigned>(std::min(chunks, kSortBlocks));
    const unsigned cell_blocks = static_cast<unsigned>(std::min<long long>((N + NT - 1) / NT, kSortBlocks));
    hipLaunchKernelGGL(k_sort_keys, dim3(chunk_blocks), dim3(NT), 0, s, items0, hist, bounds, u, v, R, P, dt0, chunks,
                       max_bits);
    const int bits = sort_key_bits(R);
    const GridGate gate{max_bits, dt0, R};
    unsigned long long* src = items0;
    unsigned long long* dst = items1;
    for (int shift = 0; shift < bits; shift += DB) {
        if (shift > 0)
            hipLaunchKernelGGL(k_sort_hist<GridGate>, dim3(chunk_blocks), dim3(NT), 0, s, src, N, shift, hist, chunks,
                               gate);
        hipLaunchKernelGGL(k_sort_scan<GridGate>, dim3(1), dim3(1024), 0, s, hist, ND * chunks, gate);
        hipLaunchKernelGGL(k_sort_scatter<GridGate>, dim3(chunk_blocks), dim3(NT), 0, s, src, dst, N, shift, hist,
                           chunks, gate);
        std::swap(src, dst);
    }
    hipLaunchKernelGGL(k_sort_bounds<GridGate>, dim3(cell_blocks), dim3(NT), 0, s, src, N, bounds, (unsigned)nb, gate);
    hipLaunchKernelGGL(k_scatter_merge, dim3(cell_blocks), dim3(NT), 0, s, out, rho0, u, v, src, bounds, R, P, dt0,
                       max_bits);
    return hipGetLastError();
}

size_t batch_scatter_bytes(long long cells, long long targets) {
    const size_t N = static_cast<size_t>(cells);
    const size_t chunks = std::max<size_t>((N + CH - 1) / CH, 1);
    return 2 * N * sizeof(unsigned long long) + static_cast<size_t>(targets) * sizeof(int2) +
           ND * chunks * sizeof(unsigned) + 4096;
}

int batch_scatter_launches(long long targets) {
    const int passes = (key_bits(static_cast<unsigned long long>(targets)) + DB - 1) / DB;
    return 4 + 3 * passes;  // window, gate, keys, (passes - 1) histograms, passes x (scan, scatter), bounds, merge
}

hipError_t launch_batch_scatter(hipStream_t s, const BatchScatter& B, float* out, const float* rho0, const float* u,
                                const float* v, const unsigned* max_bits) {
    if (B.cells <= 0 || B.cells >= (1ll << 31) || B.targets >= (1ll << 32) - 1 || !B.scratch || !B.flag)
        return hipErrorInvalidValue;
    if (B.window_waves > 0)
        hipLaunchKernelGGL(k_bscatter_window, dim3((B.window_waves + kStripWaves - 1) / kStripWaves), dim3(NT), 0, s,
                           out, rho0, u, v, B.sims, B.windows, B.window_waves, max_bits);
    hipLaunchKernelGGL(k_bsort_gate, dim3(1), dim3(256), 0, s, B.flag, B.sims, B.count, max_bits);
    // sorted path (returns at once unless the gate is set)
    const long long N = B.cells;
    const int chunks = static_cast<int>(std::max<long long>((N + CH - 1) / CH, 1));
    auto* items0 = static_cast<unsigned long long*>(B.scratch);
    auto* items1 = items0 + N;
    auto* bounds = reinterpret_cast<int2*>(items1 + N);
    auto* hist = reinterpret_cast<unsigned*>(bounds + B.targets);
    const unsigned nb = static_cast<unsigned>(B.targets);
    const unsigned chunk_blocks = static_cast<unsigned>(std::min(chunks, kSortBlocks));
    const unsigned cell_blocks = static_cast<unsigned>(std::min<long long>((N + NT - 1) / NT, kSortBlocks));
    const FlagGate gate{B.flag};
    hipLaunchKernelGGL(k_bsort_keys, dim3(chunk_blocks), dim3(NT), 0, s, items0, hist, bounds, u, v, B.sims, B.count,
                       N, nb, chunks, max_bits, gate);
    const int bits = key_bits(nb);
    unsigned long long* src = items0;
    unsigned long long* dst = items1;
    for (int shift = 0; shift < bits; shift += DB) {
        if (shift > 0)
            hipLaunchKernelGGL(k_sort_hist<FlagGate>, dim3(chunk_blocks), dim3(NT), 0, s, src, N, shift, hist, chunks,
                               gate);
        hipLaunchKernelGGL(k_sort_scan<FlagGate>, dim3(1), dim3(1024), 0, s, hist, ND * chunks, gate);
        hipLaunchKernelGGL(k_sort_scatter<FlagGate>, dim3(chunk_blocks), dim3(NT), 0, s, src, dst, N, shift, hist,
                           chunks, gate);
        std::swap(src, dst);
    }
    hipLaunchKernelGGL(k_sort_bounds<FlagGate>, dim3(cell_blocks), dim3(NT), 0, s, src, N, bounds, nb, gate);
    hipLaunchKernelGGL(k_bscatter_merge, dim3(cell_blocks), dim3(NT), 0, s, out, rho0, u, v, src, bounds, B.sims,
                       B.count, N, max_bits, gate);
    return hipGetLastError();
}

}  // namespace fs2d
