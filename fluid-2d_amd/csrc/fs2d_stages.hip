// fs2d_stages.hip -- single-pass stages of the stable-fluids step on gfx950.
//
// Each kernel is one HBM pass over the grid (2-D blocks of 64 x 4 threads,
// one 64-lane wave per row segment so loads and stores are coalesced), and
// each fuses the boundary ring of the field it produces: the thread that
// writes interior cell (1, j) also writes ring cell (0, j), and so on, so no
// separate set_boundary launch follows (src/fluid_solver_cpu.cpp:33-83 are
// closed forms of the interior).
#include "fs2d_internal.hpp"

#include <cfloat>

namespace fs2d {
namespace {

constexpr int BX = 64, BY = 4;

__device__ __forceinline__ float sgn(bool neg, float x) { return neg ? -x : x; }

// Ring cells that depend on interior cell (i, j) holding final value x.
__device__ __forceinline__ void write_ring(float* f, int i, int j, float x, int R, int P, Ring rg) {
    const float tb = sgn(rg.neg_tb, x), lr = sgn(rg.neg_lr, x);
    if (i == 1) f[j] = tb;
    if (i == R - 2) f[(size_t)(R - 1) * P + j] = tb;
    if (j == 1) f[(size_t)i * P] = lr;
    if (j == R - 2) f[(size_t)i * P + R - 1] = lr;
    if (i == 1 || i == R - 2) {
        if (j == 1 || j == R - 2) {
            const float c = 0.5f * (tb + lr);
            if (i == 1 && j == 1) f[0] = c;
            if (i == 1 && j == R - 2) f[R - 1] = c;
            if (i == R - 2 && j == 1) f[(size_t)(R - 1) * P] = c;
            if (i == R - 2 && j == R - 2) f[(size_t)(R - 1) * P + R - 1] = c;
        }
    }
}

__device__ __forceinline__ bool interior(int R, int& i, int& j) {
    j = blockIdx.x * BX + threadIdx.x + 1;
    i = blockIdx.y * BY + threadIdx.y + 1;
    return i <= R - 2 && j <= R - 2;
}

dim3 grid2d(int R) { return dim3((R - 2 + BX - 1) / BX, (R - 2 + BY - 1) / BY); }

// src/fluid_solver_cpu.cpp:85-93, plus the copy diffuse takes (:102).
__global__ __launch_bounds__(BX* BY) void k_add_sources(float* __restrict__ x, float* __restrict__ x0,
                                                        const float* __restrict__ s, int R, int P, float dt) {
    int i, j;
    if (!interior(R, i, j)) return;
    const size_t o = (size_t)i * P + j;
    const float v = x[o] + dt * s[o];
    x[o] = v;
    if (x0) x0[o] = v;
}

// The same update written to a fresh iterate: out = x + dt*s (interior),
// x0 = out, and out's ring copied from x (the stale ring the first sweep reads).
__global__ __launch_bounds__(BX* BY) void k_add_sources_split(float* __restrict__ out, float* __restrict__ x0,
                                                              const float* __restrict__ x, const float* __restrict__ s,
                                                              int R, int P, float dt) {
    int i, j;
    if (!interior(R, i, j)) return;
    const size_t o = (size_t)i * P + j;
    const float v = x[o] + dt * s[o];
    out[o] = v;
    x0[o] = v;
    if (i == 1) out[o - P] = x[o - P];
    if (i == R - 2) out[o + P] = x[o + P];
    if (j == 1) out[o - 1] = x[o - 1];
    if (j == R - 2) out[o + 1] = x[o + 1];
    if (i == 1 && j == 1) {
        const size_t last = (size_t)(R - 1) * P;
        out[0] = x[0];
        out[R - 1] = x[R - 1];
        out[last] = x[last];
        out[last + R - 1] = x[last + R - 1];
    }
}

// Largest finite |u|, |v| over the interior, as float bits (non-negative
// floats order like their bit patterns).  Bounds the scatter's search window.
__global__ __launch_bounds__(256) void k_max_speed(unsigned* __restrict__ out, const float* __restrict__ u,
                                                   const float* __restrict__ v, int R, int P) {
    float m = 0.f;
    for (int i = blockIdx.x + 1; i <= R - 2; i += gridDim.x) {
        const float* ur = u + (size_t)i * P;
        const float* vr = v + (size_t)i * P;
        for (int j = threadIdx.x + 1; j <= R - 2; j += 256) {
            const float a = fabsf(ur[j]), b = fabsf(vr[j]);
            if (a <= FLT_MAX) m = fmaxf(m, a);  // false for inf and NaN
            if (b <= FLT_MAX) m = fmaxf(m, b);
        }
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// Density advection, src/fluid_solver_cpu.cpp:127-152, as a gather.
//
// The reference scatters each source cell (i, j) bilinearly onto the four
// cells around (i + dt0*v, j + dt0*u) with `+=`, in row-major source order,
// into a zero-filled grid.  Here each destination cell walks the source cells
// that can reach it, in the same row-major order, and sums the same products
// in the same order starting from +0, so every destination gets the
// reference's exact float sum with no atomics.  The window is bounded by the
// largest displacement dt0 * max|vel| (a source moves at most that far, and
// lands on trunc(pos) and trunc(pos) + 1).  A non-finite source position is
// skipped (the reference converts NaN to size_t there: undefined behaviour).
__global__ __launch_bounds__(BX* BY) void k_scatter_density(float* __restrict__ out, const float* __restrict__ rho0,
                                                            const float* __restrict__ u, const float* __restrict__ v,
                                                            int R, int P, float dt0, const unsigned* __restrict__ max_bits) {
    int di, dj;
    if (!interior(R, di, dj)) return;
    // A source at row si lands on rows trunc(y), trunc(y) + 1 with
    // |y - si| <= |fl(dt0 v)| + ulp(y) / 2 <= dt0 * max|v| + 2^-8 (y < 2^15),
    // so it reaches row di only if |si - di| <= W = floor(that bound) + 1.
    const float mv = __uint_as_float(*max_bits);
    const float reach = dt0 * mv + 0x1p-8f;
    const int r = (reach < (float)R) ? (int)reach + 1 : R;
    const int i_lo = max(1, di - r), i_hi = min(R - 2, di + r);
    const int j_lo = max(1, dj - r), j_hi = min(R - 2, dj + r);
    const float hi = (float)R - 1.5f;
    float acc = 0.f;
    for (int si = i_lo; si <= i_hi; ++si) {
        const size_t row = (size_t)si * P;
        const float fi = (float)si;
        for (int sj = j_lo; sj <= j_hi; ++sj) {
            const float x = (float)sj + dt0 * u[row + sj];
            const float y = fi + dt0 * v[row + sj];
            if (x < 0.5f || x > hi || y < 0.5f || y > hi) continue;
            if (x != x || y != y) continue;
            const int j0 = (int)x, i0 = (int)y;
            const int ddi = di - i0, ddj = dj - j0;
            if ((unsigned)ddi > 1u || (unsigned)ddj > 1u) continue;
            const float s0 = x - (float)j0;
            const float s1 = 1.0f - s0;
            const float s2 = y - (float)i0;
            const float s3 = 1.0f - s2;
            const float w = (ddj == 0) ? ((ddi == 0) ? s1 * s3 : s1 * s2) : ((ddi == 0) ? s0 * s3 : s0 * s2);
            acc = acc + w * rho0[row + sj];
        }
    }
    out[(size_t)di * P + dj] = acc;
    write_ring(out, di, dj, acc, R, P, Ring{false, false});
}

// std::min / std::max as <algorithm> defines them (the clamp's NaN handling
// depends on the operand order, src/fluid_solver_cpu.cpp:158-159).
__device__ __forceinline__ float std_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float std_max(float a, float b) { return (a < b) ? b : a; }

// Velocity self-advection, src/fluid_solver_cpu.cpp:26-29,153-174: u and v are
// gathered at the same clamped backtrace point, so both share the index and
// weight arithmetic.  Reads the pre-advection (u0, v0); writes (u1, v1) with
// their opposite rings.
__global__ __launch_bounds__(BX* BY) void k_gather_velocity(float* __restrict__ u1, float* __restrict__ v1,
                                                            const float* __restrict__ u0, const float* __restrict__ v0,
                                                            int R, int P, float dt0) {
    int i, j;
    if (!interior(R, i, j)) return;
    const size_t o = (size_t)i * P + j;
    const float hi = (float)R - 1.5f;
    float x = (float)j - dt0 * u0[o];
    float y = (float)i - dt0 * v0[o];
    x = std_max(1.5f, std_min(hi, x));
    y = std_max(1.5f, std_min(hi, y));
    const int j0 = (int)x, i0 = (int)y;
    const float s0 = x - (float)j0;
    const float s1 = 1.0f - s0;
    const float s2 = y - (float)i0;
    const float s3 = 1.0f - s2;
    const size_t a00 = (size_t)i0 * P + j0, a10 = a00 + P;
    const float un = s3 * (s1 * u0[a00] + s0 * u0[a00 + 1]) + s2 * (s1 * u0[a10] + s0 * u0[a10 + 1]);
    const float vn = s3 * (s1 * v0[a00] + s0 * v0[a00 + 1]) + s2 * (s1 * v0[a10] + s0 * v0[a10 + 1]);
    u1[o] = un;
    v1[o] = vn;
    write_ring(u1, i, j, un, R, P, Ring{false, true});
    write_ring(v1, i, j, vn, R, P, Ring{true, false});
}

// src/fluid_solver_cpu.cpp:188-193: d = (-0.5f * h) * (((uE - uW) + vS) - vN).
// The divergence ring (:194) is never read and is not written.  When p is
// given, the pressure iterate is zeroed too, ring included (:186).
__global__ __launch_bounds__(BX* BY) void k_divergence(float* __restrict__ d, const float* __restrict__ u,
                                                       const float* __restrict__ v, int R, int P, float neg_half_h,
                                                       float* __restrict__ p) {
    int i, j;
    if (!interior(R, i, j)) return;
    const size_t o = (size_t)i * P + j;
    d[o] = neg_half_h * (u[o + 1] - u[o - 1] + v[o + P] - v[o - P]);
    if (p) {
        p[o] = 0.f;
        if (i == 1) p[o - P] = 0.f;
        if (i == R - 2) p[o + P] = 0.f;
        if (j == 1) p[o - 1] = 0.f;
        if (j == R - 2) p[o + 1] = 0.f;
        if (i == 1 && j == 1) {
            p[0] = 0.f;
            p[R - 1] = 0.f;
            p[(size_t)(R - 1) * P] = 0.f;
            p[(size_t)(R - 1) * P + R - 1] = 0.f;
        }
    }
}

// src/fluid_solver_cpu.cpp:206-214: u -= (0.5f * (pE - pW)) / h,
// v -= (0.5f * (pS - pN)) / h, then the opposite rings of u and v.
__global__ __launch_bounds__(BX* BY) void k_gradient(float* __restrict__ u, float* __restrict__ v,
                                                     const float* __restrict__ p, int R, int P, float h,
                                                     int p_is_zero, unsigned* __restrict__ speed_slots) {
    int i, j;
    const bool in = interior(R, i, j);
    float m = 0.f;
    if (in) {
        const size_t o = (size_t)i * P + j;
        float pe = 0.f, pw = 0.f, ps = 0.f, pn = 0.f;
        if (!p_is_zero) {
            pe = p[o + 1];
            pw = p[o - 1];
            ps = p[o + P];
            pn = p[o - P];
        }
        const float un = u[o] - 0.5f * (pe - pw) / h;
        const float vn = v[o] - 0.5f * (ps - pn) / h;
        u[o] = un;
        v[o] = vn;
        write_ring(u, i, j, un, R, P, Ring{false, true});
        write_ring(v, i, j, vn, R, P, Ring{true, false});
        if (fabsf(un) <= FLT_MAX) m = fabsf(un);  // finite values only
        if (fabsf(vn) <= FLT_MAX) m = fmaxf(m, fabsf(vn));
    }
    if (speed_slots) {
        // largest finite |u|, |v| of the result, for the next scatter: wave,
        // then block reduction, one atomic per block spread over kSpeedSlots
        // slots on separate cache lines (k_fold_speed folds them)
        __shared__ float wmax[BX * BY / 64];
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        const int t = threadIdx.y * BX + threadIdx.x;
        if ((t & 63) == 0) wmax[t >> 6] = m;
        __syncthreads();
        if (t == 0) {
            for (int q = 1; q < BX * BY / 64; ++q) m = fmaxf(m, wmax[q]);
            const unsigned slot = (blockIdx.y * gridDim.x + blockIdx.x) % kSpeedSlots;
            atomicMax(speed_slots + slot * kSlotStride, __float_as_uint(m));
        }
    }
}

// Fold the per-slot maxima into the scalar the scatter reads, and clear the
// slots for the next gradient.
__global__ __launch_bounds__(kSpeedSlots) void k_fold_speed(unsigned* __restrict__ max_bits,
                                                             unsigned* __restrict__ speed_slots) {
    __shared__ unsigned part[kSpeedSlots / 64];
    const int t = threadIdx.x;
    unsigned m = speed_slots[t * kSlotStride];
    speed_slots[t * kSlotStride] = 0u;
    for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
    if ((t & 63) == 0) part[t >> 6] = m;
    __syncthreads();
    if (t == 0) {
        for (int q = 1; q < kSpeedSlots / 64; ++q) m = max(m, part[q]);
        *max_bits = m;
    }
}

// Standalone ring (src/fluid_solver_cpu.cpp:33-83); every ring cell is a
// function of the interior only, so one thread per ring position suffices.
__global__ __launch_bounds__(256) void k_set_boundary(float* __restrict__ x, int R, int P, Ring rg) {
    const int t = blockIdx.x * 256 + threadIdx.x + 1;  // 1 .. R-2
    if (t > R - 2) return;
    const size_t last = (size_t)(R - 1) * P;
    x[t] = sgn(rg.neg_tb, x[(size_t)P + t]);
    x[last + t] = sgn(rg.neg_tb, x[last - P + t]);
    x[(size_t)t * P] = sgn(rg.neg_lr, x[(size_t)t * P + 1]);
    x[(size_t)t * P + R - 1] = sgn(rg.neg_lr, x[(size_t)t * P + R - 2]);
    if (t == 1) {
        const float c00 = x[(size_t)P + 1], c01 = x[(size_t)P + R - 2];
        const float c10 = x[last - P + 1], c11 = x[last - P + R - 2];
        x[0] = 0.5f * (sgn(rg.neg_tb, c00) + sgn(rg.neg_lr, c00));
        x[R - 1] = 0.5f * (sgn(rg.neg_tb, c01) + sgn(rg.neg_lr, c01));
        x[last] = 0.5f * (sgn(rg.neg_tb, c10) + sgn(rg.neg_lr, c10));
        x[last + R - 1] = 0.5f * (sgn(rg.neg_tb, c11) + sgn(rg.neg_lr, c11));
    }
}

__global__ __launch_bounds__(256) void k_fill(uint32_t* __restrict__ p, uint32_t value, size_t n) {
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256) p[t] = value;
}

}  // namespace

hipError_t launch_add_sources(hipStream_t s, float* x, float* x0, const float* src, int R, int P, float dt) {
    hipLaunchKernelGGL(k_add_sources, grid2d(R), dim3(BX, BY), 0, s, x, x0, src, R, P, dt);
    return hipGetLastError();
}

hipError_t launch_add_sources_split(hipStream_t s, float* out, float* x0, const float* x, const float* src, int R,
                                    int P, float dt) {
    hipLaunchKernelGGL(k_add_sources_split, grid2d(R), dim3(BX, BY), 0, s, out, x0, x, src, R, P, dt);
    return hipGetLastError();
}

hipError_t launch_max_speed(hipStream_t s, unsigned* out_bits, const float* u, const float* v, int R, int P) {
    hipLaunchKernelGGL(k_max_speed, dim3(std::min(R - 2, 2048)), dim3(256), 0, s, out_bits, u, v, R, P);
    return hipGetLastError();
}

hipError_t launch_scatter_density(hipStream_t s, float* out, const float* rho0, const float* u, const float* v,
                                  int R, int P, float dt0, const unsigned* max_bits) {
    hipLaunchKernelGGL(k_scatter_density, grid2d(R), dim3(BX, BY), 0, s, out, rho0, u, v, R, P, dt0, max_bits);
    return hipGetLastError();
}

hipError_t launch_gather_velocity(hipStream_t s, float* u1, float* v1, const float* u0, const float* v0, int R,
                                  int P, float dt0) {
    hipLaunchKernelGGL(k_gather_velocity, grid2d(R), dim3(BX, BY), 0, s, u1, v1, u0, v0, R, P, dt0);
    return hipGetLastError();
}

hipError_t launch_divergence(hipStream_t s, float* d, const float* u, const float* v, int R, int P,
                             float neg_half_h, float* p_zero) {
    hipLaunchKernelGGL(k_divergence, grid2d(R), dim3(BX, BY), 0, s, d, u, v, R, P, neg_half_h, p_zero);
    return hipGetLastError();
}

hipError_t launch_gradient(hipStream_t s, float* u, float* v, const float* p, int R, int P, float h,
                           int p_is_zero, unsigned* max_bits, unsigned* speed_slots) {
    hipLaunchKernelGGL(k_gradient, grid2d(R), dim3(BX, BY), 0, s, u, v, p, R, P, h, p_is_zero,
                       max_bits ? speed_slots : nullptr);
    if (max_bits) hipLaunchKernelGGL(k_fold_speed, dim3(1), dim3(kSpeedSlots), 0, s, max_bits, speed_slots);
    return hipGetLastError();
}

hipError_t launch_set_boundary(hipStream_t s, float* x, int R, int P, Ring ring) {
    hipLaunchKernelGGL(k_set_boundary, dim3((R - 2 + 255) / 256), dim3(256), 0, s, x, R, P, ring);
    return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, void* p, uint32_t value, size_t n_words) {
    const size_t blocks = std::min<size_t>((n_words + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s,
                       static_cast<uint32_t*>(p), value, n_words);
    return hipGetLastError();
}

}  // namespace fs2d
