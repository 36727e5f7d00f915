// fs2d_cells.hpp -- per-cell bodies of the single-pass stages, shared by the
// one-grid kernels (fs2d_stages.hip) and the batched kernels (fs2d_batch.hip).
//
// Each function computes one interior cell (i, j) of a stage of
// fluid_solver_cpu (src/fluid_solver_cpu.cpp) on a grid of R x R floats with
// row pitch P, and writes the boundary-ring cells that depend on it.
#pragma once

#include "fs2d_internal.hpp"

#include <cfloat>

namespace fs2d {
namespace cells {

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

// src/fluid_solver_cpu.cpp:85-93, plus the copy diffuse takes (:102).
__device__ __forceinline__ void add_sources(float* __restrict__ x, float* __restrict__ x0,
                                            const float* __restrict__ s, int P, float dt, int i, int j) {
    const size_t o = (size_t)i * P + j;
    const float v = x[o] + dt * s[o];
    x[o] = v;
    if (x0) x0[o] = v;
}

// The same update written to a fresh iterate: out = x + dt*s (interior),
// x0 = out, and out's ring copied from x (the stale ring the first sweep reads).
__device__ __forceinline__ void add_sources_split(float* __restrict__ out, float* __restrict__ x0,
                                                  const float* __restrict__ x, const float* __restrict__ s, int R,
                                                  int P, float dt, int i, int j) {
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

// Largest finite magnitude of a and b folded into m (inf and NaN are skipped).
__device__ __forceinline__ float fold_speed(float m, float a, float b) {
    a = fabsf(a);
    b = fabsf(b);
    if (a <= FLT_MAX) m = fmaxf(m, a);
    if (b <= FLT_MAX) m = fmaxf(m, b);
    return m;
}

// std::min / std::max as <algorithm> defines them (the clamp's NaN handling
// depends on the operand order, src/fluid_solver_cpu.cpp:158-159).
__device__ __forceinline__ float std_min(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float std_max(float a, float b) { return (a < b) ? b : a; }

// Velocity self-advection, src/fluid_solver_cpu.cpp:26-29,153-174: u and v are
// gathered at the same clamped backtrace point, so both share the index and
// weight arithmetic.  Reads the pre-advection (u0, v0); writes (u1, v1) with
// their opposite rings.
__device__ __forceinline__ void gather_velocity(float* __restrict__ u1, float* __restrict__ v1,
                                                const float* __restrict__ u0, const float* __restrict__ v0, int R,
                                                int P, float dt0, int i, int j) {
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
__device__ __forceinline__ void divergence(float* __restrict__ d, const float* __restrict__ u,
                                           const float* __restrict__ v, int R, int P, float neg_half_h,
                                           float* __restrict__ p, int i, int j) {
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
// v -= (0.5f * (pS - pN)) / h, then the opposite rings of u and v.  Returns
// the largest finite |u|, |v| of the cell's result.
__device__ __forceinline__ float gradient(float* __restrict__ u, float* __restrict__ v, const float* __restrict__ p,
                                          int R, int P, float h, bool p_is_zero, int i, int j) {
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
    return fold_speed(0.f, un, vn);
}

// Block-wide maximum of m (blockDim.x * blockDim.y threads, a multiple of 64,
// at most 1024); the result is valid in thread 0.
__device__ __forceinline__ float block_max(float m) {
    __shared__ float wmax[16];
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    const int t = threadIdx.y * blockDim.x + threadIdx.x;
    const int nw = (blockDim.x * blockDim.y) >> 6;
    if ((t & 63) == 0) wmax[t >> 6] = m;
    __syncthreads();
    if (t == 0)
        for (int q = 1; q < nw; ++q) m = fmaxf(m, wmax[q]);
    return m;
}

}  // namespace cells
}  // namespace fs2d
