// fs2d_stages.hip -- single-pass stages of the stable-fluids step on gfx950.
//
// Each kernel is one HBM pass over the grid (2-D blocks of 64 x 4 threads,
// one 64-lane wave per row segment so loads and stores are coalesced), and
// each fuses the boundary ring of the field it produces: the thread that
// writes interior cell (1, j) also writes ring cell (0, j), and so on, so no
// separate set_boundary launch follows (src/fluid_solver_cpu.cpp:33-83 are
// closed forms of the interior).
#include "fs2d_cells.hpp"
#include "fs2d_internal.hpp"

namespace fs2d {
namespace {

using cells::sgn;

constexpr int BX = 64, BY = 4;

__device__ __forceinline__ bool interior(int R, int& i, int& j) {
    j = blockIdx.x * BX + threadIdx.x + 1;
    i = blockIdx.y * BY + threadIdx.y + 1;
    return i <= R - 2 && j <= R - 2;
}

dim3 grid2d(int R) { return dim3((R - 2 + BX - 1) / BX, (R - 2 + BY - 1) / BY); }

__global__ __launch_bounds__(BX* BY) void k_add_sources(float* __restrict__ x, float* __restrict__ x0,
                                                        const float* __restrict__ s, int R, int P, float dt) {
    int i, j;
    if (interior(R, i, j)) cells::add_sources(x, x0, s, P, dt, i, j);
}

__global__ __launch_bounds__(BX* BY) void k_add_sources_split(float* __restrict__ out, float* __restrict__ x0,
                                                              const float* __restrict__ x, const float* __restrict__ s,
                                                              int R, int P, float dt) {
    int i, j;
    if (interior(R, i, j)) cells::add_sources_split(out, x0, x, s, R, P, dt, i, j);
}

// The step's first stage in one pass (src/fluid_solver_cpu.cpp:16, :21-22 with
// the copies diffuse takes, :102): density in place, rho += dt * rs, with its
// copy; u and v into fresh iterates, un = u + dt * us, with their copies, and
// the fresh iterates' ring copied from u and v (the stale ring the first sweep
// reads).  A thread takes 4 columns of a row, with 16-byte accesses where all
// 4 are interior (rows start 256-byte aligned, so quads are).
__global__ __launch_bounds__(256) void k_add_sources3(float* __restrict__ rho, float* __restrict__ x0,
                                                     const float* __restrict__ rs, float* __restrict__ un,
                                                     float* __restrict__ x0u, const float* __restrict__ u,
                                                     const float* __restrict__ us, float* __restrict__ vn,
                                                     float* __restrict__ x0v, const float* __restrict__ v,
                                                     const float* __restrict__ vs, int R, int P, float dt) {
    const int j0 = 4 * (blockIdx.x * 256 + threadIdx.x), i = blockIdx.y;
    if (j0 >= R) return;
    const size_t o = (size_t)i * P + j0;
    const int n = R - 2;
    if (i >= 1 && i <= n && j0 >= 1 && j0 + 3 <= n) {
        const float4 a = *reinterpret_cast<const float4*>(rho + o), sa = *reinterpret_cast<const float4*>(rs + o);
        const float4 b = *reinterpret_cast<const float4*>(u + o), sb = *reinterpret_cast<const float4*>(us + o);
        const float4 c = *reinterpret_cast<const float4*>(v + o), sc = *reinterpret_cast<const float4*>(vs + o);
        const float4 ra = make_float4(a.x + dt * sa.x, a.y + dt * sa.y, a.z + dt * sa.z, a.w + dt * sa.w);
        const float4 rb = make_float4(b.x + dt * sb.x, b.y + dt * sb.y, b.z + dt * sb.z, b.w + dt * sb.w);
        const float4 rc = make_float4(c.x + dt * sc.x, c.y + dt * sc.y, c.z + dt * sc.z, c.w + dt * sc.w);
        *reinterpret_cast<float4*>(rho + o) = ra;
        *reinterpret_cast<float4*>(x0 + o) = ra;
        *reinterpret_cast<float4*>(un + o) = rb;
        *reinterpret_cast<float4*>(x0u + o) = rb;
        *reinterpret_cast<float4*>(vn + o) = rc;
        *reinterpret_cast<float4*>(x0v + o) = rc;
        return;
    }
    for (int j = j0; j < min(j0 + 4, R); ++j) {
        const size_t q = (size_t)i * P + j;
        if (i >= 1 && i <= n && j >= 1 && j <= n) {
            const float ra = rho[q] + dt * rs[q], rb = u[q] + dt * us[q], rc = v[q] + dt * vs[q];
            rho[q] = ra;
            x0[q] = ra;
            un[q] = rb;
            x0u[q] = rb;
            vn[q] = rc;
            x0v[q] = rc;
        } else {  // the ring, corners included
            un[q] = u[q];
            vn[q] = v[q];
        }
    }
}

// Largest finite |u|, |v| over the interior, as float bits (non-negative
// floats order like their bit patterns).  Bounds how far a source moves in the
// density scatter.
__global__ __launch_bounds__(256) void k_max_speed(unsigned* __restrict__ out, const float* __restrict__ u,
                                                   const float* __restrict__ v, int R, int P) {
    float m = 0.f;
    for (int i = blockIdx.x + 1; i <= R - 2; i += gridDim.x) {
        const float* ur = u + (size_t)i * P;
        const float* vr = v + (size_t)i * P;
        for (int j = threadIdx.x + 1; j <= R - 2; j += 256) m = cells::fold_speed(m, ur[j], vr[j]);
    }
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

__global__ __launch_bounds__(BX* BY) void k_gather_velocity(float* __restrict__ u1, float* __restrict__ v1,
                                                            const float* __restrict__ u0, const float* __restrict__ v0,
                                                            int R, int P, float dt0) {
    int i, j;
    if (interior(R, i, j)) cells::gather_velocity(u1, v1, u0, v0, R, P, dt0, i, j);
}

__global__ __launch_bounds__(BX* BY) void k_divergence(float* __restrict__ d, const float* __restrict__ u,
                                                       const float* __restrict__ v, int R, int P, float neg_half_h,
                                                       float* __restrict__ p) {
    int i, j;
    if (interior(R, i, j)) cells::divergence(d, u, v, R, P, neg_half_h, p, i, j);
}

// The gradient, with the largest finite |u|, |v| of the result reduced for
// the next scatter: block reduction, then one atomic per block spread over
// kSpeedSlots slots on separate cache lines (k_fold_speed folds them).
// Four cells per thread (columns j0 .. j0 + 3 of interior row i, j0 a multiple
// of 4): u, v and the pressure rows as float4; cells::gradient's arithmetic
// and ring writes per cell.  Quads that meet column 0 or pass column n take
// the scalar body cell by cell.
__global__ __launch_bounds__(256) void k_gradient(float* __restrict__ u, float* __restrict__ v,
                                                  const float* __restrict__ p, int R, int P, float h,
                                                  int p_is_zero, unsigned* __restrict__ speed_slots) {
    const int j0 = 4 * (blockIdx.x * 256 + threadIdx.x), i = blockIdx.y + 1;
    const int n = R - 2;
    float m = 0.f;
    if (j0 >= 1 && j0 + 3 <= n) {
        const size_t o = (size_t)i * P + j0;
        const float4 uq = *reinterpret_cast<const float4*>(u + o), vq = *reinterpret_cast<const float4*>(v + o);
        float pw[4] = {0.f, 0.f, 0.f, 0.f}, pe[4] = {0.f, 0.f, 0.f, 0.f};
        float pn[4] = {0.f, 0.f, 0.f, 0.f}, ps[4] = {0.f, 0.f, 0.f, 0.f};
        if (!p_is_zero) {
            const float4 c = *reinterpret_cast<const float4*>(p + o);
            const float4 sq = *reinterpret_cast<const float4*>(p + o + P);
            const float4 nq = *reinterpret_cast<const float4*>(p + o - P);
            pw[0] = p[o - 1]; pw[1] = c.x; pw[2] = c.y; pw[3] = c.z;
            pe[0] = c.y; pe[1] = c.z; pe[2] = c.w; pe[3] = p[o + 4];
            ps[0] = sq.x; ps[1] = sq.y; ps[2] = sq.z; ps[3] = sq.w;
            pn[0] = nq.x; pn[1] = nq.y; pn[2] = nq.z; pn[3] = nq.w;
        }
        const float uo[4] = {uq.x, uq.y, uq.z, uq.w}, vo[4] = {vq.x, vq.y, vq.z, vq.w};
        float un[4], vn[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            un[k] = uo[k] - 0.5f * (pe[k] - pw[k]) / h;
            vn[k] = vo[k] - 0.5f * (ps[k] - pn[k]) / h;
            m = cells::fold_speed(m, un[k], vn[k]);
        }
        *reinterpret_cast<float4*>(u + o) = make_float4(un[0], un[1], un[2], un[3]);
        *reinterpret_cast<float4*>(v + o) = make_float4(vn[0], vn[1], vn[2], vn[3]);
        if (i == 1 || i == n || j0 + 3 == n) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                cells::write_ring(u, i, j0 + k, un[k], R, P, Ring{false, true});
                cells::write_ring(v, i, j0 + k, vn[k], R, P, Ring{true, false});
            }
        }
    } else if (j0 < R) {
        for (int j = max(j0, 1); j <= min(j0 + 3, n); ++j)
            m = fmaxf(m, cells::gradient(u, v, p, R, P, h, p_is_zero != 0, i, j));
    }
    if (speed_slots) {
        m = cells::block_max(m);
        if (threadIdx.x == 0) {
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

__global__ __launch_bounds__(256) void k_pressure_residual(const float* __restrict__ d, const float* __restrict__ p,
                                                          int R, int P, double* __restrict__ sums) {
    const int n = R - 2;
    double rr = 0, dd = 0;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < static_cast<long>(n) * n;
         t += static_cast<long>(gridDim.x) * 256) {
        const size_t o = (size_t)(t / n + 1) * P + t % n + 1;
        const double r = (double)d[o] - (4.0 * p[o] - (((double)p[o - P] + p[o + P]) + ((double)p[o - 1] + p[o + 1])));
        rr += r * r;
        dd += (double)d[o] * d[o];
    }
    for (int off = 32; off > 0; off >>= 1) {
        rr += __shfl_xor(rr, off);
        dd += __shfl_xor(dd, off);
    }
    __shared__ double part[2][4];
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = rr;
        part[1][threadIdx.x >> 6] = dd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // one pair of atomics per block
        atomicAdd(sums, (part[0][0] + part[0][1]) + (part[0][2] + part[0][3]));
        atomicAdd(sums + 1, (part[1][0] + part[1][1]) + (part[1][2] + part[1][3]));
    }
}

__global__ __launch_bounds__(256) void k_copy4(float4* __restrict__ dst, const float4* __restrict__ src, size_t n) {
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256) dst[t] = src[t];
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

hipError_t launch_add_sources3(hipStream_t s, float* rho, float* x0, const float* rs, float* un, float* x0u,
                               const float* u, const float* us, float* vn, float* x0v, const float* v,
                               const float* vs, int R, int P, float dt) {
    const dim3 grid((R + 4 * 256 - 1) / (4 * 256), R);
    hipLaunchKernelGGL(k_add_sources3, grid, dim3(256), 0, s, rho, x0, rs, un, x0u, u, us, vn, x0v, v, vs, R, P, dt);
    return hipGetLastError();
}

hipError_t launch_max_speed(hipStream_t s, unsigned* out_bits, const float* u, const float* v, int R, int P) {
    hipLaunchKernelGGL(k_max_speed, dim3(std::min(R - 2, 2048)), dim3(256), 0, s, out_bits, u, v, R, P);
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
    hipLaunchKernelGGL(k_gradient, dim3((R + 1023) / 1024, R - 2), dim3(256), 0, s, u, v, p, R, P, h, p_is_zero,
                       max_bits ? speed_slots : nullptr);
    if (max_bits) hipLaunchKernelGGL(k_fold_speed, dim3(1), dim3(kSpeedSlots), 0, s, max_bits, speed_slots);
    return hipGetLastError();
}

hipError_t launch_set_boundary(hipStream_t s, float* x, int R, int P, Ring ring) {
    hipLaunchKernelGGL(k_set_boundary, dim3((R - 2 + 255) / 256), dim3(256), 0, s, x, R, P, ring);
    return hipGetLastError();
}

hipError_t launch_pressure_residual(hipStream_t s, const float* d, const float* p, int R, int P, double* sums) {
    const long cells = static_cast<long>(R - 2) * (R - 2);
    const unsigned blocks = static_cast<unsigned>(std::max<long>(1, std::min<long>((cells + 255) / 256, 1024)));
    hipLaunchKernelGGL(k_pressure_residual, dim3(blocks), dim3(256), 0, s, d, p, R, P, sums);
    return hipGetLastError();
}

hipError_t launch_copy4(hipStream_t s, float4* dst, const float4* src, size_t n4) {
    const size_t blocks = std::min<size_t>((n4 + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(k_copy4, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s, dst, src, n4);
    return hipGetLastError();
}

hipError_t launch_fill(hipStream_t s, void* p, uint32_t value, size_t n_words) {
    const size_t blocks = std::min<size_t>((n_words + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill, dim3((unsigned)std::max<size_t>(blocks, 1)), dim3(256), 0, s,
                       static_cast<uint32_t*>(p), value, n_words);
    return hipGetLastError();
}

}  // namespace fs2d
