// fs2d_mg.hip -- geometric multigrid pieces for a pressure seed (outside
// parity mode): the pressure equation 4 p - sum(nbrs) = b of
// src/fluid_solver_cpu.cpp:196-204 with the continuous ring, on a hierarchy
// of grids whose sides halve.
//
// On a grid coarsened by two the same 5-point stencil holds with the
// right-hand side scaled by 4 (the operator carries h^2), so the coarse
// right-hand side of a cell is the sum of its four fine residuals.  The
// smoother is weighted Jacobi (omega = 4/5 damps the upper half of the 5-point
// spectrum); the correction is interpolated bilinearly (weights 9/16, 3/16,
// 3/16, 1/16 of the four nearest coarse centres).  Each kernel also writes the
// continuous ring of what it writes: a ring cell is the value of the interior
// cell beside it (corners: of the diagonal cell), so the next stencil finds a
// consistent ring.
#include "fs2d_internal.hpp"

namespace fs2d {
namespace {

constexpr int kMgThreads = 256;

__device__ __forceinline__ void write_with_ring(float* g, int i, int j, int n, int P, float v) {
    const size_t o = (size_t)i * P + j;
    g[o] = v;
    if (i == 1) g[o - P] = v;
    if (i == n) g[o + P] = v;
    if (j == 1) g[o - 1] = v;
    if (j == n) g[o + 1] = v;
    if (i == 1 && j == 1) g[o - P - 1] = v;
    if (i == 1 && j == n) g[o - P + 1] = v;
    if (i == n && j == 1) g[o + P - 1] = v;
    if (i == n && j == n) g[o + P + 1] = v;
}

__global__ __launch_bounds__(kMgThreads) void k_mg_smooth(const float* __restrict__ src, float* __restrict__ dst,
                                                          const float* __restrict__ b, int R, int P, float omega) {
    const int n = R - 2, i = blockIdx.y + 1, j = blockIdx.x * kMgThreads + threadIdx.x + 1;
    if (j > n) return;
    const size_t o = (size_t)i * P + j;
    const float c = src[o];
    const float jac = 0.25f * (b[o] + src[o - P] + src[o + P] + src[o - 1] + src[o + 1]);
    write_with_ring(dst, i, j, n, P, c + omega * (jac - c));
}

__global__ __launch_bounds__(kMgThreads) void k_mg_restrict(const float* __restrict__ p, const float* __restrict__ b,
                                                            int Pf, float* __restrict__ bc, int Rc, int Pc) {
    const int nc = Rc - 2, I = blockIdx.y + 1, J = blockIdx.x * kMgThreads + threadIdx.x + 1;
    if (J > nc) return;
    float sum = 0.f;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
            const size_t o = (size_t)(2 * I - 1 + di) * Pf + (2 * J - 1 + dj);
            sum += b[o] - (4.f * p[o] - p[o - Pf] - p[o + Pf] - p[o - 1] - p[o + 1]);
        }
    bc[(size_t)I * Pc + J] = sum;
}

__global__ __launch_bounds__(kMgThreads) void k_mg_prolong(const float* __restrict__ pc, int Pc,
                                                           float* __restrict__ pf, int Rf, int Pf) {
    const int nf = Rf - 2, i = blockIdx.y + 1, j = blockIdx.x * kMgThreads + threadIdx.x + 1;
    if (j > nf) return;
    // fine cell i lies a quarter coarse cell from the centre of coarse cell
    // (i + 1) / 2, towards coarse cell i0 (the ring at the edges)
    const int I = (i + 1) / 2, J = (j + 1) / 2;
    const int I0 = (i & 1) ? I - 1 : I + 1, J0 = (j & 1) ? J - 1 : J + 1;
    const float corr = 0.5625f * pc[(size_t)I * Pc + J] + 0.1875f * pc[(size_t)I * Pc + J0] +
                       0.1875f * pc[(size_t)I0 * Pc + J] + 0.0625f * pc[(size_t)I0 * Pc + J0];
    write_with_ring(pf, i, j, nf, Pf, pf[(size_t)i * Pf + j] + corr);
}

dim3 interior_grid(int n) { return dim3((n + kMgThreads - 1) / kMgThreads, n); }

}  // namespace

hipError_t launch_mg_smooth(hipStream_t s, const float* src, float* dst, const float* b, int R, int P, float omega) {
    hipLaunchKernelGGL(k_mg_smooth, interior_grid(R - 2), dim3(kMgThreads), 0, s, src, dst, b, R, P, omega);
    return hipGetLastError();
}

hipError_t launch_mg_restrict(hipStream_t s, const float* p, const float* b, int Pf, float* bc, int Rc, int Pc) {
    hipLaunchKernelGGL(k_mg_restrict, interior_grid(Rc - 2), dim3(kMgThreads), 0, s, p, b, Pf, bc, Rc, Pc);
    return hipGetLastError();
}

hipError_t launch_mg_prolong(hipStream_t s, const float* pc, int Pc, float* pf, int Rf, int Pf) {
    hipLaunchKernelGGL(k_mg_prolong, interior_grid(Rf - 2), dim3(kMgThreads), 0, s, pc, Pc, pf, Rf, Pf);
    return hipGetLastError();
}

}  // namespace fs2d
