// fs2d_render.hip -- visual output of the resident state: the density field as
// RGBA8 pixels and the velocity field as line vertices.
//
// These replace the two kernels of the reference's renderers
// (src/density_grid_renderer.cu:10-29, src/velocity_grid_renderer.cu:8-44).
// Both read the handle's fields where they already live in HBM, so, unlike the
// reference, nothing is uploaded before drawing.  Both are streaming kernels
// bound by HBM: one row of cells per block row, coalesced loads and stores.
#include "fs2d_internal.hpp"

namespace fs2d {
namespace {

constexpr int kRenderThreads = 256;

// static_cast<uint8_t>(max(0.f, min(255.f, p))) with the device max/min of
// floats (fmaxf/fminf: a NaN product clamps to 255); the conversion truncates.
__device__ __forceinline__ unsigned channel(float p) { return static_cast<unsigned>(fmaxf(0.f, fminf(255.f, p))); }

__global__ __launch_bounds__(kRenderThreads) void k_render_density(const float* __restrict__ d, int R, int P, float mr,
                                                                   float mg, float mb, uint32_t* __restrict__ rgba) {
    const int i = blockIdx.y, j = blockIdx.x * kRenderThreads + threadIdx.x;
    if (j >= R) return;
    const float x = d[(size_t)i * P + j];
    // bytes R, G, B, A in memory order
    rgba[(size_t)i * R + j] = channel(mr * x) | (channel(mg * x) << 8) | (channel(mb * x) << 16) | (255u << 24);
}

// One line per cell: start and end vertex of 5 words each (x, y, RGBA, texture
// x, texture y).  start = end = (j hs, i vs); on every 8th row and column the
// end follows the velocity by 250000 (u, v) / s, s = sqrtf(rows * cols).  The
// block's lines (256 cells of one row, 10240 contiguous bytes) are assembled in
// LDS and written as consecutive 8-byte words, so each store instruction
// covers 512 contiguous bytes.
__global__ __launch_bounds__(kRenderThreads) void k_render_velocity(const float* __restrict__ u,
                                                                    const float* __restrict__ v, int R, int P,
                                                                    float hs, float vs, float s,
                                                                    float2* __restrict__ lines) {
    constexpr int W = 5;  // 8-byte words per line
    __shared__ float2 stage[kRenderThreads * W];
    const int i = blockIdx.y, j0 = blockIdx.x * kRenderThreads, j = j0 + threadIdx.x;
    if (j < R) {
        const float x = static_cast<float>(j) * hs, y = static_cast<float>(i) * vs;
        float ex = x, ey = y;
        if ((i & 7) == 0 && (j & 7) == 0) {
            ex = ex + 250000.f * u[(size_t)i * P + j] / s;
            ey = ey + 250000.f * v[(size_t)i * P + j] / s;
        }
        const float white = __uint_as_float(0xffffffffu);
        float2* o = stage + W * threadIdx.x;
        o[0] = make_float2(x, y);
        o[1] = make_float2(white, 0.f);
        o[2] = make_float2(0.f, ex);
        o[3] = make_float2(ey, white);
        o[4] = make_float2(0.f, 0.f);
    }
    __syncthreads();
    const int words = W * min(kRenderThreads, R - j0);
    float2* dst = lines + (size_t)W * ((size_t)i * R + j0);
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int w = threadIdx.x + k * kRenderThreads;
        if (w < words) dst[w] = stage[w];
    }
}

dim3 render_grid(int R) { return dim3((R + kRenderThreads - 1) / kRenderThreads, R); }

}  // namespace

hipError_t launch_render_density(hipStream_t s, const float* d, int R, int P, float mr, float mg, float mb,
                                 void* rgba) {
    hipLaunchKernelGGL(k_render_density, render_grid(R), dim3(kRenderThreads), 0, s, d, R, P, mr, mg, mb,
                       static_cast<uint32_t*>(rgba));
    return hipGetLastError();
}

hipError_t launch_render_velocity(hipStream_t s, const float* u, const float* v, int R, int P, float hs, float vs,
                                  void* lines) {
    // the reference's sqrtf(rows * cols): the size_t product converted to float
    const float sq = std::sqrt(static_cast<float>(static_cast<size_t>(R) * static_cast<size_t>(R)));
    hipLaunchKernelGGL(k_render_velocity, render_grid(R), dim3(kRenderThreads), 0, s, u, v, R, P, hs, vs, sq,
                       static_cast<float2*>(lines));
    return hipGetLastError();
}

}  // namespace fs2d
