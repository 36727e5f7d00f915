// fs2d_cnn_kernels.hpp -- device kernels of the pressure-guess network.
//
// Activations are NHWC bfloat16 with a channel count that is a multiple of
// 16; pixel p = (n * H + y) * W + x.  A 3x3 convolution (zero padding,
// stride 1) is an implicit GEMM D[p][co] = sum_{tap, ci} X[p + tap][ci] *
// Wt[tap][co][ci] on v_mfma_f32_16x16x16bf16_1k: lane l supplies
// A[m = l % 16][k = 4 (l / 16) + i] and B[k = 4 (l / 16) + i][n = l % 16] and
// holds D[m = 4 (l / 16) + i][n = l % 16], i = 0..3 (the network built from
// these kernels is checked on the device against a PyTorch reference).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace fs2d {
namespace cnn {

typedef uint16_t bf16;
typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__host__ __device__ inline float bf2f(bf16 b) {
    const uint32_t u = static_cast<uint32_t>(b) << 16;
    float f;
    __builtin_memcpy(&f, &u, 4);
    return f;
}

// Round to nearest even; NaN stays a quiet NaN.
__host__ __device__ inline bf16 f2bf(float f) {
    uint32_t u;
    __builtin_memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return static_cast<bf16>((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<bf16>(u >> 16);
}

constexpr int CONV_WAVES = 4;  // pixel groups of 16 per block

// Forward 3x3 convolution, NT output-channel tiles of 16 per wave.
// w: [9][Cout][Cin] bf16; bias: [Cout] fp32 (may be null); y: bf16 or, with
// OUT_F32, fp32 [pixels][Cout].  relu applies max(., 0).
template <int NT, bool OUT_F32>
__global__ __launch_bounds__(64 * CONV_WAVES) void k_conv3x3(const bf16* __restrict__ x, int N, int H, int W, int Cin,
                                                            const bf16* __restrict__ w, const float* __restrict__ bias,
                                                            void* __restrict__ y, int Cout, int relu) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const long npix = static_cast<long>(N) * H * W;
    const long p0 = (static_cast<long>(blockIdx.x) * CONV_WAVES + wave) * 16;
    if (p0 >= npix) return;  // whole wave; npix is a multiple of 16
    const int co0 = blockIdx.y * 16 * NT;
    const long pa = p0 + (l & 15);
    const int xa = static_cast<int>(pa % W);
    const long t0 = pa / W;
    const int ya = static_cast<int>(t0 % H);
    const long na = t0 / H;
    const int kq = (l >> 4) * 4;
    float4_t acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < 9; ++tap) {
        const int yy = ya + tap / 3 - 1, xx = xa + tap % 3 - 1;
        const bool in = (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W;
        // clamped address: always inside the tensor; out-of-image taps read zero
        const int yc = min(max(yy, 0), H - 1), xc = min(max(xx, 0), W - 1);
        const bf16* xp = x + ((na * H + yc) * W + xc) * Cin + kq;
        const bf16* wp = w + (static_cast<long>(tap) * Cout + co0 + (l & 15)) * Cin + kq;
        for (int c0 = 0; c0 < Cin; c0 += 16) {
            short4_t a = *reinterpret_cast<const short4_t*>(xp + c0);
            if (!in) a = short4_t{0, 0, 0, 0};
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const short4_t b = *reinterpret_cast<const short4_t*>(wp + static_cast<long>(t) * 16 * Cin + c0);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = co0 + 16 * t + (l & 15);
        const float bv = bias ? bias[co] : 0.f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            float v = acc[t][i] + bv;
            if (relu) v = fmaxf(v, 0.f);
            const long p = p0 + kq + i;
            if (OUT_F32)
                static_cast<float*>(y)[p * Cout + co] = v;
            else
                static_cast<bf16*>(y)[p * Cout + co] = f2bf(v);
        }
    }
}

// The same convolution with the operands staged in LDS.  A block computes an
// 8-row x 16-column pixel tile (W a multiple of 16, H of 8) for 16 NT output
// channels; wave w takes rows 2w and 2w + 1, so each B fragment feeds two
// MFMAs.  Per 16-channel input chunk the block stages the tile's 10 x 18 input
// halo (zero outside the image) and the chunk's 9 x 16 NT x 16 weights in LDS.
template <int NT, bool OUT_F32>
__global__ __launch_bounds__(256) void k_conv3x3_lds(const bf16* __restrict__ x, int N, int H, int W, int Cin,
                                                    const bf16* __restrict__ w, const float* __restrict__ bias,
                                                    void* __restrict__ y, int Cout, int relu) {
    constexpr int TR = 8, TC = 16, HR = TR + 2, HC = TC + 2;
    constexpr int CO = 16 * NT;
    __shared__ short4_t xs[HR * HC * 4];  // [row][col][16 channels as 4 x short4]
    __shared__ short4_t ws[9 * CO * 4];   // [tap][co][16 channels as 4 x short4]
    const int tid = threadIdx.x, wave = tid >> 6, l = tid & 63;
    const int tiles_x = W / TC, tiles_y = H / TR;
    int b = blockIdx.x;
    const int tx = b % tiles_x;
    b /= tiles_x;
    const int ty = b % tiles_y;
    const long n = b / tiles_y;
    const int x0 = tx * TC, y0 = ty * TR;
    const int co0 = blockIdx.y * CO;
    const int kq4 = l >> 4;  // which short4 of the 16-channel chunk this lane supplies
    float4_t acc[2][NT];
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[g][t] = float4_t{0.f, 0.f, 0.f, 0.f};
    for (int c0 = 0; c0 < Cin; c0 += 16) {
        __syncthreads();  // the previous chunk's reads are done
        for (int q = tid; q < HR * HC * 4; q += 256) {
            const int part = q & 3, pix = q >> 2;
            const int r = pix / HC, c = pix % HC;
            const int yy = y0 + r - 1, xx = x0 + c - 1;
            short4_t v = short4_t{0, 0, 0, 0};
            if ((unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W)
                v = *reinterpret_cast<const short4_t*>(x + ((n * H + yy) * W + xx) * Cin + c0 + 4 * part);
            xs[q] = v;
        }
        for (int q = tid; q < 9 * CO * 4; q += 256) {
            const int part = q & 3, tc = q >> 2;
            const int co = tc % CO, tap = tc / CO;
            ws[q] = *reinterpret_cast<const short4_t*>(w + (static_cast<long>(tap) * Cout + co0 + co) * Cin + c0 +
                                                       4 * part);
        }
        __syncthreads();
#pragma unroll
        for (int tap = 0; tap < 9; ++tap) {
            const int dy = tap / 3, dx = tap % 3;  // offsets into the halo (0..2)
            short4_t a[2];
#pragma unroll
            for (int g = 0; g < 2; ++g) a[g] = xs[((2 * wave + g + dy) * HC + (l & 15) + dx) * 4 + kq4];
#pragma unroll
            for (int t = 0; t < NT; ++t) {
                const short4_t bb = ws[(tap * CO + 16 * t + (l & 15)) * 4 + kq4];
#pragma unroll
                for (int g = 0; g < 2; ++g)
                    acc[g][t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[g], bb, acc[g][t], 0, 0, 0);
            }
        }
    }
    // D[m = 4 (l / 16) + i][n = l % 16]: pixel column x0 + m of row y0 + 2 wave + g
#pragma unroll
    for (int g = 0; g < 2; ++g) {
        const long prow = (n * H + y0 + 2 * wave + g) * W + x0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int co = co0 + 16 * t + (l & 15);
            const float bv = bias ? bias[co] : 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                float v = acc[g][t][i] + bv;
                if (relu) v = fmaxf(v, 0.f);
                const long p = prow + 4 * (l >> 4) + i;
                if (OUT_F32)
                    static_cast<float*>(y)[p * Cout + co] = v;
                else
                    static_cast<bf16*>(y)[p * Cout + co] = f2bf(v);
            }
        }
    }
}

// Weight gradient: gw[tap][co][ci] += sum_p dz[p][co] * x[p + tap][ci] (fp32
// atomics).  Block: (tap, co tile, ci tile) in blockIdx.x, pixel chunk in
// blockIdx.y; each of the CONV_WAVES waves takes a quarter of the chunk.
__global__ __launch_bounds__(64 * CONV_WAVES) void k_conv3x3_wgrad(const bf16* __restrict__ x,
                                                                  const bf16* __restrict__ dz, int N, int H, int W,
                                                                  int Cin, int Cout, float* __restrict__ gw,
                                                                  long chunk) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int nco = Cout / 16, nci = Cin / 16;
    int b = blockIdx.x;
    const int ci0 = (b % nci) * 16;
    b /= nci;
    const int co0 = (b % nco) * 16;
    const int tap = b / nco;
    const int dy = tap / 3 - 1, dx = tap % 3 - 1;
    const long npix = static_cast<long>(N) * H * W;
    const long sub = chunk / CONV_WAVES;
    const long pb = static_cast<long>(blockIdx.y) * chunk + wave * sub;
    const long pe = min(pb + sub, npix);
    const int kq = (l >> 4) * 4;
    float4_t acc = {0.f, 0.f, 0.f, 0.f};
    for (long pp = pb; pp < pe; pp += 16) {
        short4_t a, bb;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const long p = pp + kq + i;  // < npix: chunks are multiples of 16 within npix
            a[i] = static_cast<short>(dz[p * Cout + co0 + (l & 15)]);
            const int xp = static_cast<int>(p % W);
            const long t = p / W;
            const int yp = static_cast<int>(t % H);
            const long np = t / H;
            const int yy = yp + dy, xx = xp + dx;
            const bool in = (unsigned)yy < (unsigned)H && (unsigned)xx < (unsigned)W;
            const int yc = min(max(yy, 0), H - 1), xc = min(max(xx, 0), W - 1);
            const bf16 v = x[((np * H + yc) * W + xc) * Cin + ci0 + (l & 15)];
            bb[i] = in ? static_cast<short>(v) : static_cast<short>(0);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
        atomicAdd(gw + (static_cast<long>(tap) * Cout + co0 + kq + i) * Cin + ci0 + (l & 15), acc[i]);
}

// gb[co] += sum_p dz[p][co]; thread = channel, block = pixel chunk.
__global__ __launch_bounds__(256) void k_bias_grad(const bf16* __restrict__ dz, long npix, int C,
                                                  float* __restrict__ gb, long chunk) {
    const int c = threadIdx.x;
    if (c >= C) return;
    const long pb = static_cast<long>(blockIdx.x) * chunk, pe = min(pb + chunk, npix);
    float s = 0.f;
    for (long p = pb; p < pe; ++p) s += bf2f(dz[p * C + c]);
    atomicAdd(gb + c, s);
}

// dy <- dy where y > 0, else 0 (the ReLU's derivative, y its output).
__global__ __launch_bounds__(256) void k_relu_mask(bf16* __restrict__ dy, const bf16* __restrict__ y, long n) {
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256)
        if (!(bf2f(y[t]) > 0.f)) dy[t] = 0;
}

// wt[8 - tap][ci][co] = w[tap][co][ci]: the weights of the data gradient (a
// 3x3 convolution of dz with the flipped, transposed kernel).
__global__ __launch_bounds__(256) void k_flip_weights(const bf16* __restrict__ w, bf16* __restrict__ wt, int Cin,
                                                     int Cout) {
    const long n = 9L * Cin * Cout;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256) {
        const int ci = static_cast<int>(t % Cin);
        const long r = t / Cin;
        const int co = static_cast<int>(r % Cout);
        const int tap = static_cast<int>(r / Cout);
        wt[(static_cast<long>(8 - tap) * Cin + ci) * Cout + co] = w[t];
    }
}

// 2x2 average: y[n][i][j][c] from x[n][2i..2i+1][2j..2j+1][c].
__global__ __launch_bounds__(256) void k_avgpool2(const bf16* __restrict__ x, bf16* __restrict__ y, int N, int H, int W,
                                                 int C) {
    const int h = H / 2, w = W / 2;
    const long n = static_cast<long>(N) * h * w * C;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256) {
        const int c = static_cast<int>(t % C);
        long r = t / C;
        const int j = static_cast<int>(r % w);
        r /= w;
        const int i = static_cast<int>(r % h);
        const long b = r / h;
        const bf16* q = x + ((b * H + 2 * i) * W + 2 * j) * C + c;
        const float s = (bf2f(q[0]) + bf2f(q[C])) + (bf2f(q[static_cast<long>(W) * C]) + bf2f(q[static_cast<long>(W) * C + C]));
        y[t] = f2bf(0.25f * s);
    }
}

// out[n][y][x][c] = dy[n][y/2][x/2][c] / 4 (+ skip[n][y][x][c]): the pool's
// backward, summed with the gradient the same activation gets through its
// skip connection.
__global__ __launch_bounds__(256) void k_pool_bwd_add(const bf16* __restrict__ dy, const bf16* __restrict__ skip,
                                                     bf16* __restrict__ out, int N, int H, int W, int C) {
    const long n = static_cast<long>(N) * H * W * C;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256) {
        const int c = static_cast<int>(t % C);
        long r = t / C;
        const int x = static_cast<int>(r % W);
        r /= W;
        const int yv = static_cast<int>(r % H);
        const long b = r / H;
        float v = 0.25f * bf2f(dy[((b * (H / 2) + yv / 2) * (W / 2) + x / 2) * C + c]);
        if (skip) v += bf2f(skip[t]);
        out[t] = f2bf(v);
    }
}

// Bilinear x2 (half-pixel centres, edge-clamped): output row Y samples input
// rows y0 = floor(Y / 2 - 1/4) and y0 + 1 with weights 3/4 and 1/4 (or 1/4 and
// 3/4), both clamped into [0, h - 1].
__device__ inline void up_taps(int Y, int h, int& a, int& b, float& wa, float& wb) {
    const int i = Y >> 1;
    if (Y & 1) {
        a = i;
        b = min(i + 1, h - 1);
        wa = 0.75f;
        wb = 0.25f;
    } else {
        a = max(i - 1, 0);
        b = i;
        wa = 0.25f;
        wb = 0.75f;
    }
}

// The decoder's input in one pass: y[p] = (bilinear x2 of x at p -- up_taps'
// two rows and two columns, wy_a (wx_a A + wx_b B) + wy_b (wx_a C + wx_b D) in
// FP32, rounded to bfloat16 -- channels [0, Ca); skip[p], channels [Ca, Ca + Cb)).
// One thread per (pixel, 8-channel group) with 16-byte loads and stores; Ca and
// Cb are multiples of 8.
typedef short short8_t __attribute__((ext_vector_type(8)));

__global__ __launch_bounds__(256) void k_upcat(const bf16* __restrict__ x, const bf16* __restrict__ skip,
                                              bf16* __restrict__ y, int N, int h, int w, int Ca, int Cb) {
    const int H = 2 * h, W = 2 * w, C = Ca + Cb, G = C / 8;
    const long n = static_cast<long>(N) * H * W * G;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256) {
        const int g = static_cast<int>(t % G);
        const long p = t / G;
        short8_t out;
        if (8 * g < Ca) {
            const int X = static_cast<int>(p % W);
            const long r = p / W;
            const int Y = static_cast<int>(r % H);
            const long b = r / H;
            int ya, yb, xa, xb;
            float wya, wyb, wxa, wxb;
            up_taps(Y, h, ya, yb, wya, wyb);
            up_taps(X, w, xa, xb, wxa, wxb);
            const bf16* q = x + b * h * w * Ca + 8 * g;
            const short8_t A = *reinterpret_cast<const short8_t*>(q + (static_cast<long>(ya) * w + xa) * Ca);
            const short8_t B = *reinterpret_cast<const short8_t*>(q + (static_cast<long>(ya) * w + xb) * Ca);
            const short8_t Cc = *reinterpret_cast<const short8_t*>(q + (static_cast<long>(yb) * w + xa) * Ca);
            const short8_t D = *reinterpret_cast<const short8_t*>(q + (static_cast<long>(yb) * w + xb) * Ca);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const float v = wya * (wxa * bf2f(static_cast<bf16>(A[e])) + wxb * bf2f(static_cast<bf16>(B[e]))) +
                                wyb * (wxa * bf2f(static_cast<bf16>(Cc[e])) + wxb * bf2f(static_cast<bf16>(D[e])));
                out[e] = static_cast<short>(f2bf(v));
            }
        } else {
            out = *reinterpret_cast<const short8_t*>(skip + p * Cb + 8 * g - Ca);
        }
        *reinterpret_cast<short8_t*>(y + p * C + 8 * g) = out;
    }
}

// Its backward as a gather: input cell (i, j) collects every output cell whose
// taps include it (output rows 2i-2 .. 2i+3 cover all clamped cases).
__global__ __launch_bounds__(256) void k_upsample2_bwd(const bf16* __restrict__ dy, bf16* __restrict__ dx, int N, int h,
                                                      int w, int C) {
    const int H = 2 * h, W = 2 * w;
    const long n = static_cast<long>(N) * h * w * C;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256) {
        const int c = static_cast<int>(t % C);
        long r = t / C;
        const int j = static_cast<int>(r % w);
        r /= w;
        const int i = static_cast<int>(r % h);
        const long b = r / h;
        float s = 0.f;
        for (int Y = max(2 * i - 2, 0); Y <= min(2 * i + 3, H - 1); ++Y) {
            int ya, yb;
            float wya, wyb;
            up_taps(Y, h, ya, yb, wya, wyb);
            const float wy = (ya == i ? wya : 0.f) + (yb == i ? wyb : 0.f);
            if (wy == 0.f) continue;
            for (int X = max(2 * j - 2, 0); X <= min(2 * j + 3, W - 1); ++X) {
                int xa, xb;
                float wxa, wxb;
                up_taps(X, w, xa, xb, wxa, wxb);
                const float wx = (xa == j ? wxa : 0.f) + (xb == j ? wxb : 0.f);
                if (wx == 0.f) continue;
                s += wy * wx * bf2f(dy[((b * H + Y) * W + X) * C + c]);
            }
        }
        dx[t] = f2bf(s);
    }
}

// The concatenation's backward: da = dy[.., 0..Ca), db = dy[.., Ca..C).
__global__ __launch_bounds__(256) void k_split(const bf16* __restrict__ dy, bf16* __restrict__ da,
                                              bf16* __restrict__ db, long npix, int Ca, int Cb) {
    const int C = Ca + Cb;
    const long n = npix * C;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256) {
        const long p = t / C;
        const int c = static_cast<int>(t % C);
        if (c < Ca)
            da[p * Ca + c] = dy[t];
        else
            db[p * Cb + c - Ca] = dy[t];
    }
}

// Loss of the normalised prediction pred = o[p][0] + prev[p] against target[p]:
// sum of squares into *loss (one atomic per block), and the gradient of the
// mean, 2 (pred - target) / npix, into channel 0 of dout (16 channels).
__global__ __launch_bounds__(256) void k_mse(const float* __restrict__ o, const float* __restrict__ prev,
                                            const float* __restrict__ target, long npix, bf16* __restrict__ dout,
                                            float* __restrict__ loss) {
    __shared__ float part[4];
    float s = 0.f;
    for (long p = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; p < npix; p += static_cast<long>(gridDim.x) * 256) {
        const float d = (o[p * 16] + prev[p]) - target[p];
        s += d * d;
        dout[p * 16] = f2bf(2.f * d / static_cast<float>(npix));
#pragma unroll
        for (int c = 1; c < 16; ++c) dout[p * 16 + c] = 0;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(loss, (part[0] + part[1]) + (part[2] + part[3]));
}

// Adam (Kingma and Ba 2015) on fp32 masters; refreshes the bf16 copy.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                                             float* __restrict__ v, bf16* __restrict__ w16, long n, float lr, float b1,
                                             float b2, float eps, float c1, float c2) {
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256) {
        const float gt = g[t];
        const float mt = b1 * m[t] + (1.f - b1) * gt;
        const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
        m[t] = mt;
        v[t] = vt;
        const float wn = w[t] - lr * (mt / c1) / (sqrtf(vt / c2) + eps);
        w[t] = wn;
        if (w16) w16[t] = f2bf(wn);
    }
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ a, bf16* __restrict__ b, long n) {
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256)
        b[t] = f2bf(a[t]);
}

__global__ __launch_bounds__(256) void k_zero_f32(float* __restrict__ a, long n) {
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < n; t += static_cast<long>(gridDim.x) * 256)
        a[t] = 0.f;
}

}  // namespace cnn
}  // namespace fs2d
