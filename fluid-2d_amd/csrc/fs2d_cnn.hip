// fs2d_cnn.hip -- the learned pressure guess: network, training step, and the
// prediction that seeds a projection (declared in include/fs2d_cnn.h).
//
// U-Net over (divergence, previous pressure), both divided by one per-sample
// scale s (the RMS of the previous pressure, else of the divergence), so the
// map stays linear in its inputs:
//   level 1 (H):    e1a 2(16)->16, e1b 16->16        -> e1
//   level 2 (H/2):  pool, e2a 16->32, e2b 32->32      -> e2
//   level 3 (H/4):  pool, e3a 32->64, e3b 64->64      -> e3
//   level 4 (H/8):  pool, e4a 64->128, e4b 128->128   -> e4
//   up(e4) ++ e3 -> d3a 192->64, d3b 64->64
//   up(d3) ++ e2 -> d2a 96->32,  d2b 32->32
//   up(d2) ++ e1 -> d1a 48->16,  d1b 16->16
//   out 16->1(16), no ReLU; prediction = previous pressure / s + out.
// Every convolution is 3x3 with bias, ReLU except the last.  The input's 2
// channels and the output's 1 are padded to 16 (their extra weights start at
// zero and receive zero gradients).
#include "../../include/fs2d_cnn.h"
#include "../../include/fs2d_hip.h"
#include "fs2d_cnn_kernels.hpp"
#include "fs2d_internal.hpp"

#include <cmath>
#include <cstring>
#include <map>
#include <exception>
#include <new>
#include <string>
#include <tuple>
#include <vector>

using namespace fs2d;
using namespace fs2d::cnn;

namespace {

thread_local std::string g_net_create_error;

struct LayerSpec {
    int cin, cout;
    bool relu;
};

constexpr LayerSpec kLayers[] = {
    {16, 16, true},   {16, 16, true},    // e1a, e1b
    {16, 32, true},   {32, 32, true},    // e2a, e2b
    {32, 64, true},   {64, 64, true},    // e3a, e3b
    {64, 128, true},  {128, 128, true},  // e4a, e4b
    {192, 64, true},  {64, 64, true},    // d3a, d3b
    {96, 32, true},   {32, 32, true},    // d2a, d2b
    {48, 16, true},   {16, 16, true},    // d1a, d1b
    {16, 16, false},                     // out
};
constexpr int kNumLayers = sizeof(kLayers) / sizeof(kLayers[0]);
enum { E1A, E1B, E2A, E2B, E3A, E3B, E4A, E4B, D3A, D3B, D2A, D2B, D1A, D1B, OUT };

struct Layer {
    int cin = 0, cout = 0;
    bool relu = false;
    long wn = 0;  // 9 * cout * cin
    // fp32 masters, gradients and Adam moments: weights then biases, contiguous
    float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    bf16 *w16 = nullptr, *wt16 = nullptr;
    float* w() const { return p; }
    float* b() const { return p + wn; }
    long count() const { return wn + cout; }
};

// Activation and gradient buffers of one (N, H, W) shape.
struct Acts {
    int N = 0, H = 0, W = 0;
    std::map<std::string, void*> buf;
    bf16* operator()(const char* k) const { return static_cast<bf16*>(buf.at(k)); }
    float* f(const char* k) const { return static_cast<float*>(buf.at(k)); }
};

}  // namespace

struct fs2d_net {
    hipStream_t stream = nullptr;
    Layer layer[kNumLayers];
    float* params = nullptr;  // all layers' p, g, m, v live in four arenas
    float* grads = nullptr;
    float* m1 = nullptr;
    float* m2 = nullptr;
    long total = 0;
    long adam_t = 0;
    float* loss = nullptr;  // device scalar
    double* scal = nullptr;  // device: sums of squares and the scale (prediction on fields)
    std::map<std::tuple<int, int, int>, Acts*> acts;
    std::string err;
};

namespace {

int nfail(fs2d_net* n, int code, const std::string& msg) {
    if (n) n->err = msg;
    else g_net_create_error = msg;
    return code;
}

#define NHIP(n, expr)                                                                         \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return nfail((n), FS2D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

unsigned blocks_for(long n) { return static_cast<unsigned>(std::max<long>(1, std::min<long>((n + 255) / 256, 8192))); }

// SplitMix64 and Box-Muller (FP64), for the weight initialisation.
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        const double u1 = 1.0 - uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
};

// Real fan-in of layer l: the input layer reads 2 channels, the others all cin.
int real_cin(int l) { return l == E1A ? 2 : kLayers[l].cin; }
int real_cout(int l) { return l == OUT ? 1 : kLayers[l].cout; }

int refresh_bf16(fs2d_net* n) {
    for (auto& L : n->layer)
        hipLaunchKernelGGL(k_to_bf16, dim3(blocks_for(L.wn)), dim3(256), 0, n->stream, L.p, L.w16, L.wn);
    NHIP(n, hipGetLastError());
    return FS2D_OK;
}

// One 3x3 convolution launch: the LDS-staged kernel where the image tiles into
// 8 x 16 pixel blocks, the direct one otherwise (the small 1/8-scale levels).
int launch_conv(fs2d_net* n, const bf16* x, int N, int H, int W, int Cin, const bf16* w, const float* bias, void* y,
                int Cout, int relu, bool f32_out) {
    const int tiles = Cout / 16;
    const int nt = tiles % 4 == 0 ? 4 : (tiles % 3 == 0 ? 3 : (tiles % 2 == 0 ? 2 : 1));
    const dim3 blk(256);
    if (H % 8 == 0 && W % 16 == 0) {
        const dim3 grid(static_cast<unsigned>(static_cast<long>(N) * (H / 8) * (W / 16)), tiles / nt);
#define FS2D_CONV_LDS(NT_, F32_) \
    hipLaunchKernelGGL((k_conv3x3_lds<NT_, F32_>), grid, blk, 0, n->stream, x, N, H, W, Cin, w, bias, y, Cout, relu)
        if (f32_out && nt == 1) FS2D_CONV_LDS(1, true);
        else if (f32_out) return nfail(n, FS2D_ERR_ARGUMENT, "conv: fp32 output needs 16 channels");
        else if (nt == 4) FS2D_CONV_LDS(4, false);
        else if (nt == 3) FS2D_CONV_LDS(3, false);
        else if (nt == 2) FS2D_CONV_LDS(2, false);
        else FS2D_CONV_LDS(1, false);
#undef FS2D_CONV_LDS
    } else {
        const long npix = static_cast<long>(N) * H * W;
        const dim3 grid(static_cast<unsigned>((npix / 16 + CONV_WAVES - 1) / CONV_WAVES), tiles / nt);
#define FS2D_CONV(NT_, F32_) \
    hipLaunchKernelGGL((k_conv3x3<NT_, F32_>), grid, blk, 0, n->stream, x, N, H, W, Cin, w, bias, y, Cout, relu)
        if (f32_out && nt == 1) FS2D_CONV(1, true);
        else if (f32_out) return nfail(n, FS2D_ERR_ARGUMENT, "conv: fp32 output needs 16 channels");
        else if (nt == 4) FS2D_CONV(4, false);
        else if (nt == 3) FS2D_CONV(3, false);
        else if (nt == 2) FS2D_CONV(2, false);
        else FS2D_CONV(1, false);
#undef FS2D_CONV
    }
    NHIP(n, hipGetLastError());
    return FS2D_OK;
}

int conv(fs2d_net* n, int l, const bf16* x, void* y, int N, int H, int W, bool f32_out = false) {
    const Layer& L = n->layer[l];
    return launch_conv(n, x, N, H, W, L.cin, L.w16, L.b(), y, L.cout, L.relu ? 1 : 0, f32_out);
}

// Data gradient of layer l: dx = conv(dz, flipped transposed weights).
int conv_dx(fs2d_net* n, int l, const bf16* dz, bf16* dx, int N, int H, int W) {
    const Layer& L = n->layer[l];
    hipLaunchKernelGGL(k_flip_weights, dim3(blocks_for(L.wn)), dim3(256), 0, n->stream, L.w16, L.wt16, L.cin, L.cout);
    return launch_conv(n, dz, N, H, W, L.cout, L.wt16, nullptr, dx, L.cin, 0, false);
}

// Weight and bias gradients of layer l (accumulated into its gradient arena).
int conv_dw(fs2d_net* n, int l, const bf16* x, const bf16* dz, int N, int H, int W) {
    const Layer& L = n->layer[l];
    const long npix = static_cast<long>(N) * H * W;
    const long chunk = std::max<long>(64, std::min<long>(16384, npix / 16 / 16 * 16));  // a multiple of 64
    const long chunk64 = (chunk + 63) / 64 * 64;
    const dim3 grid(9 * (L.cout / 16) * (L.cin / 16), static_cast<unsigned>((npix + chunk64 - 1) / chunk64));
    hipLaunchKernelGGL(k_conv3x3_wgrad, grid, dim3(64 * CONV_WAVES), 0, n->stream, x, dz, N, H, W, L.cin, L.cout, L.g,
                       chunk64);
    const long bchunk = 4096;
    hipLaunchKernelGGL(k_bias_grad, dim3(static_cast<unsigned>((npix + bchunk - 1) / bchunk)), dim3(256), 0, n->stream,
                       dz, npix, L.cout, L.g + L.wn, bchunk);
    NHIP(n, hipGetLastError());
    return FS2D_OK;
}

int relu_mask(fs2d_net* n, bf16* dy, const bf16* y, long count) {
    hipLaunchKernelGGL(k_relu_mask, dim3(blocks_for(count)), dim3(256), 0, n->stream, dy, y, count);
    NHIP(n, hipGetLastError());
    return FS2D_OK;
}

int acts_for(fs2d_net* n, int N, int H, int W, Acts** out) {
    const auto key = std::make_tuple(N, H, W);
    auto it = n->acts.find(key);
    if (it != n->acts.end()) {
        *out = it->second;
        return FS2D_OK;
    }
    Acts* a = new Acts();
    a->N = N;
    a->H = H;
    a->W = W;
    const long p1 = static_cast<long>(N) * H * W, p2 = p1 / 4, p3 = p1 / 16, p4 = p1 / 64;
    struct B {
        const char* name;
        long elems;
        int bytes;
    };
    const B list[] = {
        {"x0", p1 * 16, 2},  {"a1", p1 * 16, 2},  {"e1", p1 * 16, 2},  {"p1", p2 * 16, 2},  {"a2", p2 * 32, 2},
        {"e2", p2 * 32, 2},  {"p2", p3 * 32, 2},  {"a3", p3 * 64, 2},  {"e3", p3 * 64, 2},  {"p3", p4 * 64, 2},
        {"a4", p4 * 128, 2}, {"e4", p4 * 128, 2}, {"c3", p3 * 192, 2}, {"b3", p3 * 64, 2},
        {"d3", p3 * 64, 2},  {"c2", p2 * 96, 2},  {"b2", p2 * 32, 2},  {"d2", p2 * 32, 2},
        {"c1", p1 * 48, 2},  {"b1", p1 * 16, 2},  {"d1", p1 * 16, 2},  {"o", p1 * 16, 4},
        {"prev", p1, 4},     {"tgt", p1, 4},      {"scale", N, 4},
        // gradients
        {"dO", p1 * 16, 2},  {"gA", p1 * 48, 2},  {"gB", p1 * 48, 2},  {"gS1", p1 * 16, 2}, {"gS2", p2 * 32, 2},
        {"gS3", p3 * 64, 2},
    };
    for (const B& b : list) {
        void* ptr = nullptr;
        hipError_t e = hipMalloc(&ptr, static_cast<size_t>(b.elems) * b.bytes);
        if (e != hipSuccess) {
            for (auto& kv : a->buf) (void)hipFree(kv.second);
            delete a;
            return nfail(n, FS2D_ERR_HIP, std::string("hipMalloc(activations): ") + hipGetErrorString(e));
        }
        a->buf[b.name] = ptr;
    }
    n->acts[key] = a;
    *out = a;
    return FS2D_OK;
}

// Forward pass on a->x0 into a->o (fp32, channel 0 meaningful).
int forward(fs2d_net* n, Acts& a) {
    const int N = a.N, H = a.H, W = a.W;
    int rc;
    auto pool = [&](const bf16* x, bf16* y, int h, int w, int C) {
        hipLaunchKernelGGL(k_avgpool2, dim3(blocks_for(static_cast<long>(N) * h * w * C / 4)), dim3(256), 0, n->stream,
                           x, y, N, h, w, C);
    };
    // upsample the coarse level and concatenate the skip in one pass
    auto upcat = [&](const bf16* coarse, const bf16* skip, bf16* out, int h, int w, int Ca, int Cb) {
        hipLaunchKernelGGL(k_upcat, dim3(blocks_for(static_cast<long>(N) * h * w * 4 * ((Ca + Cb) / 8))), dim3(256), 0,
                           n->stream, coarse, skip, out, N, h, w, Ca, Cb);
    };
    if ((rc = conv(n, E1A, a("x0"), a("a1"), N, H, W))) return rc;
    if ((rc = conv(n, E1B, a("a1"), a("e1"), N, H, W))) return rc;
    pool(a("e1"), a("p1"), H, W, 16);
    if ((rc = conv(n, E2A, a("p1"), a("a2"), N, H / 2, W / 2))) return rc;
    if ((rc = conv(n, E2B, a("a2"), a("e2"), N, H / 2, W / 2))) return rc;
    pool(a("e2"), a("p2"), H / 2, W / 2, 32);
    if ((rc = conv(n, E3A, a("p2"), a("a3"), N, H / 4, W / 4))) return rc;
    if ((rc = conv(n, E3B, a("a3"), a("e3"), N, H / 4, W / 4))) return rc;
    pool(a("e3"), a("p3"), H / 4, W / 4, 64);
    if ((rc = conv(n, E4A, a("p3"), a("a4"), N, H / 8, W / 8))) return rc;
    if ((rc = conv(n, E4B, a("a4"), a("e4"), N, H / 8, W / 8))) return rc;
    upcat(a("e4"), a("e3"), a("c3"), H / 8, W / 8, 128, 64);
    if ((rc = conv(n, D3A, a("c3"), a("b3"), N, H / 4, W / 4))) return rc;
    if ((rc = conv(n, D3B, a("b3"), a("d3"), N, H / 4, W / 4))) return rc;
    upcat(a("d3"), a("e2"), a("c2"), H / 4, W / 4, 64, 32);
    if ((rc = conv(n, D2A, a("c2"), a("b2"), N, H / 2, W / 2))) return rc;
    if ((rc = conv(n, D2B, a("b2"), a("d2"), N, H / 2, W / 2))) return rc;
    upcat(a("d2"), a("e1"), a("c1"), H / 2, W / 2, 32, 16);
    if ((rc = conv(n, D1A, a("c1"), a("b1"), N, H, W))) return rc;
    if ((rc = conv(n, D1B, a("b1"), a("d1"), N, H, W))) return rc;
    if ((rc = conv(n, OUT, a("d1"), a.f("o"), N, H, W, true))) return rc;
    NHIP(n, hipGetLastError());
    return FS2D_OK;
}

// Backward pass from a->dO (set by the loss); accumulates every layer's gradients.
int backward(fs2d_net* n, Acts& a) {
    const int N = a.N, H = a.H, W = a.W;
    const long p1 = static_cast<long>(N) * H * W, p2 = p1 / 4, p3 = p1 / 16, p4 = p1 / 64;
    int rc;
    bf16 *gA = a("gA"), *gB = a("gB");
    auto split = [&](const bf16* dy, bf16* da, bf16* db, long np, int Ca, int Cb) {
        hipLaunchKernelGGL(k_split, dim3(blocks_for(np * (Ca + Cb))), dim3(256), 0, n->stream, dy, da, db, np, Ca, Cb);
    };
    auto up_bwd = [&](const bf16* dy, bf16* dx, int h, int w, int C) {
        hipLaunchKernelGGL(k_upsample2_bwd, dim3(blocks_for(static_cast<long>(N) * h * w * C)), dim3(256), 0,
                           n->stream, dy, dx, N, h, w, C);
    };
    auto pool_bwd = [&](const bf16* dy, const bf16* skip, bf16* out, int h, int w, int C) {
        hipLaunchKernelGGL(k_pool_bwd_add, dim3(blocks_for(static_cast<long>(N) * h * w * C)), dim3(256), 0,
                           n->stream, dy, skip, out, N, h, w, C);
    };
    // layer step: mask dy by the layer's ReLU output y, weight gradients from
    // its input x, data gradient into dx (if given)
    auto layer_bwd = [&](int l, const bf16* x, const bf16* y, bf16* dy, bf16* dx, int h, int w) -> int {
        const Layer& L = n->layer[l];
        const long np = static_cast<long>(N) * h * w;
        int r;
        if (L.relu && (r = relu_mask(n, dy, y, np * L.cout))) return r;
        if ((r = conv_dw(n, l, x, dy, N, h, w))) return r;
        if (dx && (r = conv_dx(n, l, dy, dx, N, h, w))) return r;
        return FS2D_OK;
    };
    // level 1 decoder
    if ((rc = layer_bwd(OUT, a("d1"), nullptr, a("dO"), gA, H, W))) return rc;  // gA: dd1 (16)
    if ((rc = layer_bwd(D1B, a("b1"), a("d1"), gA, gB, H, W))) return rc;       // gB: db1 (16)
    if ((rc = layer_bwd(D1A, a("c1"), a("b1"), gB, gA, H, W))) return rc;       // gA: dc1 (48)
    split(gA, gB, a("gS1"), p1, 32, 16);                                         // gB: du1 (32), gS1: e1 skip
    up_bwd(gB, gA, H / 2, W / 2, 32);                                            // gA: dd2 (32 @ H/2)
    if ((rc = layer_bwd(D2B, a("b2"), a("d2"), gA, gB, H / 2, W / 2))) return rc;  // gB: db2 (32)
    if ((rc = layer_bwd(D2A, a("c2"), a("b2"), gB, gA, H / 2, W / 2))) return rc;  // gA: dc2 (96)
    split(gA, gB, a("gS2"), p2, 64, 32);                                           // gB: du2 (64), gS2
    up_bwd(gB, gA, H / 4, W / 4, 64);                                              // gA: dd3 (64 @ H/4)
    if ((rc = layer_bwd(D3B, a("b3"), a("d3"), gA, gB, H / 4, W / 4))) return rc;  // gB: db3
    if ((rc = layer_bwd(D3A, a("c3"), a("b3"), gB, gA, H / 4, W / 4))) return rc;  // gA: dc3 (192)
    split(gA, gB, a("gS3"), p3, 128, 64);                                          // gB: du3 (128), gS3
    up_bwd(gB, gA, H / 8, W / 8, 128);                                             // gA: de4 (128 @ H/8)
    if ((rc = layer_bwd(E4B, a("a4"), a("e4"), gA, gB, H / 8, W / 8))) return rc;  // gB: da4
    if ((rc = layer_bwd(E4A, a("p3"), a("a4"), gB, gA, H / 8, W / 8))) return rc;  // gA: dp3 (64 @ H/8)
    pool_bwd(gA, a("gS3"), gB, H / 4, W / 4, 64);                                  // gB: de3
    if ((rc = layer_bwd(E3B, a("a3"), a("e3"), gB, gA, H / 4, W / 4))) return rc;  // gA: da3
    if ((rc = layer_bwd(E3A, a("p2"), a("a3"), gA, gB, H / 4, W / 4))) return rc;  // gB: dp2 (32 @ H/4)
    pool_bwd(gB, a("gS2"), gA, H / 2, W / 2, 32);                                  // gA: de2
    if ((rc = layer_bwd(E2B, a("a2"), a("e2"), gA, gB, H / 2, W / 2))) return rc;  // gB: da2
    if ((rc = layer_bwd(E2A, a("p1"), a("a2"), gB, gA, H / 2, W / 2))) return rc;  // gA: dp1 (16 @ H/2)
    pool_bwd(gA, a("gS1"), gB, H, W, 16);                                          // gB: de1
    if ((rc = layer_bwd(E1B, a("a1"), a("e1"), gB, gA, H, W))) return rc;          // gA: da1
    if ((rc = layer_bwd(E1A, a("x0"), a("a1"), gA, nullptr, H, W))) return rc;
    (void)p4;
    NHIP(n, hipGetLastError());
    return FS2D_OK;
}

// Pack host samples: s = RMS of the previous pressure (else of the
// divergence, else 1); x0 = (d / s, prev / s, 0...) in bf16; prev and target
// in fp32 divided by s.
void pack_host(int N, int H, int W, const float* d, const float* prev, const float* target, std::vector<bf16>& x0,
               std::vector<float>& pv, std::vector<float>& tg, std::vector<float>& scale) {
    const long hw = static_cast<long>(H) * W;
    x0.assign(static_cast<size_t>(N) * hw * 16, 0);
    pv.resize(static_cast<size_t>(N) * hw);
    tg.resize(target ? static_cast<size_t>(N) * hw : 0);
    scale.resize(N);
    for (int b = 0; b < N; ++b) {
        double sp = 0, sd = 0;
        for (long q = 0; q < hw; ++q) {
            sp += static_cast<double>(prev[b * hw + q]) * prev[b * hw + q];
            sd += static_cast<double>(d[b * hw + q]) * d[b * hw + q];
        }
        double s = std::sqrt(sp / hw);
        if (!(s > 0)) s = std::sqrt(sd / hw);
        if (!(s > 0)) s = 1.0;
        const float inv = static_cast<float>(1.0 / s);
        scale[b] = static_cast<float>(s);
        for (long q = 0; q < hw; ++q) {
            const long i = b * hw + q;
            x0[i * 16 + 0] = f2bf(d[i] * inv);
            x0[i * 16 + 1] = f2bf(prev[i] * inv);
            pv[i] = prev[i] * inv;
            if (target) tg[i] = target[i] * inv;
        }
    }
}

int check_shape(fs2d_net* net, int n, int h, int w) {
    if (!net) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net: NULL network");
    // the 1/8-scale level must hold whole 16-pixel MFMA row groups
    if (n < 1 || h < 32 || w < 32 || h % 32 || w % 32 || static_cast<long>(n) * h * w > (1L << 28))
        return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net: sides must be multiples of 32 (>= 32), n >= 1");
    return FS2D_OK;
}

// ---- prediction on device fields (for the seeded projection) -------------
// Sums of squares of the interior of d and prev into scal[0], scal[1] (FP64).
__global__ __launch_bounds__(256) void k_sumsq2(const float* __restrict__ d, const float* __restrict__ prev, int R,
                                               int P, double* __restrict__ scal) {
    const int n = R - 2;
    double a = 0, b = 0;
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < static_cast<long>(n) * n;
         t += static_cast<long>(gridDim.x) * 256) {
        const long o = (t / n + 1) * P + t % n + 1;
        a += static_cast<double>(d[o]) * d[o];
        b += static_cast<double>(prev[o]) * prev[o];
    }
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    __shared__ double part[2][4];
    if ((threadIdx.x & 63) == 0) {
        part[0][threadIdx.x >> 6] = a;
        part[1][threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {  // one pair of atomics per block
        atomicAdd(scal, (part[0][0] + part[0][1]) + (part[0][2] + part[0][3]));
        atomicAdd(scal + 1, (part[1][0] + part[1][1]) + (part[1][2] + part[1][3]));
    }
}

__global__ void k_scale(double* scal, long count, float* scale) {
    double s = sqrt(scal[1] / count);
    if (!(s > 0)) s = sqrt(scal[0] / count);
    if (!(s > 0)) s = 1.0;
    scale[0] = static_cast<float>(s);
    scal[0] = scal[1] = 0.0;  // ready for the next prediction
}

__global__ __launch_bounds__(256) void k_pack_fields(const float* __restrict__ d, const float* __restrict__ prev, int R,
                                                    int P, const float* __restrict__ scale, bf16* __restrict__ x0,
                                                    float* __restrict__ pv) {
    const int n = R - 2;
    const float inv = 1.0f / scale[0];
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < static_cast<long>(n) * n;
         t += static_cast<long>(gridDim.x) * 256) {
        const long o = (t / n + 1) * P + t % n + 1;
        x0[t * 16] = f2bf(d[o] * inv);
        x0[t * 16 + 1] = f2bf(prev[o] * inv);
#pragma unroll
        for (int c = 2; c < 16; ++c) x0[t * 16 + c] = 0;
        pv[t] = prev[o] * inv;
    }
}

// p = (o + prev / s) * s on the interior, and the continuous ring
// (src/fluid_solver_cpu.cpp:33-48) written by the interior cells next to it.
__global__ __launch_bounds__(256) void k_unpack_fields(const float* __restrict__ o, const float* __restrict__ pv,
                                                      const float* __restrict__ scale, int R, int P,
                                                      float* __restrict__ p) {
    const int n = R - 2;
    const float s = scale[0];
    for (long t = static_cast<long>(blockIdx.x) * 256 + threadIdx.x; t < static_cast<long>(n) * n;
         t += static_cast<long>(gridDim.x) * 256) {
        const int i = static_cast<int>(t / n) + 1, j = static_cast<int>(t % n) + 1;
        const float v = (o[t * 16] + pv[t]) * s;
        p[static_cast<long>(i) * P + j] = v;
        if (i == 1) p[j] = v;
        if (i == n) p[static_cast<long>(R - 1) * P + j] = v;
        if (j == 1) p[static_cast<long>(i) * P] = v;
        if (j == n) p[static_cast<long>(i) * P + R - 1] = v;
        if (i == 1 && j == 1) p[0] = v;
        if (i == 1 && j == n) p[R - 1] = v;
        if (i == n && j == 1) p[static_cast<long>(R - 1) * P] = v;
        if (i == n && j == n) p[static_cast<long>(R - 1) * P + R - 1] = v;
    }
}

}  // namespace

namespace fs2d {

// The network's prediction from the handle's divergence d and previous
// pressure prev (pitched R x R fields) into p, enqueued on `stream` (the
// handle's), which then waits for nothing else.
int net_predict_fields(fs2d_net* net, hipStream_t stream, const float* d, const float* prev, int R, int P, float* p) {
    const int n = R - 2;
    int rc;
    if ((rc = check_shape(net, 1, n, n))) return rc;
    Acts* a = nullptr;
    if ((rc = acts_for(net, 1, n, n, &a))) return rc;
    hipStream_t saved = net->stream;
    net->stream = stream;
    const long cells = static_cast<long>(n) * n;
    hipLaunchKernelGGL(k_sumsq2, dim3(std::min(blocks_for(cells), 1024u)), dim3(256), 0, stream, d, prev, R, P,
                       net->scal);
    hipLaunchKernelGGL(k_scale, dim3(1), dim3(1), 0, stream, net->scal, cells, a->f("scale"));
    hipLaunchKernelGGL(k_pack_fields, dim3(blocks_for(cells)), dim3(256), 0, stream, d, prev, R, P, a->f("scale"),
                       (*a)("x0"), a->f("prev"));
    rc = forward(net, *a);
    if (rc == FS2D_OK)
        hipLaunchKernelGGL(k_unpack_fields, dim3(blocks_for(cells)), dim3(256), 0, stream, a->f("o"), a->f("prev"),
                           a->f("scale"), R, P, p);
    net->stream = saved;
    if (rc) return rc;
    NHIP(net, hipGetLastError());
    return FS2D_OK;
}

const char* net_error(const fs2d_net* net) { return net ? net->err.c_str() : g_net_create_error.c_str(); }

// Exceptions thrown on the host (allocation in plan building, maps and
// vectors) never cross the C ABI: they become an error code and message.
template <class F>
int guarded(fs2d_net* h, const char* what, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return nfail(h, FS2D_ERR_HIP, std::string(what) + ": out of host memory");
    } catch (const std::exception& e) {
        return nfail(h, FS2D_ERR_STATE, std::string(what) + ": " + e.what());
    } catch (...) {
        return nfail(h, FS2D_ERR_STATE, std::string(what) + ": unexpected exception");
    }
}

}  // namespace fs2d

extern "C" {

int fs2d_net_destroy(fs2d_net* net) {
    if (!net) return FS2D_OK;
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    for (auto& kv : net->acts) {
        for (auto& b : kv.second->buf) (void)hipFree(b.second);
        delete kv.second;
    }
    for (float* p : {net->params, net->grads, net->m1, net->m2, net->loss})
        if (p) (void)hipFree(p);
    if (net->scal) (void)hipFree(net->scal);
    for (auto& L : net->layer) {
        if (L.w16) (void)hipFree(L.w16);
        if (L.wt16) (void)hipFree(L.wt16);
    }
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
    return FS2D_OK;
}

static int fs2d_net_create_impl(uint64_t seed, fs2d_net** out) {
    if (!out) return nfail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_net_create: out is NULL");
    *out = nullptr;
    fs2d_net* net = new fs2d_net();
    long total = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        Layer& L = net->layer[l];
        L.cin = kLayers[l].cin;
        L.cout = kLayers[l].cout;
        L.relu = kLayers[l].relu;
        L.wn = 9L * L.cin * L.cout;
        total += L.count();
    }
    net->total = total;
    auto bail = [&](hipError_t e, const char* what) {
        std::string msg = std::string("fs2d_net_create: ") + what + ": " + hipGetErrorString(e);
        fs2d_net_destroy(net);
        return nfail(nullptr, FS2D_ERR_HIP, msg);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking)) != hipSuccess) return bail(e, "stream");
    for (float** p : {&net->params, &net->grads, &net->m1, &net->m2})
        if ((e = hipMalloc(p, sizeof(float) * total)) != hipSuccess) return bail(e, "hipMalloc(parameters)");
    if ((e = hipMalloc(&net->loss, sizeof(float))) != hipSuccess) return bail(e, "hipMalloc(loss)");
    if ((e = hipMalloc(&net->scal, 2 * sizeof(double))) != hipSuccess) return bail(e, "hipMalloc(scal)");
    if ((e = hipMemsetAsync(net->scal, 0, 2 * sizeof(double), net->stream)) != hipSuccess) return bail(e, "memset");
    // He initialisation on the real fan-in; padded channels start at zero
    std::vector<float> host(total, 0.f);
    Rng rng{seed ^ 0x5f3759df00c0ffeeull};
    long off = 0;
    for (int l = 0; l < kNumLayers; ++l) {
        Layer& L = net->layer[l];
        L.p = net->params + off;
        L.g = net->grads + off;
        L.m = net->m1 + off;
        L.v = net->m2 + off;
        const int ci_r = real_cin(l), co_r = real_cout(l);
        const double sd = std::sqrt(2.0 / (9.0 * ci_r)) * (l == OUT ? 0.1 : 1.0);
        for (int tap = 0; tap < 9; ++tap)
            for (int co = 0; co < co_r; ++co)
                for (int ci = 0; ci < ci_r; ++ci)
                    host[off + (static_cast<long>(tap) * L.cout + co) * L.cin + ci] = static_cast<float>(sd * rng.normal());
        off += L.count();
        if ((e = hipMalloc(&L.w16, sizeof(bf16) * L.wn)) != hipSuccess) return bail(e, "hipMalloc(bf16 weights)");
        if ((e = hipMalloc(&L.wt16, sizeof(bf16) * L.wn)) != hipSuccess) return bail(e, "hipMalloc(bf16 weights)");
    }
    if ((e = hipMemcpy(net->params, host.data(), sizeof(float) * total, hipMemcpyHostToDevice)) != hipSuccess)
        return bail(e, "hipMemcpy(parameters)");
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks_for(total)), dim3(256), 0, net->stream, net->m1, total);
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks_for(total)), dim3(256), 0, net->stream, net->m2, total);
    if (refresh_bf16(net)) return bail(hipErrorUnknown, net->err.c_str());
    if ((e = hipStreamSynchronize(net->stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
    *out = net;
    return FS2D_OK;
}

int fs2d_net_create(uint64_t seed, fs2d_net** out) {
    return guarded(nullptr, "fs2d_net_create", [&] { return fs2d_net_create_impl(seed, out); });
}

static int fs2d_net_param_count_impl(fs2d_net* net, int64_t* count) {
    if (!net || !count) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net_param_count: bad argument");
    *count = net->total;
    return FS2D_OK;
}

int fs2d_net_param_count(fs2d_net* net, int64_t* count) {
    return guarded(net, "fs2d_net_param_count", [&] { return fs2d_net_param_count_impl(net, count); });
}

static int fs2d_net_get_params_impl(fs2d_net* net, float* host) {
    if (!net || !host) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net_get_params: bad argument");
    NHIP(net, hipStreamSynchronize(net->stream));
    NHIP(net, hipMemcpy(host, net->params, sizeof(float) * net->total, hipMemcpyDeviceToHost));
    return FS2D_OK;
}

int fs2d_net_get_params(fs2d_net* net, float* host) {
    return guarded(net, "fs2d_net_get_params", [&] { return fs2d_net_get_params_impl(net, host); });
}

static int fs2d_net_get_grads_impl(fs2d_net* net, float* host) {
    if (!net || !host) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net_get_grads: bad argument");
    NHIP(net, hipStreamSynchronize(net->stream));
    NHIP(net, hipMemcpy(host, net->grads, sizeof(float) * net->total, hipMemcpyDeviceToHost));
    return FS2D_OK;
}

int fs2d_net_get_grads(fs2d_net* net, float* host) {
    return guarded(net, "fs2d_net_get_grads", [&] { return fs2d_net_get_grads_impl(net, host); });
}

static int fs2d_net_set_params_impl(fs2d_net* net, const float* host) {
    if (!net || !host) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net_set_params: bad argument");
    NHIP(net, hipStreamSynchronize(net->stream));
    NHIP(net, hipMemcpy(net->params, host, sizeof(float) * net->total, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks_for(net->total)), dim3(256), 0, net->stream, net->m1, net->total);
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks_for(net->total)), dim3(256), 0, net->stream, net->m2, net->total);
    net->adam_t = 0;
    int rc = refresh_bf16(net);
    if (rc) return rc;
    NHIP(net, hipStreamSynchronize(net->stream));
    return FS2D_OK;
}

int fs2d_net_set_params(fs2d_net* net, const float* host) {
    return guarded(net, "fs2d_net_set_params", [&] { return fs2d_net_set_params_impl(net, host); });
}

static int fs2d_net_train_step_impl(fs2d_net* net, int n, int h, int w, const float* d, const float* prev, const float* target,
                        float lr, float* loss) {
    int rc;
    if ((rc = check_shape(net, n, h, w))) return rc;
    if (!d || !prev || !target || !loss) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net_train_step: NULL argument");
    Acts* a = nullptr;
    if ((rc = acts_for(net, n, h, w, &a))) return rc;
    std::vector<bf16> x0;
    std::vector<float> pv, tg, sc;
    pack_host(n, h, w, d, prev, target, x0, pv, tg, sc);
    const long npix = static_cast<long>(n) * h * w;
    NHIP(net, hipMemcpyAsync((*a)("x0"), x0.data(), sizeof(bf16) * x0.size(), hipMemcpyHostToDevice, net->stream));
    NHIP(net, hipMemcpyAsync(a->f("prev"), pv.data(), sizeof(float) * pv.size(), hipMemcpyHostToDevice, net->stream));
    NHIP(net, hipMemcpyAsync(a->f("tgt"), tg.data(), sizeof(float) * tg.size(), hipMemcpyHostToDevice, net->stream));
    if ((rc = forward(net, *a))) return rc;
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(256), 0, net->stream, net->loss, 1L);
    hipLaunchKernelGGL(k_mse, dim3(blocks_for(npix)), dim3(256), 0, net->stream, a->f("o"), a->f("prev"), a->f("tgt"),
                       npix, (*a)("dO"), net->loss);
    hipLaunchKernelGGL(k_zero_f32, dim3(blocks_for(net->total)), dim3(256), 0, net->stream, net->grads, net->total);
    if ((rc = backward(net, *a))) return rc;
    ++net->adam_t;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, static_cast<float>(net->adam_t));
    const float c2 = 1.f - std::pow(b2, static_cast<float>(net->adam_t));
    for (auto& L : net->layer) {
        hipLaunchKernelGGL(k_adam, dim3(blocks_for(L.wn)), dim3(256), 0, net->stream, L.p, L.g, L.m, L.v, L.w16, L.wn,
                           lr, b1, b2, 1e-8f, c1, c2);
        hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, net->stream, L.p + L.wn, L.g + L.wn, L.m + L.wn,
                           L.v + L.wn, static_cast<bf16*>(nullptr), static_cast<long>(L.cout), lr, b1, b2, 1e-8f, c1,
                           c2);
    }
    NHIP(net, hipGetLastError());
    float sum = 0.f;
    NHIP(net, hipMemcpyAsync(&sum, net->loss, sizeof(float), hipMemcpyDeviceToHost, net->stream));
    NHIP(net, hipStreamSynchronize(net->stream));
    *loss = sum / static_cast<float>(npix);
    return FS2D_OK;
}

int fs2d_net_train_step(fs2d_net* net, int n, int h, int w, const float* d, const float* prev, const float* target,
                        float lr, float* loss) {
    return guarded(net, "fs2d_net_train_step", [&] { return fs2d_net_train_step_impl(net, n, h, w, d, prev, target, lr, loss); });
}

static int fs2d_net_predict_host_impl(fs2d_net* net, int n, int h, int w, const float* d, const float* prev, float* out) {
    int rc;
    if ((rc = check_shape(net, n, h, w))) return rc;
    if (!d || !prev || !out) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net_predict_host: NULL argument");
    Acts* a = nullptr;
    if ((rc = acts_for(net, n, h, w, &a))) return rc;
    std::vector<bf16> x0;
    std::vector<float> pv, tg, sc;
    pack_host(n, h, w, d, prev, nullptr, x0, pv, tg, sc);
    const long npix = static_cast<long>(n) * h * w;
    NHIP(net, hipMemcpyAsync((*a)("x0"), x0.data(), sizeof(bf16) * x0.size(), hipMemcpyHostToDevice, net->stream));
    if ((rc = forward(net, *a))) return rc;
    std::vector<float> o(static_cast<size_t>(npix) * 16);
    NHIP(net, hipMemcpyAsync(o.data(), a->f("o"), sizeof(float) * o.size(), hipMemcpyDeviceToHost, net->stream));
    NHIP(net, hipStreamSynchronize(net->stream));
    const long hw = static_cast<long>(h) * w;
    for (long i = 0; i < npix; ++i) out[i] = (o[i * 16] + pv[i]) * sc[i / hw];
    return FS2D_OK;
}

int fs2d_net_predict_host(fs2d_net* net, int n, int h, int w, const float* d, const float* prev, float* out) {
    return guarded(net, "fs2d_net_predict_host", [&] { return fs2d_net_predict_host_impl(net, n, h, w, d, prev, out); });
}

static int fs2d_net_time_forward_impl(fs2d_net* net, int n, int h, int w, int reps, float* ms) {
    int rc;
    if ((rc = check_shape(net, n, h, w))) return rc;
    if (reps < 1 || !ms) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net_time_forward: bad argument");
    auto it = net->acts.find(std::make_tuple(n, h, w));
    if (it == net->acts.end()) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net_time_forward: no activations of that shape");
    hipEvent_t e0, e1;
    NHIP(net, hipEventCreate(&e0));
    NHIP(net, hipEventCreate(&e1));
    if ((rc = forward(net, *it->second))) return rc;  // warm
    NHIP(net, hipEventRecord(e0, net->stream));
    for (int r = 0; r < reps && rc == FS2D_OK; ++r) rc = forward(net, *it->second);
    NHIP(net, hipEventRecord(e1, net->stream));
    NHIP(net, hipEventSynchronize(e1));
    float t = 0.f;
    NHIP(net, hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *ms = t / reps;
    return FS2D_OK;
}

int fs2d_net_time_forward(fs2d_net* net, int n, int h, int w, int reps, float* ms) {
    return guarded(net, "fs2d_net_time_forward", [&] { return fs2d_net_time_forward_impl(net, n, h, w, reps, ms); });
}

static int fs2d_net_forward_flops_impl(fs2d_net* net, int n, int h, int w, double* flops) {
    int rc;
    if ((rc = check_shape(net, n, h, w))) return rc;
    if (!flops) return nfail(net, FS2D_ERR_ARGUMENT, "fs2d_net_forward_flops: bad argument");
    // pixels of each layer's level: e1*, d1*, out at 1; e2*, d2* at 1/4; e3*, d3* at 1/16; e4* at 1/64
    const int level[kNumLayers] = {0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0};
    double f = 0;
    const double p1 = static_cast<double>(n) * h * w;
    for (int l = 0; l < kNumLayers; ++l)
        f += 2.0 * 9.0 * kLayers[l].cin * kLayers[l].cout * p1 / static_cast<double>(1 << (2 * level[l]));
    *flops = f;
    return FS2D_OK;
}

int fs2d_net_forward_flops(fs2d_net* net, int n, int h, int w, double* flops) {
    return guarded(net, "fs2d_net_forward_flops", [&] { return fs2d_net_forward_flops_impl(net, n, h, w, flops); });
}

void* fs2d_net_stream(fs2d_net* net) { return net ? static_cast<void*>(net->stream) : nullptr; }

const char* fs2d_net_last_error(const fs2d_net* net) { return net_error(net); }

}  // extern "C"
