// fs2d_batch.hip -- many independent simulations of different sizes, one
// launch per stage (declared in include/fs2d_hip.h, fs2d_hip_batch_*).
//
// Layout: every field kind (density, u, v, the sources, the diffuse and
// pressure iterates and right-hand sides, the Jacobi scratch pairs) is one
// arena holding all simulations' grids; simulation s occupies the same range
// [off_s, off_s + R_s * P_s) in every arena, with its own row pitch P_s (R_s
// rounded up to 64 floats).  Swapping two arenas therefore swaps that role
// for every simulation at once, so the buffer rotation of the step stays a
// host-side permutation of arena pointers, identical for all simulations.
//
// Each stage is one launch over a work list: the single-pass stages over
// (simulation, 64 x 4 cell block) entries, the Jacobi sweeps over
// (simulation, field, tile) entries with the register-strip tile of the
// one-grid solver.  Simulation s runs its K_s sweeps as K_s / M launches of M
// fused sweeps; launch j covers the simulations with more than j launches
// (the work list is ordered by launch count, so that is a prefix of it), and
// ping-pongs home -> S1 -> S2 -> S1 ... -> home, so that every solve ends in
// its home field whatever its launch count (K_s / M >= 2).
//
// The step follows fluid_solver_cpu::solve (src/fluid_solver_cpu.cpp:6-31)
// in the Jacobi order, exactly as fs2d_handle's step does: each simulation's
// fields come out bit-identical to a one-grid handle with the same inputs.
#include "../../include/fs2d_hip.h"
#include "fs2d_cells.hpp"
#include "fs2d_internal.hpp"
#include "fs2d_jacobi.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

using namespace fs2d;

namespace {

thread_local std::string g_batch_create_error;

// Arena roles.  The diffuse solves of density, u and v use (home, rhs,
// scratch pair) = (RHO, X0, S1A/S2A), (UN, X0U, S1B/S2B), (VN, X0V, S1C/S2C);
// the pressure solve uses (P, D, S1A/S2A).
enum {
    A_RHO, A_U, A_V, A_RS, A_US, A_VS, A_X0, A_X0U, A_X0V, A_UN, A_VN, A_T1, A_T2, A_D, A_P,
    A_S1A, A_S2A, A_S1B, A_S2B, A_S1C, A_S2C, A_COUNT
};

struct Roles {
    float* a[A_COUNT];
};

// Per-simulation constants, computed on the host exactly as the one-grid
// handle computes them.
struct BatchSim {
    size_t off;  // floats, 256-byte aligned
    int R, P;
    int launches;     // K / M Jacobi launches per solve (diffuse and pressure)
    int cell_blocks;  // 64 x 4 blocks over the interior
    float dt, dt0, h, neg_half_h;
    float a_rho, denom_rho, inv_rho, lo_rho, inv_lo_rho;
    float a_vel, denom_vel, inv_vel, lo_vel, inv_lo_vel;
};

constexpr int CBX = 64, CBY = 4;      // single-pass stage blocks
constexpr int kSimSlots = 16;         // speed-reduction slots per simulation
constexpr int kSimSlotStride = 32;    // words between slots (separate cache lines)

__device__ __forceinline__ int xcd_block(int b, int nb) { return jac::xcd_block(b, nb); }

// Single-pass stages are persistent: block b walks entries b, b + G, ... of
// the work list (G blocks, a multiple of 8: each block's entries then map,
// through xcd_block, into one contiguous run of the list per XCD).
// Entry e: x = simulation, y = (by << 16) | bx.
__device__ __forceinline__ bool cell_of(const BatchSim* __restrict__ sims, const int2* __restrict__ work, int e,
                                        int nb, int& s, int& i, int& j, int& blk) {
    const int2 w = work[xcd_block(e, nb)];
    s = w.x;
    const int bx = w.y & 0xffff, by = w.y >> 16;
    j = bx * CBX + threadIdx.x + 1;
    i = by * CBY + threadIdx.y + 1;
    blk = w.y;
    const int R = sims[s].R;
    return i <= R - 2 && j <= R - 2;
}

__global__ __launch_bounds__(CBX* CBY) void kb_add_sources(const BatchSim* __restrict__ sims,
                                                           const int2* __restrict__ work, int nwork, Roles r) {
    for (int e = blockIdx.x; e < nwork; e += gridDim.x) {
        int s, i, j, blk;
        if (!cell_of(sims, work, e, nwork, s, i, j, blk)) continue;
        const BatchSim& S = sims[s];
        const size_t o = S.off;
        cells::add_sources(r.a[A_RHO] + o, r.a[A_X0] + o, r.a[A_RS] + o, S.P, S.dt, i, j);
        cells::add_sources_split(r.a[A_UN] + o, r.a[A_X0U] + o, r.a[A_U] + o, r.a[A_US] + o, S.R, S.P, S.dt, i, j);
        cells::add_sources_split(r.a[A_VN] + o, r.a[A_X0V] + o, r.a[A_V] + o, r.a[A_VS] + o, S.R, S.P, S.dt, i, j);
    }
}

// Block maximum of the largest finite |u|, |v| into one of the simulation's
// slots (k_fold folds them).  Ends with a barrier, so a persistent block may
// reduce its next entry at once.
__device__ __forceinline__ void speed_to_slot(float m, unsigned* __restrict__ slots, int s, int blk) {
    m = cells::block_max(m);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
        const unsigned slot = (unsigned)((blk & 0xffff) + 131 * (blk >> 16)) % kSimSlots;
        atomicMax(slots + ((size_t)s * kSimSlots + slot) * kSimSlotStride, __float_as_uint(m));
    }
    __syncthreads();
}

__global__ __launch_bounds__(CBX* CBY) void kb_max_speed(const BatchSim* __restrict__ sims,
                                                         const int2* __restrict__ work, int nwork, Roles r,
                                                         unsigned* __restrict__ slots) {
    for (int e = blockIdx.x; e < nwork; e += gridDim.x) {
        int s, i, j, blk;
        const bool in = cell_of(sims, work, e, nwork, s, i, j, blk);
        float m = 0.f;
        if (in) {
            const size_t o = sims[s].off + (size_t)i * sims[s].P + j;
            m = cells::fold_speed(0.f, r.a[A_U][o], r.a[A_V][o]);
        }
        speed_to_slot(m, slots, s, blk);
    }
}

// One 64-lane block per simulation: fold its slots into max_bits[s] and clear them.
__global__ __launch_bounds__(64) void kb_fold_speed(unsigned* __restrict__ max_bits, unsigned* __restrict__ slots) {
    const int s = blockIdx.x, t = threadIdx.x;
    unsigned m = 0u;
    if (t < kSimSlots) {
        unsigned* p = slots + ((size_t)s * kSimSlots + t) * kSimSlotStride;
        m = *p;
        *p = 0u;
    }
    for (int off = 32; off > 0; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
    if (t == 0) max_bits[s] = m;
}

__global__ __launch_bounds__(CBX* CBY) void kb_gather_velocity(const BatchSim* __restrict__ sims,
                                                               const int2* __restrict__ work, int nwork, Roles r) {
    for (int e = blockIdx.x; e < nwork; e += gridDim.x) {
        int s, i, j, blk;
        if (!cell_of(sims, work, e, nwork, s, i, j, blk)) continue;
        const BatchSim& S = sims[s];
        const size_t o = S.off;
        cells::gather_velocity(r.a[A_T1] + o, r.a[A_T2] + o, r.a[A_U] + o, r.a[A_V] + o, S.R, S.P, S.dt0, i, j);
    }
}

__global__ __launch_bounds__(CBX* CBY) void kb_divergence(const BatchSim* __restrict__ sims,
                                                          const int2* __restrict__ work, int nwork, Roles r) {
    for (int e = blockIdx.x; e < nwork; e += gridDim.x) {
        int s, i, j, blk;
        if (!cell_of(sims, work, e, nwork, s, i, j, blk)) continue;
        const BatchSim& S = sims[s];
        const size_t o = S.off;
        // the pressure iterate needs no zero fill: its first Jacobi launch reads zeros itself
        cells::divergence(r.a[A_D] + o, r.a[A_U] + o, r.a[A_V] + o, S.R, S.P, S.neg_half_h, nullptr, i, j);
    }
}

__global__ __launch_bounds__(CBX* CBY) void kb_gradient(const BatchSim* __restrict__ sims,
                                                        const int2* __restrict__ work, int nwork, Roles r,
                                                        unsigned* __restrict__ slots) {
    for (int e = blockIdx.x; e < nwork; e += gridDim.x) {
        int s, i, j, blk;
        const bool in = cell_of(sims, work, e, nwork, s, i, j, blk);
        float m = 0.f;
        if (in) {
            const BatchSim& S = sims[s];
            const size_t o = S.off;
            m = cells::gradient(r.a[A_U] + o, r.a[A_V] + o, r.a[A_P] + o, S.R, S.P, S.h, false, i, j);
        }
        speed_to_slot(m, slots, s, blk);
    }
}

// Jacobi launch j of M fused sweeps over (simulation, field, tile) entries:
// x = simulation | field << 24, y = tile (row-major over tiles_y x tiles_x).
// Persistent: wave g of the grid walks entries g, g + waves, ... (waves = all
// of the grid's), each entry one tile of one simulation's field; no
// block-wide barrier, so a wave leaves when its entries are done.
template <int M, int TY, int KIND>
__device__ __forceinline__ void kb_jacobi_entry(const BatchSim* __restrict__ sims, const int2 w, Roles r, int j);

template <int M, int TY, int KIND>
__global__ __launch_bounds__(64 * jac::JW_WAVES) void kb_jacobi(const BatchSim* __restrict__ sims,
                                                               const int2* __restrict__ work, int n_work, Roles r,
                                                               int j) {
    const int waves = (int)gridDim.x * jac::JW_WAVES;
    for (int t = xcd_block(blockIdx.x, gridDim.x) * jac::JW_WAVES + threadIdx.x / 64; t < n_work; t += waves)
        kb_jacobi_entry<M, TY, KIND>(sims, work[t], r, j);
}

template <int M, int TY, int KIND>
__device__ __forceinline__ void kb_jacobi_entry(const BatchSim* __restrict__ sims, const int2 w, Roles r, int j) {
    const int s = w.x & 0xffffff, f = w.x >> 24;
    const BatchSim& S = sims[s];
    constexpr int OX = 64 - 2 * M;
    const int n = S.R - 2;
    const int tiles_x = (n + OX - 1) / OX;
    const int ty = w.y / tiles_x, tx = w.y - ty * tiles_x;
    const size_t o = S.off;
    SweepArgs A{};
    A.R = S.R;
    A.P = S.P;
    A.kind = KIND;
    float *home, *s1, *s2;
    if (KIND == SWEEP_PRESSURE) {
        home = r.a[A_P] + o;
        A.b = r.a[A_D] + o;
        s1 = r.a[A_S1A] + o;
        s2 = r.a[A_S2A] + o;
        A.ring = ring_of(BND_CONT);
    } else {
        const bool rho = f == 0;
        home = r.a[f == 0 ? A_RHO : (f == 1 ? A_UN : A_VN)] + o;
        A.b = r.a[f == 0 ? A_X0 : (f == 1 ? A_X0U : A_X0V)] + o;
        s1 = r.a[f == 0 ? A_S1A : (f == 1 ? A_S1B : A_S1C)] + o;
        s2 = r.a[f == 0 ? A_S2A : (f == 1 ? A_S2B : A_S2C)] + o;
        A.a = rho ? S.a_rho : S.a_vel;
        A.denom = rho ? S.denom_rho : S.denom_vel;
        A.inv_denom = rho ? S.inv_rho : S.inv_vel;
        A.inv_denom_lo = rho ? S.inv_lo_rho : S.inv_lo_vel;
        A.div_lo = rho ? S.lo_rho : S.lo_vel;
        A.ring = ring_of(f == 0 ? BND_CONT : (f == 1 ? BND_OPP_H : BND_OPP_V));
    }
    const float* src = (j == 0) ? home : (((j - 1) & 1) ? s2 : s1);
    float* dst = (j == S.launches - 1) ? home : ((j & 1) ? s2 : s1);
    const bool zero_src = KIND == SWEEP_PRESSURE && j == 0;
    const int gi0 = 1 + TY * ty - M, gj0 = 1 + OX * tx - M;
    const int R = S.R;
    const bool plain = gi0 >= 1 && gi0 + TY + 2 * M - 1 <= R - 2 && gj0 >= 1 && gj0 + 63 <= R - 2;
    const int lane = threadIdx.x & 63;
    if (plain)
        jac::jacobi_strip<M, TY, KIND, false>(A, src, dst, zero_src, gi0, gj0, lane);
    else
        jac::jacobi_strip<M, TY, KIND, true>(A, src, dst, zero_src, gi0, gj0, lane);
}

__global__ __launch_bounds__(256) void kb_fill(uint32_t* __restrict__ p, size_t n) {
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256) p[t] = 0u;
}

// Candidate (depth, tile rows) pairs; the tuner keeps the fastest.
constexpr int kDepths[] = {1, 2, 4, 8};
constexpr int kTileRows[] = {16, 32};

// Blocks of a persistent launch over `n` entries of `per_block` each, at most
// `limit` (a multiple of 8 when it binds, so that every block's entries fall
// in one XCD's run of the list).
unsigned persistent_grid(int n, int per_block, int limit) {
    const int need = (n + per_block - 1) / per_block;
    if (limit <= 0 || need <= limit) return static_cast<unsigned>(std::max(need, 1));
    return static_cast<unsigned>(std::max(8, limit - limit % 8));
}

template <int M, int TY>
void launch_kb_jacobi_mt(hipStream_t st, int kind, const BatchSim* sims, const int2* w, int n, Roles r, int j,
                         int limit) {
    const dim3 grid(persistent_grid(n, jac::JW_WAVES, limit));
    if (kind == SWEEP_DIFFUSE)
        hipLaunchKernelGGL((kb_jacobi<M, TY, SWEEP_DIFFUSE>), grid, dim3(64 * jac::JW_WAVES), 0, st, sims, w, n, r, j);
    else
        hipLaunchKernelGGL((kb_jacobi<M, TY, SWEEP_PRESSURE>), grid, dim3(64 * jac::JW_WAVES), 0, st, sims, w, n, r,
                           j);
}

template <int TY>
bool launch_kb_jacobi_t(hipStream_t st, int M, int kind, const BatchSim* sims, const int2* w, int n, Roles r, int j,
                        int limit) {
    switch (M) {
        case 1: launch_kb_jacobi_mt<1, TY>(st, kind, sims, w, n, r, j, limit); return true;
        case 2: launch_kb_jacobi_mt<2, TY>(st, kind, sims, w, n, r, j, limit); return true;
        case 4: launch_kb_jacobi_mt<4, TY>(st, kind, sims, w, n, r, j, limit); return true;
        case 8: launch_kb_jacobi_mt<8, TY>(st, kind, sims, w, n, r, j, limit); return true;
        default: return false;
    }
}

hipError_t launch_kb_jacobi(hipStream_t st, int M, int TY, int kind, const BatchSim* sims, const int2* w, int n,
                            Roles r, int j, int limit) {
    if (n <= 0) return hipSuccess;
    const bool ok = TY == 16 ? launch_kb_jacobi_t<16>(st, M, kind, sims, w, n, r, j, limit)
                             : (TY == 32 ? launch_kb_jacobi_t<32>(st, M, kind, sims, w, n, r, j, limit) : false);
    return ok ? hipGetLastError() : hipErrorInvalidValue;
}

}  // namespace

// A Jacobi configuration: depth M, tile rows TY, and the work lists of the
// diffuse (3 fields) and pressure (1 field) solves with, for each launch j,
// the number of leading entries whose simulation runs launch j.
struct JacobiConfig {
    int M = 0, TY = 0;
    int2* diffuse_work = nullptr;
    int2* pressure_work = nullptr;
    std::vector<int> diffuse_count, pressure_count;  // per launch j
};

struct BatchGraph {
    float* pre[A_COUNT];
    float* post[A_COUNT];
    bool pre_speed;
    hipGraph_t graph = nullptr;  // kept while its executable lives (tools inspect it at launch)
    hipGraphExec_t exec = nullptr;
};

struct fs2d_batch {
    int count = 0;
    fs2d_params prm{};
    std::vector<fs2d_batch_sim> spec;
    std::vector<BatchSim> host_sims;  // launches filled in for the chosen depth
    std::vector<int> order;           // simulations by Jacobi launch count, descending
    size_t arena_floats = 0;
    long long interior = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    Roles roles{};
    BatchSim* sims = nullptr;     // device copy of host_sims
    int2* cell_work = nullptr;    // single-pass stage blocks, all simulations
    int cell_blocks = 0;
    unsigned* max_bits = nullptr;  // per simulation
    unsigned* slots = nullptr;     // per simulation kSimSlots * kSimSlotStride words, zero between uses
    ScatterSim* scatter_sims = nullptr;  // the density scatter's descriptors (fs2d_scatter.hip)
    int2* scatter_windows = nullptr;     // its window path's work list, one entry per wave
    void* scatter_scratch = nullptr;     // its sorted path's buffers
    int* scatter_flag = nullptr;         // its sorted-path gate
    BatchScatter scatter{};
    int cus = 256;                // compute units of the device
    int jac_per_cu = 0;           // persistent Jacobi blocks per CU (0 until tuned: one block per 4 entries)
    int cell_per_cu = 0;          // persistent single-pass blocks per CU (0 until tuned: one per entry)
    bool occupancy_fixed = false; // set by fs2d_hip_batch_set_occupancy: the tuner keeps them
    JacobiConfig cfg;
    bool tuned = false;
    bool speed_valid = false;
    long long steps_done = 0;
    std::vector<BatchGraph*> graphs;
    std::string err;
};

namespace {

int bfail(fs2d_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    else g_batch_create_error = msg;
    return code;
}

#define FS2D_BHIP(b, expr)                                                                     \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return bfail((b), FS2D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

float cells_of(int R) { return static_cast<float>(static_cast<size_t>(R) * static_cast<size_t>(R)); }

int jac_limit(const fs2d_batch* b) { return b->jac_per_cu > 0 ? b->cus * b->jac_per_cu : 0; }
dim3 cell_grid(const fs2d_batch* b) { return dim3(persistent_grid(b->cell_blocks, 1, b->cell_per_cu > 0 ? b->cus * b->cell_per_cu : 0)); }

void free_config(JacobiConfig& c) {
    if (c.diffuse_work) (void)hipFree(c.diffuse_work);
    if (c.pressure_work) (void)hipFree(c.pressure_work);
    c = JacobiConfig{};
}

// Work lists of depth M and tile rows TY.  Entries are ordered by the
// simulation's launch count (descending), so launch j takes a prefix.
int build_config(fs2d_batch* b, int M, int TY, JacobiConfig* out) {
    JacobiConfig c;
    c.M = M;
    c.TY = TY;
    const int OX = 64 - 2 * M;
    std::vector<int2> dw, pw;
    std::vector<int> launches(b->count);
    int max_l = 0;
    for (int s = 0; s < b->count; ++s) {
        launches[s] = b->spec[s].iterations / M;
        max_l = std::max(max_l, launches[s]);
    }
    c.diffuse_count.assign(max_l, 0);
    c.pressure_count.assign(max_l, 0);
    for (int s : b->order) {
        const int n = b->spec[s].rows - 2;
        const int tiles = ((n + OX - 1) / OX) * ((n + TY - 1) / TY);
        for (int f = 0; f < 3; ++f)
            for (int t = 0; t < tiles; ++t) dw.push_back(make_int2(s | (f << 24), t));
        for (int t = 0; t < tiles; ++t) pw.push_back(make_int2(s, t));
        for (int j = 0; j < launches[s]; ++j) {
            c.diffuse_count[j] += 3 * tiles;
            c.pressure_count[j] += tiles;
        }
    }
    FS2D_BHIP(b, hipMalloc(&c.diffuse_work, sizeof(int2) * dw.size()));
    FS2D_BHIP(b, hipMalloc(&c.pressure_work, sizeof(int2) * pw.size()));
    FS2D_BHIP(b, hipMemcpy(c.diffuse_work, dw.data(), sizeof(int2) * dw.size(), hipMemcpyHostToDevice));
    FS2D_BHIP(b, hipMemcpy(c.pressure_work, pw.data(), sizeof(int2) * pw.size(), hipMemcpyHostToDevice));
    *out = c;
    return FS2D_OK;
}

// Upload the per-simulation constants with launch counts for depth M.
int upload_sims(fs2d_batch* b, int M) {
    for (int s = 0; s < b->count; ++s) b->host_sims[s].launches = b->spec[s].iterations / M;
    // on the batch's stream: kernels already queued there still read the descriptors
    FS2D_BHIP(b, hipMemcpyAsync(b->sims, b->host_sims.data(), sizeof(BatchSim) * b->count, hipMemcpyHostToDevice,
                                b->stream));
    FS2D_BHIP(b, hipStreamSynchronize(b->stream));
    return FS2D_OK;
}

bool depth_fits(const fs2d_batch* b, int M) {
    for (const auto& s : b->spec)
        if (s.iterations % M != 0 || s.iterations / M < 2) return false;
    return true;
}

int solve(fs2d_batch* b, int kind) {
    const JacobiConfig& c = b->cfg;
    const auto& counts = kind == SWEEP_DIFFUSE ? c.diffuse_count : c.pressure_count;
    const int2* w = kind == SWEEP_DIFFUSE ? c.diffuse_work : c.pressure_work;
    for (int j = 0; j < static_cast<int>(counts.size()); ++j)
        FS2D_BHIP(b, launch_kb_jacobi(b->stream, c.M, c.TY, kind, b->sims, w, counts[j], b->roles, j, jac_limit(b)));
    return FS2D_OK;
}

hipError_t cell_launch_check() { return hipGetLastError(); }

int project(fs2d_batch* b) {
    const dim3 grid = cell_grid(b), blk(CBX, CBY);
    hipLaunchKernelGGL(kb_divergence, grid, blk, 0, b->stream, b->sims, b->cell_work, b->cell_blocks, b->roles);
    FS2D_BHIP(b, cell_launch_check());
    int rc = solve(b, SWEEP_PRESSURE);
    if (rc) return rc;
    hipLaunchKernelGGL(kb_gradient, grid, blk, 0, b->stream, b->sims, b->cell_work, b->cell_blocks, b->roles,
                       b->slots);
    hipLaunchKernelGGL(kb_fold_speed, dim3(b->count), dim3(64), 0, b->stream, b->max_bits, b->slots);
    FS2D_BHIP(b, cell_launch_check());
    return FS2D_OK;
}

// Time one launch (j = 0, writing only scratch) of each stage solve for a
// candidate configuration; returns the estimated time of both kinds of
// solves of one step (3 diffuses counted once, their launches being shared,
// and 2 pressure solves).
int time_config(fs2d_batch* b, float* est_ms) {
    const JacobiConfig& c = b->cfg;
    float t[2] = {0.f, 0.f};
    for (int k = 0; k < 2; ++k) {
        const int kind = k == 0 ? SWEEP_DIFFUSE : SWEEP_PRESSURE;
        const int2* w = kind == SWEEP_DIFFUSE ? c.diffuse_work : c.pressure_work;
        const int n0 = kind == SWEEP_DIFFUSE ? c.diffuse_count[0] : c.pressure_count[0];
        float best = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
            FS2D_BHIP(b, hipEventRecord(b->ev0, b->stream));
            FS2D_BHIP(b, launch_kb_jacobi(b->stream, c.M, c.TY, kind, b->sims, w, n0, b->roles, 0, jac_limit(b)));
            FS2D_BHIP(b, hipEventRecord(b->ev1, b->stream));
            FS2D_BHIP(b, hipEventSynchronize(b->ev1));
            float ms = 0.f;
            FS2D_BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
            if (rep == 1 || (rep > 1 && ms < best)) best = ms;
        }
        // later launches cover fewer entries: scale by the entry counts
        const auto& counts = kind == SWEEP_DIFFUSE ? c.diffuse_count : c.pressure_count;
        double total = 0;
        for (int cnt : counts) total += static_cast<double>(cnt) / n0;
        t[k] = static_cast<float>(best * total);
    }
    *est_ms = t[0] + 2.f * t[1];
    return FS2D_OK;
}

// Choose depth and tile rows: the fixed jacobi_block (tile rows tuned), or
// every depth that divides all iteration counts at least twice.
int tune(fs2d_batch* b) {
    std::vector<int> depths;
    if (b->prm.jacobi_block)
        depths.push_back(b->prm.jacobi_block);
    else
        for (int m : kDepths)
            if (depth_fits(b, m)) depths.push_back(m);
    float best = 0.f;
    JacobiConfig chosen;
    for (int m : depths)
        for (int ty : kTileRows) {
            free_config(b->cfg);
            int rc = build_config(b, m, ty, &b->cfg);
            if (rc) return rc;
            if ((rc = upload_sims(b, m))) return rc;
            float est = 0.f;
            if ((rc = time_config(b, &est))) return rc;
            if (chosen.M == 0 || est < best) {
                best = est;
                chosen.M = m;
                chosen.TY = ty;
            }
        }
    free_config(b->cfg);
    int rc = build_config(b, chosen.M, chosen.TY, &b->cfg);
    if (rc) return rc;
    if ((rc = upload_sims(b, chosen.M))) return rc;
    if (b->occupancy_fixed) {
        b->tuned = true;
        return FS2D_OK;
    }
    // waves per CU of the persistent launches: Jacobi blocks (4 waves) per CU
    // timed on the same launches as above, single-pass blocks (4 waves) per
    // CU on the divergence (it only writes the divergence arena, which every
    // projection recomputes first); 0 is one block per work entry
    float best_j = 0.f;
    int pick_j = 0;
    for (int f : {0, 1, 2, 4, 8}) {
        b->jac_per_cu = f;
        float est = 0.f;
        if ((rc = time_config(b, &est))) return rc;
        if (f == 0 || est < best_j) {
            best_j = est;
            pick_j = f;
        }
    }
    b->jac_per_cu = pick_j;
    float best_c = 0.f;
    int pick_c = 0;
    for (int f : {0, 2, 4, 8}) {
        b->cell_per_cu = f;
        float t = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
            FS2D_BHIP(b, hipEventRecord(b->ev0, b->stream));
            hipLaunchKernelGGL(kb_divergence, cell_grid(b), dim3(CBX, CBY), 0, b->stream, b->sims, b->cell_work,
                               b->cell_blocks, b->roles);
            FS2D_BHIP(b, hipGetLastError());
            FS2D_BHIP(b, hipEventRecord(b->ev1, b->stream));
            FS2D_BHIP(b, hipEventSynchronize(b->ev1));
            float ms = 0.f;
            FS2D_BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
            if (rep == 1 || (rep > 1 && ms < t)) t = ms;
        }
        if (f == 0 || t < best_c) {
            best_c = t;
            pick_c = f;
        }
    }
    b->cell_per_cu = pick_c;
    b->tuned = true;
    return FS2D_OK;
}

// One step of every simulation (fluid_solver_cpu::solve, Jacobi order).
int step(fs2d_batch* b) {
    const dim3 blk(CBX, CBY);
    int rc;
    if (!b->speed_valid) {
        hipLaunchKernelGGL(kb_max_speed, cell_grid(b), blk, 0, b->stream, b->sims, b->cell_work, b->cell_blocks,
                           b->roles, b->slots);
        hipLaunchKernelGGL(kb_fold_speed, dim3(b->count), dim3(64), 0, b->stream, b->max_bits, b->slots);
        FS2D_BHIP(b, cell_launch_check());
        b->speed_valid = true;
    }
    // sources of all three fields; u and v go to fresh iterates so that the
    // density scatter still reads the pre-step velocity (:16-24)
    hipLaunchKernelGGL(kb_add_sources, cell_grid(b), blk, 0, b->stream, b->sims, b->cell_work, b->cell_blocks,
                       b->roles);
    FS2D_BHIP(b, cell_launch_check());
    if (!b->tuned && (rc = tune(b))) return rc;
    if ((rc = solve(b, SWEEP_DIFFUSE))) return rc;
    FS2D_BHIP(b, launch_batch_scatter(b->stream, b->scatter, b->roles.a[A_T1], b->roles.a[A_RHO], b->roles.a[A_U],
                                      b->roles.a[A_V], b->max_bits));
    std::swap(b->roles.a[A_RHO], b->roles.a[A_T1]);
    std::swap(b->roles.a[A_U], b->roles.a[A_UN]);
    std::swap(b->roles.a[A_V], b->roles.a[A_VN]);
    if ((rc = project(b))) return rc;
    hipLaunchKernelGGL(kb_gather_velocity, cell_grid(b), blk, 0, b->stream, b->sims, b->cell_work, b->cell_blocks,
                       b->roles);
    FS2D_BHIP(b, cell_launch_check());
    std::swap(b->roles.a[A_U], b->roles.a[A_T1]);
    std::swap(b->roles.a[A_V], b->roles.a[A_T2]);
    return project(b);
}

constexpr size_t kMaxBatchGraphs = 16;

void drop_graphs(fs2d_batch* b) {
    for (auto* g : b->graphs) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
    }
    b->graphs.clear();
}

// As fs2d_handle's step graphs: the first step runs eagerly (it tunes), then
// each pre-step state (arena permutation, speed flag) is captured once and
// replayed.
int advance(fs2d_batch* b) {
    if (b->prm.launch_mode == FS2D_LAUNCH_EAGER || b->steps_done == 0) {
        const int rc = step(b);
        if (rc == FS2D_OK) ++b->steps_done;
        return rc;
    }
    for (auto* g : b->graphs) {
        if (std::memcmp(g->pre, b->roles.a, sizeof g->pre) != 0 || g->pre_speed != b->speed_valid) continue;
        FS2D_BHIP(b, hipGraphLaunch(g->exec, b->stream));
        std::memcpy(b->roles.a, g->post, sizeof g->post);
        b->speed_valid = true;
        ++b->steps_done;
        return FS2D_OK;
    }
    if (b->graphs.size() >= kMaxBatchGraphs) drop_graphs(b);
    auto* g = new BatchGraph();
    std::memcpy(g->pre, b->roles.a, sizeof g->pre);
    g->pre_speed = b->speed_valid;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete g;
        return bfail(b, FS2D_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    const int rc = step(b);
    e = hipStreamEndCapture(b->stream, &g->graph);
    if (rc == FS2D_OK && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (rc != FS2D_OK || e != hipSuccess) {
        std::memcpy(b->roles.a, g->pre, sizeof g->pre);  // nothing the capture recorded has run
        b->speed_valid = g->pre_speed;
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc != FS2D_OK ? rc : bfail(b, FS2D_ERR_HIP, std::string("batch graph capture: ") + hipGetErrorString(e));
    }
    std::memcpy(g->post, b->roles.a, sizeof g->post);
    b->graphs.push_back(g);
    FS2D_BHIP(b, hipGraphLaunch(g->exec, b->stream));
    ++b->steps_done;
    return FS2D_OK;
}

int copy_sim(fs2d_batch* b, int sim, float* dst_arena, const float* host) {
    const BatchSim& S = b->host_sims[sim];
    const size_t row = sizeof(float) * S.R;
    FS2D_BHIP(b, hipMemcpy2DAsync(dst_arena + S.off, sizeof(float) * S.P, host, row, row, S.R,
                                  hipMemcpyHostToDevice, b->stream));
    return FS2D_OK;
}

int copy_sim_out(fs2d_batch* b, int sim, float* host, const float* src_arena) {
    const BatchSim& S = b->host_sims[sim];
    const size_t row = sizeof(float) * S.R;
    FS2D_BHIP(b, hipMemcpy2DAsync(host, row, src_arena + S.off, sizeof(float) * S.P, row, S.R,
                                  hipMemcpyDeviceToHost, b->stream));
    return FS2D_OK;
}

// Exceptions thrown on the host (allocation in plan building, maps and
// vectors) never cross the C ABI: they become an error code and message.
template <class F>
int guarded(fs2d_batch* h, const char* what, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return bfail(h, FS2D_ERR_HIP, std::string(what) + ": out of host memory");
    } catch (const std::exception& e) {
        return bfail(h, FS2D_ERR_STATE, std::string(what) + ": " + e.what());
    } catch (...) {
        return bfail(h, FS2D_ERR_STATE, std::string(what) + ": unexpected exception");
    }
}

}  // namespace

extern "C" {

int fs2d_hip_batch_destroy(fs2d_batch* b) {
    if (!b) return FS2D_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_graphs(b);
    free_config(b->cfg);
    for (float* p : b->roles.a)
        if (p) (void)hipFree(p);
    if (b->sims) (void)hipFree(b->sims);
    if (b->cell_work) (void)hipFree(b->cell_work);
    if (b->max_bits) (void)hipFree(b->max_bits);
    if (b->slots) (void)hipFree(b->slots);
    if (b->scatter_sims) (void)hipFree(b->scatter_sims);
    if (b->scatter_windows) (void)hipFree(b->scatter_windows);
    if (b->scatter_scratch) (void)hipFree(b->scatter_scratch);
    if (b->scatter_flag) (void)hipFree(b->scatter_flag);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return FS2D_OK;
}

static int fs2d_hip_batch_create_impl(int count, const fs2d_batch_sim* sims, const fs2d_params* params, fs2d_batch** out) {
    if (!out) return bfail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_create: out is NULL");
    *out = nullptr;
    if (count < 1 || count > (1 << 20) || !sims)
        return bfail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_create: count must be in [1, 2^20] with sims given");
    fs2d_params prm;
    fs2d_default_params(&prm);
    prm.sweep_order = FS2D_ORDER_JACOBI;
    if (params) prm = *params;
    if (prm.sweep_order != FS2D_ORDER_JACOBI)
        return bfail(nullptr, FS2D_ERR_ARGUMENT,
                     "fs2d_hip_batch_create: batches run the Jacobi order only (sweep_order must be FS2D_ORDER_JACOBI)");
    if (prm.launch_mode != FS2D_LAUNCH_GRAPH && prm.launch_mode != FS2D_LAUNCH_EAGER)
        return bfail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_create: launch_mode must be 0 (graph) or 1 (eager)");
    for (int s = 0; s < count; ++s) {
        const fs2d_batch_sim& q = sims[s];
        if (q.rows < 3 || q.rows > 8194)
            return bfail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_create: rows must be in [3, 8194]");
        if (q.iterations < 2 || q.iterations > 1000000)
            return bfail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_create: iterations must be in [2, 1e6]");
    }
    if (prm.jacobi_block != 0) {
        const int m = prm.jacobi_block;
        bool ok = m == 1 || m == 2 || m == 4 || m == 8;
        for (int s = 0; s < count && ok; ++s) ok = sims[s].iterations % m == 0 && sims[s].iterations / m >= 2;
        if (!ok)
            return bfail(nullptr, FS2D_ERR_ARGUMENT,
                         "fs2d_hip_batch_create: jacobi_block must be 1, 2, 4 or 8 and divide every simulation's "
                         "iterations at least twice");
    }
    fs2d_batch* b = new fs2d_batch();
    b->count = count;
    b->prm = prm;
    b->spec.assign(sims, sims + count);
    b->host_sims.resize(count);
    size_t off = 0;
    std::vector<int2> cw, sw;  // single-pass stage blocks; the scatter's window-path waves
    std::vector<ScatterSim> ss(count);
    long long scatter_cells = 0, scatter_targets = 0;
    for (int s = 0; s < count; ++s) {
        BatchSim& S = b->host_sims[s];
        const fs2d_batch_sim& q = sims[s];
        S.R = q.rows;
        S.P = (q.rows + 63) / 64 * 64;
        S.off = off;
        off += static_cast<size_t>(S.R) * S.P;  // a multiple of 64 floats
        S.dt = q.dt;
        // scalars exactly as the one-grid handle computes them
        S.dt0 = std::sqrt(cells_of(S.R)) * q.dt;
        S.h = 1.0f / std::sqrt(cells_of(S.R));
        S.neg_half_h = -0.5f * S.h;
        S.a_rho = q.dt * cells_of(S.R) * q.diffusion_rate;
        S.denom_rho = 1.f + 4.f * S.a_rho;
        S.a_vel = q.dt * cells_of(S.R) * q.viscosity;
        S.denom_vel = 1.f + 4.f * S.a_vel;
        const SweepArgs dr = sweep_args(nullptr, nullptr, S.R, S.P, SWEEP_DIFFUSE, S.a_rho, S.denom_rho, BND_CONT, 2, 0);
        const SweepArgs dv = sweep_args(nullptr, nullptr, S.R, S.P, SWEEP_DIFFUSE, S.a_vel, S.denom_vel, BND_CONT, 2, 0);
        S.inv_rho = dr.inv_denom;
        S.lo_rho = dr.div_lo;
        S.inv_lo_rho = dr.inv_denom_lo;
        S.inv_vel = dv.inv_denom;
        S.lo_vel = dv.div_lo;
        S.inv_lo_vel = dv.inv_denom_lo;
        const int n = S.R - 2;
        const int gx = (n + CBX - 1) / CBX, gy = (n + CBY - 1) / CBY;
        S.cell_blocks = gx * gy;
        for (int by = 0; by < gy; ++by)
            for (int bx = 0; bx < gx; ++bx) cw.push_back(make_int2(s, (by << 16) | bx));
        ss[s] = ScatterSim{S.off, S.R, S.P, S.dt0, static_cast<unsigned>(scatter_cells),
                           static_cast<unsigned>(scatter_targets)};
        scatter_cells += static_cast<long long>(n) * n;
        scatter_targets += static_cast<long long>(S.R - 1) * (S.R - 1);
        for (int band = 0; band < (n + kBatchBand - 1) / kBatchBand; ++band)
            for (int strip = 0; strip < (n + kBatchStrip - 1) / kBatchStrip; ++strip)
                sw.push_back(make_int2(s, (band << 16) | strip));
        b->interior += static_cast<long long>(n) * n;
        b->order.push_back(s);
    }
    std::stable_sort(b->order.begin(), b->order.end(),
                     [&](int x, int y) { return sims[x].iterations > sims[y].iterations; });
    b->arena_floats = off;
    b->cell_blocks = static_cast<int>(cw.size());
    auto bail = [&](hipError_t e, const char* what) {
        std::string msg = std::string("fs2d_hip_batch_create: ") + what + ": " + hipGetErrorString(e);
        fs2d_hip_batch_destroy(b);
        return bfail(nullptr, FS2D_ERR_HIP, msg);
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    {
        int dev = 0, cus = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0)
            b->cus = cus;
    }
    if ((e = hipEventCreate(&b->ev0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&b->ev1)) != hipSuccess) return bail(e, "hipEventCreate");
    const size_t bytes = sizeof(float) * off;
    const size_t fill_blocks = std::max<size_t>(1, std::min<size_t>((off + 255) / 256, 4096));
    for (int a = 0; a < A_COUNT; ++a) {
        if ((e = hipMalloc(&b->roles.a[a], bytes)) != hipSuccess) return bail(e, "hipMalloc(arena)");
        hipLaunchKernelGGL(kb_fill, dim3((unsigned)fill_blocks), dim3(256), 0, b->stream,
                           reinterpret_cast<uint32_t*>(b->roles.a[a]), off);
    }
    if ((e = hipMalloc(&b->sims, sizeof(BatchSim) * count)) != hipSuccess) return bail(e, "hipMalloc(sims)");
    if ((e = hipMalloc(&b->cell_work, sizeof(int2) * cw.size())) != hipSuccess) return bail(e, "hipMalloc(work)");
    if ((e = hipMalloc(&b->max_bits, sizeof(unsigned) * count)) != hipSuccess) return bail(e, "hipMalloc(speed)");
    const size_t slot_words = static_cast<size_t>(count) * kSimSlots * kSimSlotStride;
    if ((e = hipMalloc(&b->slots, sizeof(unsigned) * slot_words)) != hipSuccess) return bail(e, "hipMalloc(slots)");
    hipLaunchKernelGGL(kb_fill, dim3((unsigned)std::min<size_t>((slot_words + 255) / 256, 4096)), dim3(256), 0,
                       b->stream, b->slots, slot_words);
    // the density scatter: descriptors, window work list, sorted-path buffers
    if (scatter_cells >= (1ll << 31) || scatter_targets >= (1ll << 32) - 1)
        return bail(hipErrorInvalidValue, "too many cells for the batched density scatter");
    if ((e = hipMalloc(&b->scatter_sims, sizeof(ScatterSim) * count)) != hipSuccess)
        return bail(e, "hipMalloc(scatter descriptors)");
    if ((e = hipMalloc(&b->scatter_windows, sizeof(int2) * sw.size())) != hipSuccess)
        return bail(e, "hipMalloc(scatter work)");
    if ((e = hipMalloc(&b->scatter_scratch, batch_scatter_bytes(scatter_cells, scatter_targets))) != hipSuccess)
        return bail(e, "hipMalloc(scatter buffers)");
    if ((e = hipMalloc(&b->scatter_flag, 256)) != hipSuccess) return bail(e, "hipMalloc(scatter gate)");
    if ((e = hipMemcpyAsync(b->scatter_sims, ss.data(), sizeof(ScatterSim) * count, hipMemcpyHostToDevice,
                            b->stream)) != hipSuccess)
        return bail(e, "hipMemcpy(scatter descriptors)");
    if ((e = hipMemcpyAsync(b->scatter_windows, sw.data(), sizeof(int2) * sw.size(), hipMemcpyHostToDevice,
                            b->stream)) != hipSuccess)
        return bail(e, "hipMemcpy(scatter work)");
    b->scatter = BatchScatter{b->scatter_sims, count, b->scatter_windows, static_cast<int>(sw.size()), scatter_cells,
                              scatter_targets, b->scatter_scratch, b->scatter_flag};
    if ((e = hipMemcpyAsync(b->cell_work, cw.data(), sizeof(int2) * cw.size(), hipMemcpyHostToDevice, b->stream)) !=
        hipSuccess)
        return bail(e, "hipMemcpy(work)");
    if ((e = hipMemcpyAsync(b->sims, b->host_sims.data(), sizeof(BatchSim) * count, hipMemcpyHostToDevice,
                            b->stream)) != hipSuccess)
        return bail(e, "hipMemcpy(sims)");
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
    *out = b;
    return FS2D_OK;
}

int fs2d_hip_batch_create(int count, const fs2d_batch_sim* sims, const fs2d_params* params, fs2d_batch** out) {
    return guarded(nullptr, "fs2d_hip_batch_create", [&] { return fs2d_hip_batch_create_impl(count, sims, params, out); });
}

static int fs2d_hip_batch_upload_impl(fs2d_batch* b, int sim, const float* rho, const float* u, const float* v) {
    if (!b || sim < 0 || sim >= (b ? b->count : 0) || !rho || !u || !v)
        return bfail(b, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_upload: bad argument");
    int rc;
    b->speed_valid = false;
    if ((rc = copy_sim(b, sim, b->roles.a[A_RHO], rho))) return rc;
    if ((rc = copy_sim(b, sim, b->roles.a[A_U], u))) return rc;
    if ((rc = copy_sim(b, sim, b->roles.a[A_V], v))) return rc;
    FS2D_BHIP(b, hipStreamSynchronize(b->stream));
    return FS2D_OK;
}

int fs2d_hip_batch_upload(fs2d_batch* b, int sim, const float* rho, const float* u, const float* v) {
    return guarded(b, "fs2d_hip_batch_upload", [&] { return fs2d_hip_batch_upload_impl(b, sim, rho, u, v); });
}

static int fs2d_hip_batch_set_sources_impl(fs2d_batch* b, int sim, const float* rs, const float* us, const float* vs) {
    if (!b || sim < 0 || sim >= b->count) return bfail(b, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_set_sources: bad argument");
    const float* src[3] = {rs, us, vs};
    const int dst[3] = {A_RS, A_US, A_VS};
    const BatchSim& S = b->host_sims[sim];
    const size_t n = static_cast<size_t>(S.R) * S.P;
    for (int i = 0; i < 3; ++i) {
        if (src[i]) {
            int rc = copy_sim(b, sim, b->roles.a[dst[i]], src[i]);
            if (rc) return rc;
        } else {
            hipLaunchKernelGGL(kb_fill, dim3((unsigned)std::max<size_t>(1, std::min<size_t>((n + 255) / 256, 4096))),
                               dim3(256), 0, b->stream, reinterpret_cast<uint32_t*>(b->roles.a[dst[i]] + S.off), n);
            FS2D_BHIP(b, hipGetLastError());
        }
    }
    FS2D_BHIP(b, hipStreamSynchronize(b->stream));
    return FS2D_OK;
}

int fs2d_hip_batch_set_sources(fs2d_batch* b, int sim, const float* rs, const float* us, const float* vs) {
    return guarded(b, "fs2d_hip_batch_set_sources", [&] { return fs2d_hip_batch_set_sources_impl(b, sim, rs, us, vs); });
}

static int fs2d_hip_batch_step_n_impl(fs2d_batch* b, int n) {
    if (!b || n < 0) return bfail(b, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_step_n: bad argument");
    for (int i = 0; i < n; ++i) {
        const int rc = advance(b);
        if (rc) return rc;
    }
    return FS2D_OK;
}

int fs2d_hip_batch_step_n(fs2d_batch* b, int n) {
    return guarded(b, "fs2d_hip_batch_step_n", [&] { return fs2d_hip_batch_step_n_impl(b, n); });
}

static int fs2d_hip_batch_download_impl(fs2d_batch* b, int sim, float* rho, float* u, float* v) {
    if (!b || sim < 0 || sim >= b->count) return bfail(b, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_download: bad argument");
    int rc;
    if (rho && (rc = copy_sim_out(b, sim, rho, b->roles.a[A_RHO]))) return rc;
    if (u && (rc = copy_sim_out(b, sim, u, b->roles.a[A_U]))) return rc;
    if (v && (rc = copy_sim_out(b, sim, v, b->roles.a[A_V]))) return rc;
    FS2D_BHIP(b, hipStreamSynchronize(b->stream));
    return FS2D_OK;
}

int fs2d_hip_batch_download(fs2d_batch* b, int sim, float* rho, float* u, float* v) {
    return guarded(b, "fs2d_hip_batch_download", [&] { return fs2d_hip_batch_download_impl(b, sim, rho, u, v); });
}

static int fs2d_hip_batch_sync_impl(fs2d_batch* b) {
    if (!b) return bfail(b, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_sync: NULL batch");
    FS2D_BHIP(b, hipStreamSynchronize(b->stream));
    return FS2D_OK;
}

int fs2d_hip_batch_sync(fs2d_batch* b) {
    return guarded(b, "fs2d_hip_batch_sync", [&] { return fs2d_hip_batch_sync_impl(b); });
}

static int fs2d_hip_batch_info_impl(fs2d_batch* b, int* depth, int* tile_rows, int* launches, int* graphs,
                        long long* interior_cells) {
    if (!b) return bfail(b, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_info: NULL batch");
    if (depth) *depth = b->cfg.M;
    if (tile_rows) *tile_rows = b->cfg.TY;
    if (launches) {
        // max speed (first step only) aside: add, diffuse launches, the scatter's
        // kernels, 2 x (divergence, pressure launches, gradient, fold), gather
        const int d = static_cast<int>(b->cfg.diffuse_count.size()), p = static_cast<int>(b->cfg.pressure_count.size());
        *launches = b->tuned ? 1 + d + batch_scatter_launches(b->scatter.targets) + 2 * (3 + p) + 1 : 0;
    }
    if (graphs) *graphs = static_cast<int>(b->graphs.size());
    if (interior_cells) *interior_cells = b->interior;
    return FS2D_OK;
}

int fs2d_hip_batch_set_occupancy(fs2d_batch* b, int jacobi_blocks_per_cu, int stage_blocks_per_cu) {
    if (!b || jacobi_blocks_per_cu < 0 || jacobi_blocks_per_cu > 64 || stage_blocks_per_cu < 0 ||
        stage_blocks_per_cu > 64)
        return bfail(b, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_set_occupancy: blocks per CU must be in [0, 64]");
    if (b->steps_done > 0)
        return bfail(b, FS2D_ERR_STATE, "fs2d_hip_batch_set_occupancy: only before the first step");
    b->jac_per_cu = jacobi_blocks_per_cu;
    b->cell_per_cu = stage_blocks_per_cu;
    b->occupancy_fixed = true;
    return FS2D_OK;
}

int fs2d_hip_batch_occupancy(fs2d_batch* b, int* jacobi_blocks_per_cu, int* stage_blocks_per_cu,
                             int* compute_units) {
    if (!b) return bfail(b, FS2D_ERR_ARGUMENT, "fs2d_hip_batch_occupancy: NULL batch");
    if (jacobi_blocks_per_cu) *jacobi_blocks_per_cu = b->jac_per_cu;
    if (stage_blocks_per_cu) *stage_blocks_per_cu = b->cell_per_cu;
    if (compute_units) *compute_units = b->cus;
    return FS2D_OK;
}

int fs2d_hip_batch_info(fs2d_batch* b, int* depth, int* tile_rows, int* launches, int* graphs,
                        long long* interior_cells) {
    return guarded(b, "fs2d_hip_batch_info", [&] { return fs2d_hip_batch_info_impl(b, depth, tile_rows, launches, graphs, interior_cells); });
}

const char* fs2d_hip_batch_last_error(const fs2d_batch* b) {
    return b ? b->err.c_str() : g_batch_create_error.c_str();
}

}  // extern "C"
