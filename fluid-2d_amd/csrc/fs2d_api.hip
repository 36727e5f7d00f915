// fs2d_api.hip -- the C ABI of libfs2d_hip.so (declared in include/fs2d_hip.h).
//
// A handle owns its HIP stream and every device field of one simulation.
// Fields are allocated once at create time with a 256-byte row pitch and stay
// resident in HBM; the step composes the stage launchers in the order of
// fluid_solver_cpu::solve (src/fluid_solver_cpu.cpp:6-31) and swaps buffer
// pointers instead of copying where the reference copies a grid
// (initial_grid in advect, temp_* before the velocity advection).
#include "../../include/fs2d_cnn.h"
#include "../../include/fs2d_hip.h"
#include "../../include/fs2d_render.h"
#include "fs2d_internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <exception>
#include <new>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

using namespace fs2d;

namespace {

thread_local std::string g_create_error;

enum { F_RHO, F_U, F_V, F_RS, F_US, F_VS, F_X0, F_T1, F_T2, F_D, F_P, F_SCRATCH, F_X0U, F_X0V, F_UN, F_VN, F_COUNT };

struct GsPolicy {
    int tw, M, lag;  // tile columns, sweeps per unit (0: one-sweep tile kernel), lag
};

}  // namespace

struct fs2d_handle {
    int R = 0, P = 0;
    size_t field_bytes = 0;
    fs2d_params prm{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float* f[F_COUNT] = {};
    unsigned* max_bits = nullptr;
    unsigned* speed_slots = nullptr;  // kSpeedSlots * kSlotStride words, zero between uses
    std::map<std::tuple<int, int, int, int>, GsPlan> gs_plans;  // by (sweep count, tile columns, block, lag)
    // Gauss-Seidel launch policy: fixed by the parameters, or (0) chosen per
    // (kind, sweep count) by gs_policy_for; gs_block 1 selects the one-sweep
    // tile kernel, anything else the unit kernels.
    int gs_tw = 0, gs_block = 0, gs_lag = 0;
    std::map<std::pair<int, int>, GsPolicy> gs_policy;  // (kind, K) -> policy in use
    std::map<std::tuple<int, int, int, int, int>, int> pipe_offsets;  // (Kd, Kp, tw, M, lag) -> pipeline offset
    std::map<std::pair<int, int>, GsPolicy> joint_policy;             // (Kd, Kp) -> policy of the joint launches
    int jacobi_kernel = FS2D_JACOBI_WAVE;
    std::map<std::pair<int, int>, int> jacobi_depth;  // (kind, K) -> sweeps per launch, when tuned
    bool pressure_zero = true;  // last projection had K_p == 0 (p is identically 0)
    bool speed_valid = false;   // max_bits holds max |u|, |v| of the current velocity
    long long steps_done = 0;
    double* res_sums = nullptr;  // 2 doubles: residual sums of the seeded projection
    void* sort_buf = nullptr;    // the density scatter's sorted path (scatter_sort_bytes), on the first advect
    void* render_buf = nullptr;  // device staging of the visual output, grown on demand
    size_t render_bytes = 0;
    struct MgLevel {
        int n, R, P;
        float *p, *t, *b;  // iterate, its ping-pong partner, right-hand side
    };
    std::vector<MgLevel> mg;  // coarse grids of the multigrid seed (side n/2, n/4, ...), on first use
    struct StepGraph;
    std::vector<StepGraph*> graphs;
    std::string err;

    float*& rho() { return f[F_RHO]; }
    float*& u() { return f[F_U]; }
    float*& v() { return f[F_V]; }
};

void fs2d::set_handle_free_error(const std::string& msg) { g_create_error = msg; }

namespace {

int fail(fs2d_handle* h, int code, const std::string& msg) {
    if (h) h->err = msg;
    else g_create_error = msg;
    return code;
}

#define FS2D_HIP(h, expr)                                                                     \
    do {                                                                                      \
        hipError_t e_ = (expr);                                                               \
        if (e_ != hipSuccess)                                                                 \
            return fail((h), FS2D_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

// Scalars exactly as the reference computes them.
struct Scalars {
    float dt0, h, neg_half_h;
};

float cells(int R) { return static_cast<float>(static_cast<size_t>(R) * static_cast<size_t>(R)); }

Scalars scalars(int R, float dt) {
    Scalars s;
    s.dt0 = std::sqrt(cells(R)) * dt;       // src/fluid_solver_cpu.cpp:126
    s.h = 1.0f / std::sqrt(cells(R));       // :183
    s.neg_half_h = -0.5f * s.h;             // :190, (-0.5f * h) * (...)
    return s;
}

bool gs_units(const fs2d_handle* h) { return h->gs_block != 1; }

hipError_t gs_plan_of(fs2d_handle* h, int K, const GsPolicy& c, const GsPlan** out) {
    const auto key = std::make_tuple(K, c.tw, c.M, c.lag);
    auto it = h->gs_plans.find(key);
    if (it == h->gs_plans.end()) {
        GsPlan plan{};
        hipError_t e = c.M > 0 ? gs_unit_plan_create(&plan, h->R, K, c.tw, c.M, c.lag) : gs_plan_create(&plan, h->R, K, c.tw);
        if (e != hipSuccess) {
            gs_plan_destroy(&plan);
            return e;
        }
        it = h->gs_plans.emplace(key, plan).first;
    }
    *out = &it->second;
    return hipSuccess;
}

// Gauss-Seidel launch policy of a solve kind and sweep count (the reference
// fixes its launch shape, kernel_launcher.hpp:10-30; every policy here gives
// identical results).  Parameters left at 0 are chosen on the first solve of
// each (kind, K): every candidate -- tile columns {64, 128} x sweeps per unit
// {4, 8} x lag {4, 8}, within the fixed parameters -- relaxes zero-started
// scratch iterates against the resident fields (three fields for a diffuse,
// as the step relaxes them together), timed with HIP events (best of two
// after a warm-up), and the fastest is kept, also process-wide for the same
// (device, rows, K, kind, fixed parameters).  A stream under capture cannot
// be timed: the default (128 columns for diffuses, 64 for pressure; 8 sweeps;
// lag 4) serves it and is not kept.  Plans are kept until the handle is
// destroyed (a captured step may hold any of them).
std::mutex g_policy_mu;
std::map<std::tuple<int, int, int, int, int, int, int>, GsPolicy> g_policy_cache;

// The launch policies open to the tuner: tile columns {64, 128} x sweeps per
// unit {4, 8} x lag {4, 8}, each parameter fixed where the handle fixes it.
std::vector<GsPolicy> policy_candidates(const fs2d_handle* h) {
    std::vector<GsPolicy> cand;
    for (int tw : {64, 128})
        for (int M : {4, 8})
            for (int lag : {4, 8}) {
                GsPolicy c{h->gs_tw ? h->gs_tw : tw, h->gs_block ? (h->gs_block == 1 ? 0 : h->gs_block) : M,
                           h->gs_lag ? h->gs_lag : lag};
                if (c.M == 0) c.lag = 1;
                if (c.M == 2 && c.tw == 128) continue;
                bool dup = false;
                for (const auto& d : cand) dup = dup || (d.tw == c.tw && d.M == c.M && d.lag == c.lag);
                if (!dup) cand.push_back(c);
            }
    return cand;
}

hipError_t gs_policy_for(fs2d_handle* h, int K, int kind, GsPolicy* out) {
    const auto hkey = std::make_pair(kind, K);
    auto it = h->gs_policy.find(hkey);
    if (it != h->gs_policy.end()) {
        *out = it->second;
        return hipSuccess;
    }
    GsPolicy def{};
    def.M = h->gs_block == 1 ? 0 : (h->gs_block ? h->gs_block : 8);
    def.tw = h->gs_tw ? h->gs_tw : ((kind == SWEEP_DIFFUSE && def.M != 2) ? 128 : 64);
    def.lag = def.M == 0 ? 1 : (h->gs_lag ? h->gs_lag : 4);
    const std::vector<GsPolicy> cand = policy_candidates(h);
    if (cand.size() <= 1 || K <= 0) {
        *out = cand.size() == 1 ? cand[0] : def;
        h->gs_policy.emplace(hkey, *out);
        return hipSuccess;
    }
    const auto gkey = std::make_tuple(h->device, h->R, K, kind, h->gs_tw, h->gs_block, h->gs_lag);
    {
        std::lock_guard<std::mutex> lock(g_policy_mu);
        auto g = g_policy_cache.find(gkey);
        if (g != g_policy_cache.end()) {
            *out = g->second;
            h->gs_policy.emplace(hkey, *out);
            return hipSuccess;
        }
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    hipError_t e = hipStreamIsCapturing(h->stream, &cs);
    if (e != hipSuccess) return e;
    if (cs != hipStreamCaptureStatusNone) {
        *out = def;
        return hipSuccess;
    }
    const int nf = kind == SWEEP_DIFFUSE ? 3 : 1;
    float* tmp[3] = {nullptr, nullptr, nullptr};
    auto release = [&]() {
        for (float* q : tmp)
            if (q) (void)hipFree(q);
    };
    for (int f = 0; f < nf; ++f)
        if ((e = hipMalloc(&tmp[f], h->field_bytes)) != hipSuccess) return release(), e;
    SweepSet set{};
    set.count = nf;
    const int modes[3] = {BND_CONT, BND_OPP_H, BND_OPP_V};
    const int rhs[3] = {F_RHO, F_U, F_V};
    for (int f = 0; f < nf; ++f)
        set.a[f] = kind == SWEEP_DIFFUSE
                       ? sweep_args(tmp[f], h->f[rhs[f]], h->R, h->P, SWEEP_DIFFUSE, 1.0f, 5.0f, modes[f], K, 0)
                       : sweep_args(tmp[f], h->f[F_D], h->R, h->P, SWEEP_PRESSURE, 0.f, 0.f, BND_CONT, K, 0);
    float best_ms = 0.f;
    GsPolicy best = def;
    for (size_t c = 0; c < cand.size(); ++c) {
        const GsPlan* plan = nullptr;
        if ((e = gs_plan_of(h, K, cand[c], &plan)) != hipSuccess) return release(), e;
        SweepSet run = set;
        if (plan->M == 0 && nf > 1) run.count = 1;  // the tile kernel relaxes one field per launch sequence
        float t = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
            for (int f = 0; f < nf && e == hipSuccess; ++f) e = launch_fill(h->stream, tmp[f], 0u, h->field_bytes / 4);
            if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
            for (int f = 0; f < nf && e == hipSuccess; ++f) {
                if (plan->M == 0) {
                    e = launch_gs_solve(h->stream, *plan, set.a[f]);
                } else {
                    e = launch_gs_solve_set(h->stream, *plan, run);
                    break;
                }
            }
            if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
            if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev0, h->ev1);
            if (e != hipSuccess) return release(), e;
            if (rep == 1 || (rep > 1 && ms < t)) t = ms;
        }
        if (c == 0 || t < best_ms) {
            best_ms = t;
            best = cand[c];
        }
    }
    release();
    {
        std::lock_guard<std::mutex> lock(g_policy_mu);
        g_policy_cache.emplace(gkey, best);
    }
    h->gs_policy.emplace(hkey, best);
    *out = best;
    return hipSuccess;
}

hipError_t gs_plan_for(fs2d_handle* h, int K, int kind, const GsPlan** out) {
    GsPolicy c{};
    hipError_t e = gs_policy_for(h, K, kind, &c);
    if (e != hipSuccess) return e;
    return gs_plan_of(h, K, c, out);
}

constexpr int kJacobiDefaultDepth = 4;

// Sweeps fused per Jacobi launch: the fixed jacobi_block if one was given,
// otherwise the depth in {1, 2, 4, 8} with the least measured time per solve,
// chosen on the first solve of each (kind, K): one launch of each depth is
// timed (best of two, after a warm-up) reading the live iterate and writing
// only the scratch field.  Results are identical for every depth.  A stream
// under capture cannot be timed; the default depth is used then (not kept).
int jacobi_depth_for(fs2d_handle* h, const SweepArgs& A, int* out) {
    if (h->prm.jacobi_block) {
        *out = h->prm.jacobi_block;
        return FS2D_OK;
    }
    const auto key = std::make_pair(A.kind, A.K);
    auto it = h->jacobi_depth.find(key);
    if (it != h->jacobi_depth.end()) {
        *out = it->second;
        return FS2D_OK;
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    FS2D_HIP(h, hipStreamIsCapturing(h->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) {
        *out = kJacobiDefaultDepth;
        return FS2D_OK;
    }
    float best = 0.f;
    int best_m = kJacobiDefaultDepth;
    for (int m : {1, 2, 4, 8}) {
        if (m > 1 && m > A.K) break;
        float t = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
            FS2D_HIP(h, hipEventRecord(h->ev0, h->stream));
            FS2D_HIP(h, launch_jacobi_pass(h->stream, A, A.x, h->f[F_SCRATCH], A.zero_start, m, h->jacobi_kernel));
            FS2D_HIP(h, hipEventRecord(h->ev1, h->stream));
            FS2D_HIP(h, hipEventSynchronize(h->ev1));
            float ms = 0.f;
            FS2D_HIP(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
            if (rep == 1 || (rep > 1 && ms < t)) t = ms;
        }
        const float cost = t * static_cast<float>(jacobi_launches(A.K, m));
        if (m == 1 || cost < best) {
            best = cost;
            best_m = m;
        }
    }
    h->jacobi_depth.emplace(key, best_m);
    *out = best_m;
    return FS2D_OK;
}

// Sweeps per Jacobi launch without tuning (for reporting): the fixed or tuned
// depth, else the default.
int jacobi_depth_known(const fs2d_handle* h, int kind, int K) {
    if (h->prm.jacobi_block) return h->prm.jacobi_block;
    auto it = h->jacobi_depth.find(std::make_pair(kind, K));
    return it != h->jacobi_depth.end() ? it->second : kJacobiDefaultDepth;
}

// K relaxation sweeps of x against b; Jacobi may leave the result in the
// scratch field, in which case the two pointers are swapped.
int linear_solve(fs2d_handle* h, float*& x, const float* b, int kind, float a, float denom, int mode, int K,
                 int zero_start) {
    if (K <= 0) return FS2D_OK;
    SweepArgs A = sweep_args(x, b, h->R, h->P, kind, a, denom, mode, K, zero_start);
    if (h->prm.sweep_order == FS2D_ORDER_GS_LEX) {
        const GsPlan* plan = nullptr;
        FS2D_HIP(h, gs_plan_for(h, K, kind, &plan));
        if (plan->M > 0 && zero_start) {
            // the unit kernel starts from a zero-filled iterate, ring included
            // (zero_start == 2: the caller has filled it already)
            if (zero_start != 2) FS2D_HIP(h, launch_fill(h->stream, x, 0u, h->field_bytes / 4));
            A.zero_start = 0;
        }
        FS2D_HIP(h, launch_gs_solve(h->stream, *plan, A));
    } else {
        int depth = kJacobiDefaultDepth, rc;
        if ((rc = jacobi_depth_for(h, A, &depth))) return rc;
        FS2D_HIP(h, launch_jacobi_solve(h->stream, A, h->f[F_SCRATCH], depth, h->jacobi_kernel));
        if (jacobi_launches(K, depth) & 1) std::swap(x, h->f[F_SCRATCH]);
    }
    return FS2D_OK;
}

int diffuse(fs2d_handle* h, float*& x, int mode, float rate, float dt, int K) {
    // a = (dt * float(R*R)) * rate, src/fluid_solver_cpu.cpp:100
    const float a = dt * cells(h->R) * rate;
    const float denom = 1.f + 4.f * a;
    return linear_solve(h, x, h->f[F_X0], SWEEP_DIFFUSE, a, denom, mode, K, 0);
}

// Zero-start pressure iterate for the solver in use: the Gauss-Seidel unit
// kernel and Jacobi take a zero-filled p; the one-sweep tile kernel reads
// zeros itself.
bool pressure_needs_fill(fs2d_handle* h) {
    return !(h->prm.sweep_order == FS2D_ORDER_GS_LEX && !gs_units(h));
}

int project(fs2d_handle* h, int K) {
    const Scalars s = scalars(h->R, 1.0f);
    const bool fill = K > 0 && pressure_needs_fill(h);
    FS2D_HIP(h, launch_divergence(h->stream, h->f[F_D], h->u(), h->v(), h->R, h->P, s.neg_half_h,
                                  fill ? h->f[F_P] : nullptr));
    int rc = linear_solve(h, h->f[F_P], h->f[F_D], SWEEP_PRESSURE, 0.f, 0.f, BND_CONT, K, fill ? 2 : 1);
    if (rc) return rc;
    h->pressure_zero = (K <= 0);
    FS2D_HIP(h, launch_gradient(h->stream, h->u(), h->v(), h->f[F_P], h->R, h->P, s.h, h->pressure_zero ? 1 : 0,
                                h->max_bits, h->speed_slots));
    h->speed_valid = true;
    return FS2D_OK;
}

int advect_density(fs2d_handle* h, float dt) {
    const Scalars s = scalars(h->R, dt);
    if (!h->speed_valid) {
        FS2D_HIP(h, launch_fill(h->stream, h->max_bits, 0u, 1));
        FS2D_HIP(h, launch_max_speed(h->stream, h->max_bits, h->u(), h->v(), h->R, h->P));
        h->speed_valid = true;
    }
    if (!h->sort_buf) {
        // A handle's first advect runs eagerly (the first step is never
        // captured), so this allocation never happens inside a capture.
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        FS2D_HIP(h, hipStreamIsCapturing(h->stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(h, FS2D_ERR_STATE, "advect_density: the scatter's buffers must exist before a capture");
        FS2D_HIP(h, hipMalloc(&h->sort_buf, scatter_sort_bytes(h->R)));
    }
    FS2D_HIP(h, launch_scatter_density(h->stream, h->f[F_T1], h->rho(), h->u(), h->v(), h->R, h->P, s.dt0,
                                       h->max_bits, h->sort_buf));
    std::swap(h->rho(), h->f[F_T1]);
    return FS2D_OK;
}

int advect_velocity(fs2d_handle* h, float dt) {
    const Scalars s = scalars(h->R, dt);
    h->speed_valid = false;
    FS2D_HIP(h, launch_gather_velocity(h->stream, h->f[F_T1], h->f[F_T2], h->u(), h->v(), h->R, h->P, s.dt0));
    std::swap(h->u(), h->f[F_T1]);
    std::swap(h->v(), h->f[F_T2]);
    return FS2D_OK;
}

// Levels the first projection's pressure solve trails the diffuses by, when
// pipelined (gs_pipeline_offset), per (Kd, Kp, policy).
int pipeline_offset_for(fs2d_handle* h, int Kd, const GsPlan& pp, const GsPolicy& c) {
    const auto key = std::make_tuple(Kd, pp.K, c.tw, c.M, c.lag);
    auto it = h->pipe_offsets.find(key);
    if (it != h->pipe_offsets.end()) return it->second;
    const int off = gs_pipeline_offset(pp, h->R, Kd, c.tw, c.M);
    h->pipe_offsets.emplace(key, off);
    return off;
}

// The level launches that relax a step's three diffuses together with its
// first pressure solve (launch_gs_pipelined): one launch policy for both kinds.
struct JointPlans {
    const GsPlan* d = nullptr;  // null: the step does not pipeline (no sweeps of a kind, or no units)
    const GsPlan* p = nullptr;
    int offset = 0;
};

// The joint launches' policy, chosen like gs_policy_for's but on what they
// run: every candidate relaxes three zero-started scratch diffuse iterates and
// a scratch pressure iterate whose right-hand side is computed on load from
// two of them, pipelined at that policy's offset, timed with HIP events (best
// of two after a warm-up); the fastest is kept, also process-wide for the
// same (device, rows, Kd, Kp, fixed parameters).  Under capture the diffuse's
// policy serves and is not kept.
std::map<std::tuple<int, int, int, int, int, int, int>, GsPolicy> g_joint_cache;

hipError_t joint_plans(fs2d_handle* h, int Kd, int Kp, JointPlans* out) {
    *out = JointPlans{};
    if (Kd <= 0 || Kp <= 0 || h->prm.sweep_order != FS2D_ORDER_GS_LEX || !gs_units(h)) return hipSuccess;
    auto finish = [&](const GsPolicy& c) -> hipError_t {
        hipError_t e = gs_plan_of(h, Kd, c, &out->d);
        if (e == hipSuccess) e = gs_plan_of(h, Kp, c, &out->p);
        if (e != hipSuccess) return *out = JointPlans{}, e;
        out->offset = pipeline_offset_for(h, Kd, *out->p, c);
        return hipSuccess;
    };
    const auto hkey = std::make_pair(Kd, Kp);
    auto it = h->joint_policy.find(hkey);
    if (it != h->joint_policy.end()) return finish(it->second);
    GsPolicy def{};
    hipError_t e = gs_policy_for(h, Kd, SWEEP_DIFFUSE, &def);
    if (e != hipSuccess) return e;
    if (def.M <= 0) return hipSuccess;
    std::vector<GsPolicy> cand;
    for (const GsPolicy& c : policy_candidates(h))
        if (c.M > 0) cand.push_back(c);
    if (cand.size() <= 1) {
        const GsPolicy c = cand.empty() ? def : cand[0];
        h->joint_policy.emplace(hkey, c);
        return finish(c);
    }
    const auto gkey = std::make_tuple(h->device, h->R, Kd, Kp, h->gs_tw, h->gs_block, h->gs_lag);
    {
        std::lock_guard<std::mutex> lock(g_policy_mu);
        auto g = g_joint_cache.find(gkey);
        if (g != g_joint_cache.end()) {
            h->joint_policy.emplace(hkey, g->second);
            return finish(g->second);
        }
    }
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if ((e = hipStreamIsCapturing(h->stream, &cs)) != hipSuccess) return e;
    if (cs != hipStreamCaptureStatusNone) return finish(def);
    float* tmp[4] = {nullptr, nullptr, nullptr, nullptr};
    auto release = [&]() {
        for (float* q : tmp)
            if (q) (void)hipFree(q);
    };
    for (float*& q : tmp)
        if ((e = hipMalloc(&q, h->field_bytes)) != hipSuccess) return release(), e;
    SweepSet set{};
    set.count = 3;
    const int modes[3] = {BND_CONT, BND_OPP_H, BND_OPP_V};
    const int rhs[3] = {F_RHO, F_U, F_V};
    for (int f = 0; f < 3; ++f)
        set.a[f] = sweep_args(tmp[f], h->f[rhs[f]], h->R, h->P, SWEEP_DIFFUSE, 1.0f, 5.0f, modes[f], Kd, 0);
    SweepArgs A = sweep_args(tmp[3], h->f[F_D], h->R, h->P, SWEEP_PRESSURE, 0.f, 0.f, BND_CONT, Kp, 0);
    A.du = tmp[1];
    A.dv = tmp[2];
    A.neg_half_h = scalars(h->R, 1.0f).neg_half_h;
    float best_ms = 0.f;
    GsPolicy best = def;
    for (size_t c = 0; c < cand.size(); ++c) {
        if ((e = finish(cand[c])) != hipSuccess) return release(), e;
        float t = 0.f;
        for (int rep = 0; rep < 3; ++rep) {
            for (int f = 0; f < 4 && e == hipSuccess; ++f) e = launch_fill(h->stream, tmp[f], 0u, h->field_bytes / 4);
            if (e == hipSuccess) e = hipEventRecord(h->ev0, h->stream);
            if (e == hipSuccess) e = launch_gs_pipelined(h->stream, *out->d, set, *out->p, A, out->offset);
            if (e == hipSuccess) e = hipEventRecord(h->ev1, h->stream);
            if (e == hipSuccess) e = hipEventSynchronize(h->ev1);
            float ms = 0.f;
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, h->ev0, h->ev1);
            if (e != hipSuccess) return release(), e;
            if (rep == 1 || (rep > 1 && ms < t)) t = ms;
        }
        if (c == 0 || t < best_ms) {
            best_ms = t;
            best = cand[c];
        }
    }
    release();
    {
        std::lock_guard<std::mutex> lock(g_policy_mu);
        g_joint_cache.emplace(gkey, best);
    }
    h->joint_policy.emplace(hkey, best);
    return finish(best);
}

// Distinct level launches of a joint run: the diffuse's levels and the
// pressure's, shifted by the offset.
int joint_levels(const JointPlans& j) {
    std::vector<int> lv(j.d->level_of, j.d->level_of + j.d->levels);
    for (int l = 0; l < j.p->levels; ++l) lv.push_back(j.p->level_of[l] + j.offset);
    std::sort(lv.begin(), lv.end());
    return static_cast<int>(std::unique(lv.begin(), lv.end()) - lv.begin());
}

// The three diffuses of a step relaxed together (Gauss-Seidel unit plans):
// density in place, u and v on fresh iterates so that the density scatter can
// still read the pre-step velocity (src/fluid_solver_cpu.cpp:16-24).  Same
// results as the sequential order: the three solves share no data.
int step_shared_diffuse(fs2d_handle* h, float diff, float visc, float dt) {
    const int Kd = h->prm.diffuse_iterations, Kp = h->prm.pressure_iterations;
    const int R = h->R, P = h->P;
    int rc;
    FS2D_HIP(h, launch_add_sources3(h->stream, h->rho(), h->f[F_X0], h->f[F_RS], h->f[F_UN], h->f[F_X0U], h->u(),
                                    h->f[F_US], h->f[F_VN], h->f[F_X0V], h->v(), h->f[F_VS], R, P, dt));
    if (Kd > 0) {
        const float ad = dt * cells(R) * diff, av = dt * cells(R) * visc;
        const float dd = 1.f + 4.f * ad, dv = 1.f + 4.f * av;
        SweepSet set{};
        set.count = 3;
        set.a[0] = sweep_args(h->rho(), h->f[F_X0], R, P, SWEEP_DIFFUSE, ad, dd, BND_CONT, Kd, 0);
        set.a[1] = sweep_args(h->f[F_UN], h->f[F_X0U], R, P, SWEEP_DIFFUSE, av, dv, BND_OPP_H, Kd, 0);
        set.a[2] = sweep_args(h->f[F_VN], h->f[F_X0V], R, P, SWEEP_DIFFUSE, av, dv, BND_OPP_V, Kd, 0);
        JointPlans joint;
        FS2D_HIP(h, joint_plans(h, Kd, Kp, &joint));
        if (joint.d) {
            // The first projection pipelined behind the diffuses: its pressure
            // solve (right-hand side the divergence of the diffused velocity,
            // computed where each unit loads it) runs in the same level
            // launches, `offset` levels behind.
            FS2D_HIP(h, launch_fill(h->stream, h->f[F_P], 0u, h->field_bytes / 4));
            SweepArgs A = sweep_args(h->f[F_P], h->f[F_D], R, P, SWEEP_PRESSURE, 0.f, 0.f, BND_CONT, Kp, 0);
            A.du = h->f[F_UN];
            A.dv = h->f[F_VN];
            A.neg_half_h = scalars(R, 1.0f).neg_half_h;
            FS2D_HIP(h, launch_gs_pipelined(h->stream, *joint.d, set, *joint.p, A, joint.offset));
            if ((rc = advect_density(h, dt))) return rc;  // reads the pre-step u, v
            std::swap(h->u(), h->f[F_UN]);
            std::swap(h->v(), h->f[F_VN]);
            h->pressure_zero = false;
            FS2D_HIP(h, launch_gradient(h->stream, h->u(), h->v(), h->f[F_P], R, P, scalars(R, 1.0f).h, 0,
                                        h->max_bits, h->speed_slots));
            h->speed_valid = true;
            if ((rc = advect_velocity(h, dt))) return rc;
            // the second projection from a stored divergence: its units then
            // read one right-hand side field instead of two (faster than the
            // on-load divergence, which only the pipelined first solve needs)
            return project(h, Kp);
        }
        const GsPlan* plan = nullptr;
        FS2D_HIP(h, gs_plan_for(h, Kd, SWEEP_DIFFUSE, &plan));
        FS2D_HIP(h, launch_gs_solve_set(h->stream, *plan, set));
    }
    if ((rc = advect_density(h, dt))) return rc;  // reads the pre-step u, v
    std::swap(h->u(), h->f[F_UN]);
    std::swap(h->v(), h->f[F_VN]);
    h->speed_valid = false;
    if ((rc = project(h, Kp))) return rc;
    if ((rc = advect_velocity(h, dt))) return rc;
    return project(h, Kp);
}

// fluid_solver_cpu::solve, src/fluid_solver_cpu.cpp:6-31.
int step(fs2d_handle* h, float diff, float visc, float dt) {
    const int Kd = h->prm.diffuse_iterations, Kp = h->prm.pressure_iterations;
    const int R = h->R, P = h->P;
    int rc;
    if (h->prm.sweep_order == FS2D_ORDER_GS_LEX && gs_units(h)) return step_shared_diffuse(h, diff, visc, dt);
    // density: add source, diffuse, scatter-advect (:16-18)
    FS2D_HIP(h, launch_add_sources(h->stream, h->rho(), h->f[F_X0], h->f[F_RS], R, P, dt));
    if ((rc = diffuse(h, h->rho(), BND_CONT, diff, dt, Kd))) return rc;
    if ((rc = advect_density(h, dt))) return rc;
    // velocity: add sources, diffuse (:21-24; u and v are independent, so
    // each field's add is followed directly by its diffuse)
    h->speed_valid = false;
    FS2D_HIP(h, launch_add_sources(h->stream, h->u(), h->f[F_X0], h->f[F_US], R, P, dt));
    if ((rc = diffuse(h, h->u(), BND_OPP_H, visc, dt, Kd))) return rc;
    FS2D_HIP(h, launch_add_sources(h->stream, h->v(), h->f[F_X0], h->f[F_VS], R, P, dt));
    if ((rc = diffuse(h, h->v(), BND_OPP_V, visc, dt, Kd))) return rc;
    // project, self-advect, project (:25-30)
    if ((rc = project(h, Kp))) return rc;
    if ((rc = advect_velocity(h, dt))) return rc;
    return project(h, Kp);
}

}  // namespace

// Everything on the host that one step reads and changes: the field buffer
// assignment (the step swaps pointers where the reference copies a grid), the
// two state flags and the step's scalars.  A step captured from one such
// state replays correctly from that same state, and leaves `post`.
struct fs2d_handle::StepGraph {
    float* pre[F_COUNT];
    float* post[F_COUNT];
    bool pre_speed, pre_pzero, post_speed, post_pzero;
    float diff, visc, dt;
    hipGraph_t graph = nullptr;  // kept while its executable lives (tools inspect it at launch)
    hipGraphExec_t exec = nullptr;
};

namespace {

constexpr size_t kMaxStepGraphs = 32;  // the buffer permutation cycles within 12 steps

void drop_graphs(fs2d_handle* h) {
    for (auto* g : h->graphs) {
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
    }
    h->graphs.clear();
}

bool same_bits(float a, float b) { return std::memcmp(&a, &b, sizeof a) == 0; }

// One step through a replayed graph.  The first step of a handle runs eagerly:
// it builds the Gauss-Seidel plans and tunes the Jacobi depth, neither of which
// may happen inside a capture.  After that each distinct pre-step state is
// captured once (the capture runs the step's host logic, so the pointer swaps
// and flags come out as the eager step leaves them), instantiated, and
// replayed whenever the state recurs.
int step_graph(fs2d_handle* h, float diff, float visc, float dt) {
    if (h->prm.launch_mode == FS2D_LAUNCH_EAGER || h->steps_done == 0) return step(h, diff, visc, dt);
    for (auto* g : h->graphs) {
        if (std::memcmp(g->pre, h->f, sizeof g->pre) != 0 || g->pre_speed != h->speed_valid ||
            g->pre_pzero != h->pressure_zero || !same_bits(g->diff, diff) || !same_bits(g->visc, visc) ||
            !same_bits(g->dt, dt))
            continue;
        FS2D_HIP(h, hipGraphLaunch(g->exec, h->stream));
        std::memcpy(h->f, g->post, sizeof g->post);
        h->speed_valid = g->post_speed;
        h->pressure_zero = g->post_pzero;
        return FS2D_OK;
    }
    if (h->graphs.size() >= kMaxStepGraphs) drop_graphs(h);
    auto* g = new fs2d_handle::StepGraph();
    std::memcpy(g->pre, h->f, sizeof g->pre);
    g->pre_speed = h->speed_valid;
    g->pre_pzero = h->pressure_zero;
    g->diff = diff;
    g->visc = visc;
    g->dt = dt;
    auto restore = [&]() {
        std::memcpy(h->f, g->pre, sizeof g->pre);
        h->speed_valid = g->pre_speed;
        h->pressure_zero = g->pre_pzero;
    };
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        delete g;
        return fail(h, FS2D_ERR_HIP, std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    }
    const int rc = step(h, diff, visc, dt);
    e = hipStreamEndCapture(h->stream, &g->graph);
    if (rc == FS2D_OK && e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (rc != FS2D_OK || e != hipSuccess) {
        restore();  // nothing the capture recorded has run
        if (g->exec) (void)hipGraphExecDestroy(g->exec);
        if (g->graph) (void)hipGraphDestroy(g->graph);
        delete g;
        return rc != FS2D_OK ? rc : fail(h, FS2D_ERR_HIP, std::string("step graph capture: ") + hipGetErrorString(e));
    }
    std::memcpy(g->post, h->f, sizeof g->post);
    g->post_speed = h->speed_valid;
    g->post_pzero = h->pressure_zero;
    h->graphs.push_back(g);
    FS2D_HIP(h, hipGraphLaunch(g->exec, h->stream));
    return FS2D_OK;
}

int advance(fs2d_handle* h, float diff, float visc, float dt) {
    const int rc = step_graph(h, diff, visc, dt);
    if (rc == FS2D_OK) ++h->steps_done;
    return rc;
}

int copy_in(fs2d_handle* h, float* dst, const float* src, hipMemcpyKind kind) {
    const size_t row = sizeof(float) * h->R;
    FS2D_HIP(h, hipMemcpy2DAsync(dst, sizeof(float) * h->P, src, row, row, h->R, kind, h->stream));
    return FS2D_OK;
}

int copy_out(fs2d_handle* h, float* dst, const float* src, hipMemcpyKind kind) {
    const size_t row = sizeof(float) * h->R;
    FS2D_HIP(h, hipMemcpy2DAsync(dst, row, src, sizeof(float) * h->P, row, h->R, kind, h->stream));
    return FS2D_OK;
}

bool field_ok(int field) { return field >= FS2D_FIELD_DENSITY && field <= FS2D_FIELD_VELOCITY_V; }

// Exceptions thrown on the host (allocation in plan building, maps and
// vectors) never cross the C ABI: they become an error code and message.
template <class F>
int guarded(fs2d_handle* h, const char* what, F&& body) {
    try {
        return body();
    } catch (const std::bad_alloc&) {
        return fail(h, FS2D_ERR_HIP, std::string(what) + ": out of host memory");
    } catch (const std::exception& e) {
        return fail(h, FS2D_ERR_STATE, std::string(what) + ": " + e.what());
    } catch (...) {
        return fail(h, FS2D_ERR_STATE, std::string(what) + ": unexpected exception");
    }
}

}  // namespace

extern "C" {

int fs2d_hip_abi_version(void) { return FS2D_ABI_VERSION; }

void fs2d_default_params(fs2d_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->diffuse_iterations = 20;
    p->pressure_iterations = 20;
    p->sweep_order = FS2D_ORDER_GS_LEX;
}

static int fs2d_hip_create_impl(int rows, int cols, const fs2d_params* params, fs2d_handle** out) {
    if (!out) return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: out is NULL");
    *out = nullptr;
    if (rows != cols)
        return fail(nullptr, FS2D_ERR_ARGUMENT,
                    "fs2d_hip_create: rows != cols; the reference solver is defined for square grids only "
                    "(grid::cols() returns rows)");
    if (rows < 3 || rows > 32768)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: rows must be in [3, 32768] (ring included)");
    fs2d_params prm;
    fs2d_default_params(&prm);
    if (params) prm = *params;
    if (prm.diffuse_iterations < 0 || prm.pressure_iterations < 0 || prm.diffuse_iterations > 1000000 ||
        prm.pressure_iterations > 1000000)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: iteration counts must be in [0, 1e6]");
    if (prm.sweep_order != FS2D_ORDER_GS_LEX && prm.sweep_order != FS2D_ORDER_JACOBI)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: unknown sweep_order");
    if (prm.jacobi_block < 0 || prm.jacobi_block > 8)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: jacobi_block must be in [0, 8]");
    if (prm.gs_tile_cols != 0 && prm.gs_tile_cols != 64 && prm.gs_tile_cols != 128)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: gs_tile_cols must be 0, 64 or 128");
    if (prm.gs_block != 0 && prm.gs_block != 1 && prm.gs_block != 2 && prm.gs_block != 4 && prm.gs_block != 8)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: gs_block must be 0, 1, 2, 4 or 8");
    if (prm.launch_mode != FS2D_LAUNCH_GRAPH && prm.launch_mode != FS2D_LAUNCH_EAGER)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: launch_mode must be 0 (graph) or 1 (eager)");
    if (prm.gs_lag != 0 && prm.gs_lag != 1 && prm.gs_lag != 4 && prm.gs_lag != 8)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: gs_lag must be 0, 1, 4 or 8");
    if (prm.gs_block == 2 && prm.gs_tile_cols == 128)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: gs_tile_cols 128 takes gs_block 1, 4 or 8");
    if (prm.jacobi_kernel != 0 && prm.jacobi_kernel != FS2D_JACOBI_WAVE && prm.jacobi_kernel != FS2D_JACOBI_LDS)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_create: jacobi_kernel must be 0, 1 (wave) or 2 (LDS)");
    if (prm.sweep_order == FS2D_ORDER_GS_LEX) {
        // The Gauss-Seidel plans enumerate tiles x sweep blocks on the host: bound
        // them (the smallest candidate tiles, 64 x 64 with 4 sweeps per unit, or
        // one sweep per tile) so plan building stays small and finite.
        const long long n = rows - 2, K = std::max(prm.diffuse_iterations, prm.pressure_iterations);
        const long long g = (n + K - 1 + 63) / 64;
        const long long blocks = prm.gs_block == 1 ? K : (K + 3) / 4;
        if (g * g * blocks > (1ll << 28))
            return fail(nullptr, FS2D_ERR_ARGUMENT,
                        "fs2d_hip_create: too many Gauss-Seidel sweeps for this grid (the unit plan would hold " +
                            std::to_string(g * g * blocks) + " units; at most 2^28); use fewer iterations or the "
                            "Jacobi order");
    }

    fs2d_handle* h = new fs2d_handle();
    h->R = rows;
    h->P = (rows + 63) / 64 * 64;
    h->field_bytes = sizeof(float) * static_cast<size_t>(h->R) * h->P;
    h->prm = prm;
    h->gs_tw = prm.gs_tile_cols;
    h->gs_block = prm.gs_block;
    h->gs_lag = prm.gs_lag;
    h->jacobi_kernel = prm.jacobi_kernel ? prm.jacobi_kernel : FS2D_JACOBI_WAVE;
    auto bail = [&](hipError_t e, const char* what) {
        std::string msg = std::string("fs2d_hip_create: ") + what + ": " + hipGetErrorString(e);
        fs2d_hip_destroy(h);
        return fail(nullptr, FS2D_ERR_HIP, msg);
    };
    hipError_t e;
    if ((e = hipGetDevice(&h->device)) != hipSuccess) return bail(e, "hipGetDevice");
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    if ((e = hipEventCreate(&h->ev0)) != hipSuccess) return bail(e, "hipEventCreate");
    if ((e = hipEventCreate(&h->ev1)) != hipSuccess) return bail(e, "hipEventCreate");
    for (int i = 0; i < F_COUNT; ++i) {
        if ((e = hipMalloc(&h->f[i], h->field_bytes)) != hipSuccess) return bail(e, "hipMalloc(field)");
        if ((e = launch_fill(h->stream, h->f[i], 0u, h->field_bytes / 4)) != hipSuccess)
            return bail(e, "zero fill");
    }
    if ((e = hipMalloc(&h->max_bits, 256)) != hipSuccess) return bail(e, "hipMalloc(scalar)");
    if ((e = hipMalloc(&h->speed_slots, sizeof(unsigned) * kSpeedSlots * kSlotStride)) != hipSuccess)
        return bail(e, "hipMalloc(speed slots)");
    if ((e = launch_fill(h->stream, h->speed_slots, 0u, kSpeedSlots * kSlotStride)) != hipSuccess)
        return bail(e, "zero fill");
    if ((e = hipStreamSynchronize(h->stream)) != hipSuccess) return bail(e, "hipStreamSynchronize");
    *out = h;
    return FS2D_OK;
}

int fs2d_hip_create(int rows, int cols, const fs2d_params* params, fs2d_handle** out) {
    return guarded(nullptr, "fs2d_hip_create", [&] { return fs2d_hip_create_impl(rows, cols, params, out); });
}

int fs2d_hip_destroy(fs2d_handle* h) {
    if (!h) return FS2D_OK;
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    drop_graphs(h);
    if (h->res_sums) (void)hipFree(h->res_sums);
    if (h->render_buf) (void)hipFree(h->render_buf);
    if (h->sort_buf) (void)hipFree(h->sort_buf);
    for (auto& L : h->mg)
        for (float* q : {L.p, L.t, L.b})
            if (q) (void)hipFree(q);
    for (auto& kv : h->gs_plans) gs_plan_destroy(&kv.second);
    for (float* p : h->f)
        if (p) (void)hipFree(p);
    if (h->max_bits) (void)hipFree(h->max_bits);
    if (h->speed_slots) (void)hipFree(h->speed_slots);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return FS2D_OK;
}

const char* fs2d_hip_last_error(const fs2d_handle* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

void* fs2d_hip_stream(fs2d_handle* h) { return h ? static_cast<void*>(h->stream) : nullptr; }

static int fs2d_hip_upload_impl(fs2d_handle* h, const float* rho, const float* u, const float* v) {
    if (!h || !rho || !u || !v) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_upload: NULL argument");
    int rc;
    h->speed_valid = false;
    if ((rc = copy_in(h, h->rho(), rho, hipMemcpyHostToDevice))) return rc;
    if ((rc = copy_in(h, h->u(), u, hipMemcpyHostToDevice))) return rc;
    if ((rc = copy_in(h, h->v(), v, hipMemcpyHostToDevice))) return rc;
    FS2D_HIP(h, hipStreamSynchronize(h->stream));
    return FS2D_OK;
}

int fs2d_hip_upload(fs2d_handle* h, const float* rho, const float* u, const float* v) {
    return guarded(h, "fs2d_hip_upload", [&] { return fs2d_hip_upload_impl(h, rho, u, v); });
}

static int fs2d_hip_upload_device_impl(fs2d_handle* h, const float* rho, const float* u, const float* v) {
    if (!h || !rho || !u || !v) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_upload_device: NULL argument");
    int rc;
    h->speed_valid = false;
    if ((rc = copy_in(h, h->rho(), rho, hipMemcpyDeviceToDevice))) return rc;
    if ((rc = copy_in(h, h->u(), u, hipMemcpyDeviceToDevice))) return rc;
    return copy_in(h, h->v(), v, hipMemcpyDeviceToDevice);
}

int fs2d_hip_upload_device(fs2d_handle* h, const float* rho, const float* u, const float* v) {
    return guarded(h, "fs2d_hip_upload_device", [&] { return fs2d_hip_upload_device_impl(h, rho, u, v); });
}

static int fs2d_hip_set_sources_impl(fs2d_handle* h, const float* rs, const float* us, const float* vs) {
    if (!h) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_set_sources: NULL handle");
    const float* src[3] = {rs, us, vs};
    const int dst[3] = {F_RS, F_US, F_VS};
    for (int i = 0; i < 3; ++i) {
        if (src[i]) {
            int rc = copy_in(h, h->f[dst[i]], src[i], hipMemcpyHostToDevice);
            if (rc) return rc;
        } else {
            FS2D_HIP(h, launch_fill(h->stream, h->f[dst[i]], 0u, h->field_bytes / 4));
        }
    }
    FS2D_HIP(h, hipStreamSynchronize(h->stream));
    return FS2D_OK;
}

int fs2d_hip_set_sources(fs2d_handle* h, const float* rs, const float* us, const float* vs) {
    return guarded(h, "fs2d_hip_set_sources", [&] { return fs2d_hip_set_sources_impl(h, rs, us, vs); });
}

static int fs2d_hip_step_n_impl(fs2d_handle* h, int n, float diff, float visc, float dt) {
    if (!h || n < 0) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_step_n: bad argument");
    for (int i = 0; i < n; ++i) {
        int rc = advance(h, diff, visc, dt);
        if (rc) return rc;
    }
    return FS2D_OK;
}

int fs2d_hip_step_n(fs2d_handle* h, int n, float diff, float visc, float dt) {
    return guarded(h, "fs2d_hip_step_n", [&] { return fs2d_hip_step_n_impl(h, n, diff, visc, dt); });
}

static int fs2d_hip_graph_count_impl(fs2d_handle* h, int* count) {
    if (!h || !count) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_graph_count: bad argument");
    *count = static_cast<int>(h->graphs.size());
    return FS2D_OK;
}

int fs2d_hip_graph_count(fs2d_handle* h, int* count) {
    return guarded(h, "fs2d_hip_graph_count", [&] { return fs2d_hip_graph_count_impl(h, count); });
}

static int fs2d_hip_sync_impl(fs2d_handle* h) {
    if (!h) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_sync: NULL handle");
    FS2D_HIP(h, hipStreamSynchronize(h->stream));
    return FS2D_OK;
}

int fs2d_hip_sync(fs2d_handle* h) {
    return guarded(h, "fs2d_hip_sync", [&] { return fs2d_hip_sync_impl(h); });
}

static int fs2d_hip_download_impl(fs2d_handle* h, float* rho, float* u, float* v) {
    if (!h) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_download: NULL handle");
    int rc;
    if (rho && (rc = copy_out(h, rho, h->rho(), hipMemcpyDeviceToHost))) return rc;
    if (u && (rc = copy_out(h, u, h->u(), hipMemcpyDeviceToHost))) return rc;
    if (v && (rc = copy_out(h, v, h->v(), hipMemcpyDeviceToHost))) return rc;
    FS2D_HIP(h, hipStreamSynchronize(h->stream));
    return FS2D_OK;
}

int fs2d_hip_download(fs2d_handle* h, float* rho, float* u, float* v) {
    return guarded(h, "fs2d_hip_download", [&] { return fs2d_hip_download_impl(h, rho, u, v); });
}

static int fs2d_hip_download_device_impl(fs2d_handle* h, float* rho, float* u, float* v) {
    if (!h) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_download_device: NULL handle");
    int rc;
    if (rho && (rc = copy_out(h, rho, h->rho(), hipMemcpyDeviceToDevice))) return rc;
    if (u && (rc = copy_out(h, u, h->u(), hipMemcpyDeviceToDevice))) return rc;
    if (v && (rc = copy_out(h, v, h->v(), hipMemcpyDeviceToDevice))) return rc;
    return FS2D_OK;
}

int fs2d_hip_download_device(fs2d_handle* h, float* rho, float* u, float* v) {
    return guarded(h, "fs2d_hip_download_device", [&] { return fs2d_hip_download_device_impl(h, rho, u, v); });
}

static int fs2d_hip_download_pressure_impl(fs2d_handle* h, float* p) {
    if (!h || !p) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_download_pressure: NULL argument");
    if (h->pressure_zero) {
        std::memset(p, 0, sizeof(float) * static_cast<size_t>(h->R) * h->R);
        return FS2D_OK;
    }
    int rc = copy_out(h, p, h->f[F_P], hipMemcpyDeviceToHost);
    if (rc) return rc;
    FS2D_HIP(h, hipStreamSynchronize(h->stream));
    return FS2D_OK;
}

int fs2d_hip_download_pressure(fs2d_handle* h, float* p) {
    return guarded(h, "fs2d_hip_download_pressure", [&] { return fs2d_hip_download_pressure_impl(h, p); });
}

static int fs2d_hip_solve_impl(fs2d_handle* h, float* rho, const float* rho_src, float diff, float* u, float* v,
                   const float* u_src, const float* v_src, float visc, float dt) {
    if (!h || !rho || !rho_src || !u || !v || !u_src || !v_src)
        return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_solve: NULL argument");
    int rc;
    if ((rc = fs2d_hip_upload(h, rho, u, v))) return rc;
    if ((rc = fs2d_hip_set_sources(h, rho_src, u_src, v_src))) return rc;
    if ((rc = advance(h, diff, visc, dt))) return rc;
    return fs2d_hip_download(h, rho, u, v);
}

int fs2d_hip_solve(fs2d_handle* h, float* rho, const float* rho_src, float diff, float* u, float* v,
                   const float* u_src, const float* v_src, float visc, float dt) {
    return guarded(h, "fs2d_hip_solve", [&] { return fs2d_hip_solve_impl(h, rho, rho_src, diff, u, v, u_src, v_src, visc, dt); });
}

// ---- stages ------------------------------------------------------------------
static float*& field_ref(fs2d_handle* h, int field) {
    return field == FS2D_FIELD_DENSITY ? h->rho() : (field == FS2D_FIELD_VELOCITY_U ? h->u() : h->v());
}

static int fs2d_hip_stage_add_sources_impl(fs2d_handle* h, int field, float dt) {
    if (!h || !field_ok(field)) return fail(h, FS2D_ERR_ARGUMENT, "stage_add_sources: bad argument");
    const int src = field == FS2D_FIELD_DENSITY ? F_RS : (field == FS2D_FIELD_VELOCITY_U ? F_US : F_VS);
    if (field != FS2D_FIELD_DENSITY) h->speed_valid = false;
    FS2D_HIP(h, launch_add_sources(h->stream, field_ref(h, field), nullptr, h->f[src], h->R, h->P, dt));
    return FS2D_OK;
}

int fs2d_hip_stage_add_sources(fs2d_handle* h, int field, float dt) {
    return guarded(h, "fs2d_hip_stage_add_sources", [&] { return fs2d_hip_stage_add_sources_impl(h, field, dt); });
}

static int fs2d_hip_stage_diffuse_impl(fs2d_handle* h, int field, float rate, float dt, int iterations) {
    if (!h || !field_ok(field) || iterations < 0)
        return fail(h, FS2D_ERR_ARGUMENT, "stage_diffuse: bad argument");
    float*& x = field_ref(h, field);
    if (field != FS2D_FIELD_DENSITY) h->speed_valid = false;
    FS2D_HIP(h, hipMemcpyAsync(h->f[F_X0], x, h->field_bytes, hipMemcpyDeviceToDevice, h->stream));
    return diffuse(h, x, field, rate, dt, iterations);
}

int fs2d_hip_stage_diffuse(fs2d_handle* h, int field, float rate, float dt, int iterations) {
    return guarded(h, "fs2d_hip_stage_diffuse", [&] { return fs2d_hip_stage_diffuse_impl(h, field, rate, dt, iterations); });
}

static int fs2d_hip_stage_advect_density_impl(fs2d_handle* h, float dt) {
    if (!h) return fail(h, FS2D_ERR_ARGUMENT, "stage_advect_density: NULL handle");
    return advect_density(h, dt);
}

int fs2d_hip_stage_advect_density(fs2d_handle* h, float dt) {
    return guarded(h, "fs2d_hip_stage_advect_density", [&] { return fs2d_hip_stage_advect_density_impl(h, dt); });
}

static int fs2d_hip_stage_advect_velocity_impl(fs2d_handle* h, float dt) {
    if (!h) return fail(h, FS2D_ERR_ARGUMENT, "stage_advect_velocity: NULL handle");
    return advect_velocity(h, dt);
}

int fs2d_hip_stage_advect_velocity(fs2d_handle* h, float dt) {
    return guarded(h, "fs2d_hip_stage_advect_velocity", [&] { return fs2d_hip_stage_advect_velocity_impl(h, dt); });
}

static int fs2d_hip_stage_project_impl(fs2d_handle* h, int iterations) {
    if (!h || iterations < 0) return fail(h, FS2D_ERR_ARGUMENT, "stage_project: bad argument");
    return project(h, iterations);
}

int fs2d_hip_stage_project(fs2d_handle* h, int iterations) {
    return guarded(h, "fs2d_hip_stage_project", [&] { return fs2d_hip_stage_project_impl(h, iterations); });
}

static int fs2d_hip_stage_set_boundary_impl(fs2d_handle* h, int field) {
    if (!h || !field_ok(field)) return fail(h, FS2D_ERR_ARGUMENT, "stage_set_boundary: bad argument");
    if (field != FS2D_FIELD_DENSITY) h->speed_valid = false;
    FS2D_HIP(h, launch_set_boundary(h->stream, field_ref(h, field), h->R, h->P, ring_of(field)));
    return FS2D_OK;
}

int fs2d_hip_stage_set_boundary(fs2d_handle* h, int field) {
    return guarded(h, "fs2d_hip_stage_set_boundary", [&] { return fs2d_hip_stage_set_boundary_impl(h, field); });
}

// ---- measurement -------------------------------------------------------------
static int fs2d_hip_time_pressure_sweeps_impl(fs2d_handle* h, int reps, float* ms) {
    if (!h || reps <= 0 || !ms) return fail(h, FS2D_ERR_ARGUMENT, "time_pressure_sweeps: bad argument");
    const Scalars s = scalars(h->R, 1.0f);
    const int K = h->prm.pressure_iterations;
    int rc;
    // as a step's second projection runs it: zero fill, then the sweeps from
    // the stored divergence
    auto once = [&]() -> int { return linear_solve(h, h->f[F_P], h->f[F_D], SWEEP_PRESSURE, 0.f, 0.f, BND_CONT, K, 1); };
    FS2D_HIP(h, launch_divergence(h->stream, h->f[F_D], h->u(), h->v(), h->R, h->P, s.neg_half_h, nullptr));
    if ((rc = once())) return rc;  // warm
    FS2D_HIP(h, hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < reps; ++i)
        if ((rc = once())) return rc;
    FS2D_HIP(h, hipEventRecord(h->ev1, h->stream));
    FS2D_HIP(h, hipEventSynchronize(h->ev1));
    float total = 0.f;
    FS2D_HIP(h, hipEventElapsedTime(&total, h->ev0, h->ev1));
    *ms = total / reps;
    return FS2D_OK;
}

int fs2d_hip_time_pressure_sweeps(fs2d_handle* h, int reps, float* ms) {
    return guarded(h, "fs2d_hip_time_pressure_sweeps", [&] { return fs2d_hip_time_pressure_sweeps_impl(h, reps, ms); });
}

static int fs2d_hip_time_joint_sweeps_impl(fs2d_handle* h, float diff, float visc, float dt, int reps, float* ms) {
    if (!h || reps <= 0 || !ms) return fail(h, FS2D_ERR_ARGUMENT, "time_joint_sweeps: bad argument");
    const int Kd = h->prm.diffuse_iterations, Kp = h->prm.pressure_iterations;
    const int R = h->R, P = h->P;
    if (h->prm.sweep_order != FS2D_ORDER_GS_LEX || !gs_units(h) || Kd <= 0 || Kp <= 0)
        return fail(h, FS2D_ERR_STATE, "time_joint_sweeps: Gauss-Seidel units with diffuse and pressure sweeps only");
    JointPlans joint;
    FS2D_HIP(h, joint_plans(h, Kd, Kp, &joint));
    if (!joint.d) return fail(h, FS2D_ERR_STATE, "time_joint_sweeps: Gauss-Seidel units only");
    // iterates and right-hand sides are copies of the state, which stays untouched;
    // the pressure iterate is the (otherwise unused) scratch field
    const size_t n4 = h->field_bytes / 16;
    const int pairs[6][2] = {{F_X0, F_RHO}, {F_X0U, F_U}, {F_X0V, F_V}, {F_T1, F_RHO}, {F_UN, F_U}, {F_VN, F_V}};
    for (const auto& pr : pairs)
        FS2D_HIP(h, launch_copy4(h->stream, reinterpret_cast<float4*>(h->f[pr[0]]),
                                 reinterpret_cast<const float4*>(h->f[pr[1]]), n4));
    const float ad = dt * cells(R) * diff, av = dt * cells(R) * visc;
    SweepSet set{};
    set.count = 3;
    set.a[0] = sweep_args(h->f[F_T1], h->f[F_X0], R, P, SWEEP_DIFFUSE, ad, 1.f + 4.f * ad, BND_CONT, Kd, 0);
    set.a[1] = sweep_args(h->f[F_UN], h->f[F_X0U], R, P, SWEEP_DIFFUSE, av, 1.f + 4.f * av, BND_OPP_H, Kd, 0);
    set.a[2] = sweep_args(h->f[F_VN], h->f[F_X0V], R, P, SWEEP_DIFFUSE, av, 1.f + 4.f * av, BND_OPP_V, Kd, 0);
    SweepArgs A = sweep_args(h->f[F_SCRATCH], h->f[F_D], R, P, SWEEP_PRESSURE, 0.f, 0.f, BND_CONT, Kp, 0);
    A.du = h->f[F_UN];
    A.dv = h->f[F_VN];
    A.neg_half_h = scalars(R, 1.0f).neg_half_h;
    auto once = [&]() -> int {
        FS2D_HIP(h, launch_fill(h->stream, h->f[F_SCRATCH], 0u, h->field_bytes / 4));
        FS2D_HIP(h, launch_gs_pipelined(h->stream, *joint.d, set, *joint.p, A, joint.offset));
        return FS2D_OK;
    };
    int rc = once();  // warm
    if (rc) return rc;
    FS2D_HIP(h, hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < reps; ++i)
        if ((rc = once())) return rc;
    FS2D_HIP(h, hipEventRecord(h->ev1, h->stream));
    FS2D_HIP(h, hipEventSynchronize(h->ev1));
    float total = 0.f;
    FS2D_HIP(h, hipEventElapsedTime(&total, h->ev0, h->ev1));
    *ms = total / reps;
    return FS2D_OK;
}

int fs2d_hip_time_joint_sweeps(fs2d_handle* h, float diff, float visc, float dt, int reps, float* ms) {
    return guarded(h, "fs2d_hip_time_joint_sweeps", [&] { return fs2d_hip_time_joint_sweeps_impl(h, diff, visc, dt, reps, ms); });
}

static int fs2d_hip_time_diffuse_sweeps_impl(fs2d_handle* h, float diff, float visc, float dt, int reps, float* ms) {
    if (!h || reps <= 0 || !ms) return fail(h, FS2D_ERR_ARGUMENT, "time_diffuse_sweeps: bad argument");
    const int Kd = h->prm.diffuse_iterations;
    const int R = h->R, P = h->P;
    const size_t n4 = h->field_bytes / 16;
    // right-hand sides and iterates are copies of the state, which stays untouched
    const int pairs[6][2] = {{F_X0, F_RHO}, {F_X0U, F_U}, {F_X0V, F_V}, {F_T1, F_RHO}, {F_UN, F_U}, {F_VN, F_V}};
    for (const auto& pr : pairs)
        FS2D_HIP(h, launch_copy4(h->stream, reinterpret_cast<float4*>(h->f[pr[0]]),
                                 reinterpret_cast<const float4*>(h->f[pr[1]]), n4));
    const float ad = dt * cells(R) * diff, av = dt * cells(R) * visc;
    const float dd = 1.f + 4.f * ad, dv = 1.f + 4.f * av;
    auto once = [&]() -> int {
        if (Kd <= 0) return FS2D_OK;
        if (h->prm.sweep_order == FS2D_ORDER_GS_LEX && gs_units(h)) {  // as the step relaxes them
            SweepSet set{};
            set.count = 3;
            set.a[0] = sweep_args(h->f[F_T1], h->f[F_X0], R, P, SWEEP_DIFFUSE, ad, dd, BND_CONT, Kd, 0);
            set.a[1] = sweep_args(h->f[F_UN], h->f[F_X0U], R, P, SWEEP_DIFFUSE, av, dv, BND_OPP_H, Kd, 0);
            set.a[2] = sweep_args(h->f[F_VN], h->f[F_X0V], R, P, SWEEP_DIFFUSE, av, dv, BND_OPP_V, Kd, 0);
            const GsPlan* plan = nullptr;
            FS2D_HIP(h, gs_plan_for(h, Kd, SWEEP_DIFFUSE, &plan));
            FS2D_HIP(h, launch_gs_solve_set(h->stream, *plan, set));
            return FS2D_OK;
        }
        int rc;
        if ((rc = linear_solve(h, h->f[F_T1], h->f[F_X0], SWEEP_DIFFUSE, ad, dd, BND_CONT, Kd, 0))) return rc;
        if ((rc = linear_solve(h, h->f[F_UN], h->f[F_X0U], SWEEP_DIFFUSE, av, dv, BND_OPP_H, Kd, 0))) return rc;
        return linear_solve(h, h->f[F_VN], h->f[F_X0V], SWEEP_DIFFUSE, av, dv, BND_OPP_V, Kd, 0);
    };
    int rc = once();  // warm
    if (rc) return rc;
    FS2D_HIP(h, hipEventRecord(h->ev0, h->stream));
    for (int i = 0; i < reps; ++i)
        if ((rc = once())) return rc;
    FS2D_HIP(h, hipEventRecord(h->ev1, h->stream));
    FS2D_HIP(h, hipEventSynchronize(h->ev1));
    float total = 0.f;
    FS2D_HIP(h, hipEventElapsedTime(&total, h->ev0, h->ev1));
    *ms = total / reps;
    return FS2D_OK;
}

int fs2d_hip_time_diffuse_sweeps(fs2d_handle* h, float diff, float visc, float dt, int reps, float* ms) {
    return guarded(h, "fs2d_hip_time_diffuse_sweeps", [&] { return fs2d_hip_time_diffuse_sweeps_impl(h, diff, visc, dt, reps, ms); });
}

static int fs2d_hip_gs_policy_impl(fs2d_handle* h, int kind, int* tile_cols, int* sweeps_per_unit, int* lag, int* levels) {
    if (!h || kind < 0 || kind > 2) return fail(h, FS2D_ERR_ARGUMENT, "gs_policy: bad argument");
    if (h->prm.sweep_order != FS2D_ORDER_GS_LEX) return fail(h, FS2D_ERR_STATE, "gs_policy: Gauss-Seidel order only");
    if (kind == 2) {  // the joint level launches of a step
        JointPlans joint;
        FS2D_HIP(h, joint_plans(h, h->prm.diffuse_iterations, h->prm.pressure_iterations, &joint));
        if (!joint.d) return fail(h, FS2D_ERR_STATE, "gs_policy: no joint launches");
        if (tile_cols) *tile_cols = joint.d->TW;
        if (sweeps_per_unit) *sweeps_per_unit = joint.d->M;
        if (lag) *lag = joint.d->lag;
        if (levels) *levels = joint_levels(joint);
        return FS2D_OK;
    }
    const int K = kind == 0 ? h->prm.diffuse_iterations : h->prm.pressure_iterations;
    if (K <= 0) return fail(h, FS2D_ERR_STATE, "gs_policy: no sweeps of that kind");
    const int sk = kind == 0 ? SWEEP_DIFFUSE : SWEEP_PRESSURE;
    GsPolicy c{};
    FS2D_HIP(h, gs_policy_for(h, K, sk, &c));
    const GsPlan* plan = nullptr;
    FS2D_HIP(h, gs_plan_of(h, K, c, &plan));
    if (tile_cols) *tile_cols = c.tw;
    if (sweeps_per_unit) *sweeps_per_unit = c.M == 0 ? 1 : c.M;
    if (lag) *lag = c.lag;
    if (levels) *levels = plan->levels;
    return FS2D_OK;
}

int fs2d_hip_gs_policy(fs2d_handle* h, int kind, int* tile_cols, int* sweeps_per_unit, int* lag, int* levels) {
    return guarded(h, "fs2d_hip_gs_policy", [&] { return fs2d_hip_gs_policy_impl(h, kind, tile_cols, sweeps_per_unit, lag, levels); });
}

int fs2d_gs_pipeline_offset(int rows, int diffuse_iterations, int pressure_iterations, int tile_cols,
                            int sweeps_per_unit, int* offset) {
    if (rows < 3 || diffuse_iterations < 1 || pressure_iterations < 1 || !offset ||
        (tile_cols != 64 && tile_cols != 128) ||
        (sweeps_per_unit != 2 && sweeps_per_unit != 4 && sweeps_per_unit != 8))
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_gs_pipeline_offset: bad argument");
    GsPlan pp{};
    const hipError_t e = gs_unit_plan_build(&pp, rows, pressure_iterations, tile_cols, sweeps_per_unit, 4, false);
    if (e != hipSuccess) {
        gs_plan_destroy(&pp);
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_gs_pipeline_offset: plan too large");
    }
    *offset = gs_pipeline_offset(pp, rows, diffuse_iterations, tile_cols, sweeps_per_unit);
    gs_plan_destroy(&pp);
    return FS2D_OK;
}

int fs2d_division_plan(float denom, float* inv, float* inv_lo, float* two_hi, float* two_lo, float* lo,
                       int* checked) {
    if (!inv || !inv_lo || !two_hi || !two_lo || !lo || !checked)
        return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_division_plan: null output");
    SweepArgs A = sweep_args(nullptr, nullptr, 3, 64, SWEEP_DIFFUSE, 0.f, denom, BND_CONT, 1, 0);
    two_product_quotient(A);
    *inv = A.inv_denom;
    *inv_lo = A.inv_denom_lo;
    *two_hi = A.inv_hi2;
    *two_lo = A.inv_lo2;
    *lo = A.div_lo;
    *checked = A.div_checked;
    return FS2D_OK;
}

static int fs2d_hip_copy_bandwidth_impl(long long bytes, int reps, float* gbs) {
    if (bytes < 1024 || reps < 1 || !gbs) return fail(nullptr, FS2D_ERR_ARGUMENT, "fs2d_hip_copy_bandwidth: bad argument");
    const size_t n4 = static_cast<size_t>(bytes) / 16;
    float4 *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipMalloc(&a, n4 * 16);
    if (e == hipSuccess) e = hipMalloc(&b, n4 * 16);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = launch_fill(st, a, 0x3f800000u, n4 * 4);
    if (e == hipSuccess) e = launch_copy4(st, b, a, n4);  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_copy4(st, (r & 1) ? a : b, (r & 1) ? b : a, n4);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) return fail(nullptr, FS2D_ERR_HIP, std::string("fs2d_hip_copy_bandwidth: ") + hipGetErrorString(e));
    *gbs = static_cast<float>(2.0 * n4 * 16 * reps / (ms * 1e-3) / 1e9);
    return FS2D_OK;
}

int fs2d_hip_copy_bandwidth(long long bytes, int reps, float* gbs) {
    return guarded(nullptr, "fs2d_hip_copy_bandwidth", [&] { return fs2d_hip_copy_bandwidth_impl(bytes, reps, gbs); });
}

static int fs2d_hip_launches_per_step_impl(fs2d_handle* h, int* launches) {
    if (!h || !launches) return fail(h, FS2D_ERR_ARGUMENT, "launches_per_step: bad argument");
    const int Kd = h->prm.diffuse_iterations, Kp = h->prm.pressure_iterations;
    auto solve_launches = [&](int K, int kind) -> int {
        if (K <= 0) return 0;
        if (h->prm.sweep_order == FS2D_ORDER_JACOBI) return jacobi_launches(K, jacobi_depth_known(h, kind, K));
        const GsPlan* plan = nullptr;
        if (gs_plan_for(h, K, kind, &plan) != hipSuccess) return -1;
        return plan->levels;
    };
    const int d = solve_launches(Kd, SWEEP_DIFFUSE), p = solve_launches(Kp, SWEEP_PRESSURE);
    if (d < 0 || p < 0) return fail(h, FS2D_ERR_HIP, "launches_per_step: plan failed");
    const bool shared = h->prm.sweep_order == FS2D_ORDER_GS_LEX && gs_units(h);
    JointPlans joint;
    if (joint_plans(h, Kd, Kp, &joint) != hipSuccess) return fail(h, FS2D_ERR_HIP, "launches_per_step: plan failed");
    if (joint.d) {
        // pipelined: the three adds (one launch), fill p, the joint level launches,
        // scatter, gradient + fold, gather, then the divergence (filling p), the
        // second solve's levels, gradient + fold
        *launches = 1 + 1 + joint_levels(joint) + scatter_launches(h->R) + 2 + 1 + 1 + p + 2;
        return FS2D_OK;
    }
    // adds (one launch shared, x3 otherwise), diffuse (x3, or x1 shared), scatter,
    // project x2 (div + solve + grad), gather
    *launches = (shared ? 1 : 3) + (shared ? d : 3 * d) + scatter_launches(h->R) + 2 * (2 + p) + 1;
    return FS2D_OK;
}

int fs2d_hip_launches_per_step(fs2d_handle* h, int* launches) {
    return guarded(h, "fs2d_hip_launches_per_step", [&] { return fs2d_hip_launches_per_step_impl(h, launches); });
}

// ---- the seeded projection (include/fs2d_cnn.h) ---------------------------
static int residual_of(fs2d_handle* h, double* rel) {
    if (!h->res_sums) FS2D_HIP(h, hipMalloc(&h->res_sums, 2 * sizeof(double)));
    FS2D_HIP(h, launch_fill(h->stream, h->res_sums, 0u, 4));
    FS2D_HIP(h, launch_pressure_residual(h->stream, h->f[F_D], h->f[F_P], h->R, h->P, h->res_sums));
    double sums[2] = {0, 0};
    FS2D_HIP(h, hipMemcpyAsync(sums, h->res_sums, sizeof sums, hipMemcpyDeviceToHost, h->stream));
    FS2D_HIP(h, hipStreamSynchronize(h->stream));
    *rel = sums[1] > 0 ? std::sqrt(sums[0] / sums[1]) : std::sqrt(sums[0]);
    return FS2D_OK;
}

static int fs2d_hip_pressure_residual_impl(fs2d_handle* h, double* rel) {
    if (!h || !rel) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_pressure_residual: bad argument");
    return residual_of(h, rel);
}

int fs2d_hip_pressure_residual(fs2d_handle* h, double* rel) {
    return guarded(h, "fs2d_hip_pressure_residual", [&] { return fs2d_hip_pressure_residual_impl(h, rel); });
}

// Jacobi launches on the handle's pressure field until the relative residual
// is <= tol (checked every `check` sweeps) or max_sweeps have run.  A
// non-finite residual stops it at once: a NaN or infinite iterate stays
// non-finite under the sweeps, so the budget would be spent for nothing.
static int jacobi_until(fs2d_handle* h, float tol, int max_sweeps, int check, int* sweeps, double* rel) {
    const int depth = jacobi_depth_known(h, SWEEP_PRESSURE, h->prm.pressure_iterations);
    const SweepArgs A = sweep_args(h->f[F_P], h->f[F_D], h->R, h->P, SWEEP_PRESSURE, 0.f, 0.f, BND_CONT, max_sweeps, 0);
    int rc;
    if ((rc = residual_of(h, rel))) return rc;
    int done = 0, since = 0;
    while (done < max_sweeps && !(*rel <= tol) && std::isfinite(*rel)) {
        const int m = std::min(depth, max_sweeps - done);
        FS2D_HIP(h, launch_jacobi_pass(h->stream, A, h->f[F_P], h->f[F_SCRATCH], 0, m, h->jacobi_kernel));
        std::swap(h->f[F_P], h->f[F_SCRATCH]);
        done += m;
        since += m;
        if (since >= check || done >= max_sweeps) {
            since = 0;
            if ((rc = residual_of(h, rel))) return rc;
        }
    }
    *sweeps = done;
    return FS2D_OK;
}

// ---- multigrid seed: V(2,2) cycles of weighted Jacobi on grids of side n, n/2, ...
constexpr float kMgOmega = 0.8f;
constexpr int kMgPre = 2, kMgPost = 2, kMgMaxCycles = 30;

// Sweeps on the coarsest grid: its slowest mode decays as 1 - omega pi^2 / (2 n^2)
// per sweep, so n^2 sweeps reduce it by ~50x (64 at least, 4096 at most).
static int mg_coarsest_sweeps(int n) { return std::min(4096, std::max(64, n * n)); }

static void mg_free(fs2d_handle* h) {
    for (auto& L : h->mg)
        for (float* q : {L.p, L.t, L.b})
            if (q) (void)hipFree(q);
    h->mg.clear();
}

// Coarse grids of the multigrid seed.  Either every level is allocated and
// cleared, or none is kept: a level enters h->mg only with all three of its
// arrays, and any failure frees what was allocated so far.
static int mg_levels(fs2d_handle* h) {
    if (!h->mg.empty()) return FS2D_OK;
    int n = h->R - 2;
    while (n % 2 == 0 && n / 2 >= 4) {
        n /= 2;
        fs2d_handle::MgLevel L{n, n + 2, (n + 2 + 63) / 64 * 64, nullptr, nullptr, nullptr};
        const size_t bytes = sizeof(float) * static_cast<size_t>(L.R) * L.P;
        hipError_t e = hipMalloc(&L.p, bytes);
        if (e == hipSuccess) e = hipMalloc(&L.t, bytes);
        if (e == hipSuccess) e = hipMalloc(&L.b, bytes);
        if (e == hipSuccess) e = launch_fill(h->stream, L.b, 0u, bytes / 4);
        if (e != hipSuccess) {
            for (float* q : {L.p, L.t, L.b})
                if (q) (void)hipFree(q);
            mg_free(h);
            return fail(h, FS2D_ERR_HIP, std::string("multigrid levels: ") + hipGetErrorString(e));
        }
        h->mg.push_back(L);  // owned by the handle from here on, freed in destroy
    }
    return FS2D_OK;
}

// One V-cycle on level l (0: the handle's pressure, scratch and divergence
// fields); the result is left in the level's iterate pointer.
static int mg_vcycle(fs2d_handle* h, size_t l) {
    float** p = l == 0 ? &h->f[F_P] : &h->mg[l - 1].p;
    float** t = l == 0 ? &h->f[F_SCRATCH] : &h->mg[l - 1].t;
    const float* b = l == 0 ? h->f[F_D] : h->mg[l - 1].b;
    const int R = l == 0 ? h->R : h->mg[l - 1].R, P = l == 0 ? h->P : h->mg[l - 1].P;
    auto smooth = [&](int count) -> int {
        for (int k = 0; k < count; ++k) {
            FS2D_HIP(h, launch_mg_smooth(h->stream, *p, *t, b, R, P, kMgOmega));
            std::swap(*p, *t);
        }
        return FS2D_OK;
    };
    int rc;
    if (l == h->mg.size()) return smooth(mg_coarsest_sweeps(R - 2));
    if ((rc = smooth(kMgPre))) return rc;
    fs2d_handle::MgLevel& C = h->mg[l];
    FS2D_HIP(h, launch_mg_restrict(h->stream, *p, b, P, C.b, C.R, C.P));
    FS2D_HIP(h, launch_fill(h->stream, C.p, 0u, static_cast<size_t>(C.R) * C.P));
    if ((rc = mg_vcycle(h, l + 1))) return rc;
    FS2D_HIP(h, launch_mg_prolong(h->stream, C.p, C.P, *p, R, P));
    return smooth(kMgPost);
}

static int fs2d_hip_project_seeded_impl(fs2d_handle* h, fs2d_net* net, int seed, float tol, int max_sweeps, int check,
                            int fallback, int* sweeps_run, double* rel_residual, int* fell_back) {
    if (!h || seed < 0 || seed > 3 || max_sweeps < 0 || check < 1 || !(tol >= 0.f))
        return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_project_seeded: bad argument");
    if (h->prm.sweep_order != FS2D_ORDER_JACOBI)
        return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_project_seeded: needs the Jacobi order");
    if (seed == 2 && (!net || (h->R - 2) % 32 != 0))
        return fail(h, FS2D_ERR_ARGUMENT,
                    "fs2d_hip_project_seeded: seed 2 needs a network and a side n = rows - 2 that is a multiple of 32");
    const Scalars s = scalars(h->R, 1.0f);
    FS2D_HIP(h, launch_divergence(h->stream, h->f[F_D], h->u(), h->v(), h->R, h->P, s.neg_half_h, nullptr));
    if (seed == 0 || seed == 3 || h->pressure_zero) FS2D_HIP(h, launch_fill(h->stream, h->f[F_P], 0u, h->field_bytes / 4));
    const bool keep_prev = seed == 2 && fallback;
    if (keep_prev)  // the previous pressure, for the fallback
        FS2D_HIP(h, hipMemcpyAsync(h->f[F_T1], h->f[F_P], h->field_bytes, hipMemcpyDeviceToDevice, h->stream));
    if (seed == 2) {
        const int rc = net_predict_fields(net, h->stream, h->f[F_D], h->f[F_P], h->R, h->P, h->f[F_SCRATCH]);
        if (rc) return fail(h, rc, std::string("fs2d_hip_project_seeded: network: ") + net_error(net));
        std::swap(h->f[F_P], h->f[F_SCRATCH]);
    }
    int sweeps = 0, rc;
    double rel = 0;
    int fine_smooth = 0;  // weighted Jacobi sweeps of the fine grid in V-cycles
    if (seed == 3) {
        if ((rc = mg_levels(h))) return rc;
        for (const auto& L : h->mg)
            if (!L.p || !L.t || !L.b) return fail(h, FS2D_ERR_STATE, "fs2d_hip_project_seeded: multigrid level not allocated");
        for (int c = 0; c < kMgMaxCycles; ++c) {
            if ((rc = mg_vcycle(h, 0))) return rc;
            fine_smooth += h->mg.empty() ? mg_coarsest_sweeps(h->R - 2) : kMgPre + kMgPost;
            if ((rc = residual_of(h, &rel))) return rc;
            if (rel <= tol) break;
        }
    }
    if ((rc = jacobi_until(h, tol, max_sweeps, check, &sweeps, &rel))) return rc;
    sweeps += fine_smooth;
    int fb = 0;
    if (keep_prev && !(rel <= tol)) {
        FS2D_HIP(h, hipMemcpyAsync(h->f[F_P], h->f[F_T1], h->field_bytes, hipMemcpyDeviceToDevice, h->stream));
        int more = 0;
        if ((rc = jacobi_until(h, tol, max_sweeps, check, &more, &rel))) return rc;
        sweeps += more;
        fb = 1;
    }
    h->pressure_zero = false;
    FS2D_HIP(h, launch_gradient(h->stream, h->u(), h->v(), h->f[F_P], h->R, h->P, s.h, 0, h->max_bits,
                                h->speed_slots));
    h->speed_valid = true;
    FS2D_HIP(h, hipStreamSynchronize(h->stream));
    if (sweeps_run) *sweeps_run = sweeps;
    if (rel_residual) *rel_residual = rel;
    if (fell_back) *fell_back = fb;
    return FS2D_OK;
}

int fs2d_hip_project_seeded(fs2d_handle* h, fs2d_net* net, int seed, float tol, int max_sweeps, int check,
                            int fallback, int* sweeps_run, double* rel_residual, int* fell_back) {
    return guarded(h, "fs2d_hip_project_seeded", [&] { return fs2d_hip_project_seeded_impl(h, net, seed, tol, max_sweeps, check, fallback, sweeps_run, rel_residual, fell_back); });
}

}  // extern "C"

// ---- visual output (src/density_grid_renderer.cu, src/velocity_grid_renderer.cu) ----
namespace {

int render_staging(fs2d_handle* h, size_t bytes, void** out) {
    if (h->render_bytes < bytes) {
        FS2D_HIP(h, hipStreamSynchronize(h->stream));
        if (h->render_buf) FS2D_HIP(h, hipFree(h->render_buf));
        h->render_buf = nullptr;
        h->render_bytes = 0;
        FS2D_HIP(h, hipMalloc(&h->render_buf, bytes));
        h->render_bytes = bytes;
    }
    *out = h->render_buf;
    return FS2D_OK;
}

size_t cell_count(const fs2d_handle* h) { return static_cast<size_t>(h->R) * static_cast<size_t>(h->R); }

}  // namespace

int fs2d_hip_render_density_device(fs2d_handle* h, const fs2d_color_multipliers* m, unsigned char* rgba) {
    if (!h || !m || !rgba) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_render_density: NULL argument");
    FS2D_HIP(h, launch_render_density(h->stream, h->rho(), h->R, h->P, m->r, m->g, m->b, rgba));
    return FS2D_OK;
}

int fs2d_hip_render_density(fs2d_handle* h, const fs2d_color_multipliers* m, unsigned char* rgba) {
    if (!h || !m || !rgba) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_render_density: NULL argument");
    const size_t bytes = 4 * cell_count(h);
    void* buf = nullptr;
    int rc = render_staging(h, bytes, &buf);
    if (rc) return rc;
    if ((rc = fs2d_hip_render_density_device(h, m, static_cast<unsigned char*>(buf)))) return rc;
    FS2D_HIP(h, hipMemcpyAsync(rgba, buf, bytes, hipMemcpyDeviceToHost, h->stream));
    FS2D_HIP(h, hipStreamSynchronize(h->stream));
    return FS2D_OK;
}

int fs2d_hip_render_velocity_device(fs2d_handle* h, float hs, float vs, fs2d_line* lines) {
    static_assert(sizeof(fs2d_vertex) == 20 && sizeof(fs2d_line) == 40, "SFML vertex layout");
    if (!h || !lines) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_render_velocity: NULL argument");
    FS2D_HIP(h, launch_render_velocity(h->stream, h->u(), h->v(), h->R, h->P, hs, vs, lines));
    return FS2D_OK;
}

int fs2d_hip_render_velocity(fs2d_handle* h, float hs, float vs, fs2d_line* lines) {
    if (!h || !lines) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_render_velocity: NULL argument");
    const size_t bytes = sizeof(fs2d_line) * cell_count(h);
    void* buf = nullptr;
    int rc = render_staging(h, bytes, &buf);
    if (rc) return rc;
    if ((rc = fs2d_hip_render_velocity_device(h, hs, vs, static_cast<fs2d_line*>(buf)))) return rc;
    FS2D_HIP(h, hipMemcpyAsync(lines, buf, bytes, hipMemcpyDeviceToHost, h->stream));
    FS2D_HIP(h, hipStreamSynchronize(h->stream));
    return FS2D_OK;
}

int fs2d_hip_time_render(fs2d_handle* h, int reps, float* ms_density, float* ms_velocity) {
    if (!h || reps < 1 || !ms_density || !ms_velocity) return fail(h, FS2D_ERR_ARGUMENT, "fs2d_hip_time_render: bad argument");
    void* buf = nullptr;
    int rc = render_staging(h, sizeof(fs2d_line) * cell_count(h), &buf);
    if (rc) return rc;
    const fs2d_color_multipliers m{1.f, 0.5f, 0.25f};
    for (int pass = 0; pass < 2; ++pass) {
        auto once = [&]() {
            return pass == 0 ? fs2d_hip_render_density_device(h, &m, static_cast<unsigned char*>(buf))
                             : fs2d_hip_render_velocity_device(h, 1.f, 1.f, static_cast<fs2d_line*>(buf));
        };
        if ((rc = once())) return rc;  // warm
        FS2D_HIP(h, hipEventRecord(h->ev0, h->stream));
        for (int r = 0; r < reps; ++r)
            if ((rc = once())) return rc;
        FS2D_HIP(h, hipEventRecord(h->ev1, h->stream));
        FS2D_HIP(h, hipEventSynchronize(h->ev1));
        float ms = 0.f;
        FS2D_HIP(h, hipEventElapsedTime(&ms, h->ev0, h->ev1));
        (pass == 0 ? *ms_density : *ms_velocity) = ms / reps;
    }
    return FS2D_OK;
}
