// fs2d_internal.hpp -- shared declarations of the HIP stable-fluids solver.
//
// Device layout: every field is an R x R float grid (ring included) stored
// row-major with a row pitch of P floats, P = R rounded up to 64 (256 bytes),
// so each row starts on a 256-byte boundary.  Cell (i, j) is at p[i * P + j].
// All arithmetic is float32 with FP contraction off and IEEE division, in the
// evaluation order of src/fluid_solver_cpu.cpp; scalar constants (a, 1 + 4a,
// dt0, h) are computed once on the host exactly as the reference computes them.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include <cstdint>

namespace fs2d {

// Boundary rule of a field (src/fluid_solver_cpu.cpp:33-83), as the signs of
// its top/bottom rows (tb) and left/right columns (lr) relative to the
// adjacent interior cell:
//   continuous  tb = +1, lr = +1   (density, pressure)
//   opposite_h  tb = +1, lr = -1   (u)
//   opposite_v  tb = -1, lr = +1   (v)
// Corners are 0.5f * (tb(x) + lr(x)) of the diagonal interior cell x.
enum Bnd : int { BND_CONT = 0, BND_OPP_H = 1, BND_OPP_V = 2 };

struct Ring {
    bool neg_tb, neg_lr;
};

__host__ __device__ inline Ring ring_of(int mode) {
    return Ring{mode == BND_OPP_V, mode == BND_OPP_H};
}

// The linear relaxation a sweep applies (src/fluid_solver_cpu.cpp:107-108, :199-200).
enum SweepKind : int { SWEEP_DIFFUSE = 0, SWEEP_PRESSURE = 1 };

struct SweepArgs {
    float* x;        // iterate (in place for Gauss-Seidel)
    const float* b;  // right-hand side: x0 for diffuse, divergence for pressure
    int R, P;
    int kind;        // SweepKind
    float a, denom;  // diffuse: a and 1 + 4a
    float inv_denom; // RN(1 / denom), for the FMA-corrected division
    Ring ring;
    int K;           // total sweeps of this solve
    int zero_start;  // iterate starts at 0 everywhere, ring included (pressure, :186)
    float div_lo;    // smallest nonzero |dividend| the FMA-corrected division takes
    // Pressure with du, dv set: the right-hand side is the divergence of
    // (du, dv), neg_half_h * (((uE - uW) + vS) - vN) (src/fluid_solver_cpu.cpp:188-193),
    // computed where the solve loads it instead of read from b.
    const float* du;
    const float* dv;
    float neg_half_h;
    float inv_denom_lo;  // RN(1 / denom - inv_denom)
    int div_checked;     // the two-product quotient is exact for this divisor (two_product_quotient)
    float inv_hi2, inv_lo2;  // its pair: inv_hi2 <= 1 / denom, inv_lo2 = RN(1 / denom - inv_hi2) > 0
};

// Lower end of the dividends the FMA-corrected division handles (zero aside):
// |x| >= 2^-100 keeps its remainder exactly representable and |x| >= d 2^-124
// keeps the quotient normal; above 2^100 the denominator's reciprocal nears
// the subnormal range, and every dividend takes the IEEE division.
inline float division_floor(float d) {
    if (!(d <= 0x1p100f)) return __builtin_inff();
    return std::max(0x1p-100f, d * 0x1p-124f);
}

inline SweepArgs sweep_args(float* x, const float* b, int R, int P, int kind, float a, float denom, int mode, int K,
                            int zero_start) {
    SweepArgs A{x, b, R, P, kind, a, denom, 0.f, ring_of(mode), K, zero_start, __builtin_inff(), nullptr, nullptr, 0.f};
    if (kind == SWEEP_DIFFUSE) {
        A.inv_denom = 1.0f / denom;
        A.div_lo = division_floor(denom);
        // 1 - d y is exact in double (d y has at most 48 significant bits)
        A.inv_denom_lo = static_cast<float>((1.0 - static_cast<double>(denom) * A.inv_denom) / denom);
    }
    return A;
}

// Whether q = RN(x y2 + RN(x y2_lo)) equals RN(x / d) for every finite
// dividend from the division floor up, with y2 the largest float <= 1/d and
// y2_lo = RN(1/d - y2) > 0; if so, sets A.div_checked, A.inv_hi2 and
// A.inv_lo2, and raises A.div_lo to where RN(x y2_lo) stays normal.  The tail
// being positive, +-inf dividends come out +-inf as well.  Decided once per
// divisor by trying every significand of one binade (the quotients of the
// other binades are exact power-of-two scalings of those) and cached.  Host
// only; a diffuse solve's first call with a new divisor costs about 4-18 ms.
void two_product_quotient(SweepArgs& A);

// Up to three independent solves with the same sweep count, relaxed by the
// same level launches (e.g. the density, u and v diffuses of one step).
struct SweepSet {
    static constexpr int kMax = 4;
    SweepArgs a[kMax];
    int count;
    int units_per_field;  // set per launch
    int level, gi, gj, gk;  // launches without a work list: the level and the unit grid (GI, GJ, GK)
};

// ---- launchers (stream-ordered, no synchronisation) ---------------------
hipError_t launch_add_sources(hipStream_t s, float* x, float* x0, const float* src, int R, int P, float dt);
// out = x + dt * src and x0 = out on the interior, out's ring = x's ring; x is
// left untouched (the density scatter still needs the pre-step velocity).
hipError_t launch_add_sources_split(hipStream_t s, float* out, float* x0, const float* x, const float* src, int R,
                                    int P, float dt);
// The density's add_sources and u's and v's add_sources_split in one launch.
hipError_t launch_add_sources3(hipStream_t s, float* rho, float* x0, const float* rs, float* un, float* x0u,
                               const float* u, const float* us, float* vn, float* x0v, const float* v,
                               const float* vs, int R, int P, float dt);
hipError_t launch_max_speed(hipStream_t s, unsigned* out_bits, const float* u, const float* v, int R, int P);
// Density advection (fs2d_scatter.hip) into out (interior and ring); max_bits
// holds max |u|, |v|.  sort_scratch (scatter_sort_bytes(R) bytes, may be null
// when every source is known to stay within the window path's radius) serves
// the sorted path taken when some source moves farther.
hipError_t launch_scatter_density(hipStream_t s, float* out, const float* rho0, const float* u, const float* v, int R,
                                  int P, float dt0, const unsigned* max_bits, void* sort_scratch);
size_t scatter_sort_bytes(int R);
int scatter_launches(int R);  // kernels launch_scatter_density enqueues with sort_scratch

// The density scatter of a batch of simulations (fs2d_batch.hip) on the
// kernels of fs2d_scatter.hip.  Simulation s's fields start `off` floats into
// the arenas; its interior cells are numbered from `cells` and its target
// cells ((R - 1)^2 of them) from `targets` across the batch.
struct ScatterSim {
    size_t off;
    int R, P;
    float dt0;
    unsigned cells, targets;
};
constexpr int kBatchStrip = 56, kBatchBand = 32;  // window-path strip columns and band rows in a batch
struct BatchScatter {
    const ScatterSim* sims;  // device, `count` entries
    int count;
    const int2* windows;     // device: (simulation, band << 16 | strip), one per wave of the window path
    int window_waves;
    long long cells, targets;  // totals over the batch
    void* scratch;           // batch_scatter_bytes(cells, targets)
    int* flag;               // device word, set every step: some simulation takes the sorted path
};
size_t batch_scatter_bytes(long long cells, long long targets);
int batch_scatter_launches(long long targets);
// max_bits[s]: the largest finite |u|, |v| of simulation s.
hipError_t launch_batch_scatter(hipStream_t s, const BatchScatter& B, float* out, const float* rho0, const float* u,
                                const float* v, const unsigned* max_bits);
hipError_t launch_gather_velocity(hipStream_t s, float* u1, float* v1, const float* u0, const float* v0,
                                  int R, int P, float dt0);
// p_zero (may be null): also zero this pressure iterate, ring included.
hipError_t launch_divergence(hipStream_t s, float* d, const float* u, const float* v, int R, int P,
                             float neg_half_h, float* p_zero);
// sums[0] += sum ||d - L p||^2 and sums[1] += sum d^2 over the interior (FP64),
// L the compact 5-point operator 4 p - (pN + pS + pW + pE) of the pressure
// sweeps (src/fluid_solver_cpu.cpp:196-204) with p's stored ring.
hipError_t launch_pressure_residual(hipStream_t s, const float* d, const float* p, int R, int P, double* sums);

// Max-speed reduction slots: kSpeedSlots words, kSlotStride words apart
// (separate cache lines), zero between uses.
constexpr int kSpeedSlots = 256, kSlotStride = 32;
// max_bits (may be null): set it to the max finite |u|, |v| of the result,
// through speed_slots (kSpeedSlots * kSlotStride words, all zero on entry,
// left zero).
hipError_t launch_gradient(hipStream_t s, float* u, float* v, const float* p, int R, int P, float h,
                           int p_is_zero, unsigned* max_bits, unsigned* speed_slots);
hipError_t launch_set_boundary(hipStream_t s, float* x, int R, int P, Ring ring);
hipError_t launch_copy4(hipStream_t s, float4* dst, const float4* src, size_t n4);
hipError_t launch_fill(hipStream_t s, void* p, uint32_t value, size_t n_words);

// Visual output (fs2d_render.hip): density -> RGBA8 (R * R * 4 bytes), velocity
// -> line vertices (R * R lines of 40 bytes), both unpitched, row-major.
hipError_t launch_render_density(hipStream_t s, const float* d, int R, int P, float mr, float mg, float mb,
                                 void* rgba);
hipError_t launch_render_velocity(hipStream_t s, const float* u, const float* v, int R, int P, float hs, float vs,
                                  void* lines);

// Multigrid pieces for the pressure equation with the continuous ring
// (fs2d_mg.hip).  smooth: one weighted Jacobi sweep src -> dst.  restrict: the
// coarse right-hand side of each coarse interior cell = the sum of the
// residuals b - (4p - sum nbrs) of its 2 x 2 fine cells (fine side = 2 x coarse
// side).  prolong: pf += bilinear interpolation of pc (pc's ring included).
// Each writes the ring of what it writes.
hipError_t launch_mg_smooth(hipStream_t s, const float* src, float* dst, const float* b, int R, int P, float omega);
hipError_t launch_mg_restrict(hipStream_t s, const float* p, const float* b, int Pf, float* bc, int Rc, int Pc);
hipError_t launch_mg_prolong(hipStream_t s, const float* pc, int Pc, float* pf, int Rf, int Pf);

// Gauss-Seidel tile wavefront: one launch per level of the tile hyperplane.
struct GsPlan {
    int TW;                // tile columns
    int M;                 // sweeps per unit (0: one sweep per tile, k_gs_tile)
    int lag;               // steps between consecutive sweeps' waves (barrier interval)
    int GI, GJ;            // tile grid
    int K;
    int levels;
    int2* work;            // device: (I << 16 | J, k) per tile-sweep, grouped by level
    int* host_offsets;     // levels + 1 offsets into work
    int* level_of;         // host: I' + J' + kappa of each (non-empty) level
    int2* host_work;       // host copy of work
};
hipError_t gs_plan_create(GsPlan* plan, int R, int K, int TW);
hipError_t gs_unit_plan_create(GsPlan* plan, int R, int K, int TW, int M, int lag);
// The same plan on the host only (no device work list) when upload is false.
hipError_t gs_unit_plan_build(GsPlan* plan, int R, int K, int TW, int M, int lag, bool upload);
void gs_plan_destroy(GsPlan* plan);
hipError_t launch_gs_solve(hipStream_t s, const GsPlan& plan, const SweepArgs& a);
// The same solve with the lagged unit kernel (M sweeps per unit, lag steps)
// but no plan: each level's units are enumerated by the kernel from its block
// index, so nothing is allocated or copied (K <= kImplicitMaxSweeps).
constexpr int kImplicitMaxSweeps = 4096;
hipError_t launch_gs_solve_implicit(hipStream_t s, int R, int K, int TW, int M, int lag, const SweepArgs& a);
// All solves of the set in shared level launches (unit plans; a tile plan takes one).
hipError_t launch_gs_solve_set(hipStream_t s, const GsPlan& plan, const SweepSet& set);
// Two solves pipelined: the diffuse fields of `diffuse` (plan pd) and the
// pressure solve p (plan pp, its right-hand side the divergence of the
// diffused velocity, p.du / p.dv) in the same level launches, pressure level
// l running with diffuse level l + offset.  offset = gs_pipeline_offset(...):
// every pressure unit then runs after the diffuse units that write the
// velocity its divergence reads.  Both plans are unit plans of one policy.
hipError_t launch_gs_pipelined(hipStream_t s, const GsPlan& pd, const SweepSet& diffuse, const GsPlan& pp,
                               const SweepArgs& p, int offset);
// The smallest such offset for a pressure plan behind a diffuse solve of Kd
// sweeps with tile columns TWd and Md sweeps per unit.
int gs_pipeline_offset(const GsPlan& pp, int R, int Kd, int TWd, int Md);

// Jacobi sweeps, `block` sweeps fused per launch, ping-ponging x -> scratch ->
// x -> ...; after an odd number of launches the result is in scratch.
// kernel: JACOBI_WAVE (register strips, one wave per tile, DPP lane exchange)
// or JACOBI_LDS (256-thread blocks, halo exchange in LDS); identical results.
enum JacobiKernel : int { JACOBI_WAVE = 1, JACOBI_LDS = 2 };
hipError_t launch_jacobi_solve(hipStream_t s, const SweepArgs& a, float* scratch, int block, int kernel);
// One launch of m fused Jacobi sweeps from src into dst (m in 1..8).
hipError_t launch_jacobi_pass(hipStream_t s, const SweepArgs& a, const float* src, float* dst, int zero_src, int m,
                              int kernel);
int jacobi_launches(int K, int block);

}  // namespace fs2d

#include <string>
namespace fs2d {
// The message fs2d_hip_last_error(NULL) returns (calls that take no handle).
void set_handle_free_error(const std::string& msg);
}  // namespace fs2d

struct fs2d_net;
namespace fs2d {
// The pressure network's prediction from pitched fields d and prev into p
// (fs2d_cnn.hip), enqueued on `stream`.
int net_predict_fields(::fs2d_net* net, hipStream_t stream, const float* d, const float* prev, int R, int P, float* p);
const char* net_error(const ::fs2d_net* net);
}  // namespace fs2d
