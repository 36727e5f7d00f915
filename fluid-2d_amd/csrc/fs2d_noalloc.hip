// fs2d_noalloc.hip -- the projection of fluid_solver_cpu
// (src/fluid_solver_cpu.cpp:179-215) on caller-owned device arrays, for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).  It runs the kernels of
// fs2d_hip_stage_project -- divergence, the relaxation sweeps, the gradient
// with the opposite rings -- but its scratch (the divergence, the pressure
// iterate and the pressure's Jacobi partner) lives in one file-scope device
// array instead of a handle's allocations, and the Gauss-Seidel levels take
// their units from the block index instead of a device work list.
//
// The array is sized for grids of up to kNoAllocMaxRows rows (ring included);
// every GPU the library runs on holds its own copy until the process ends.
// Every call on one GPU shares that GPU's copy: calls that overlap on a GPU
// (concurrent streams, host threads) race and must be serialised by the
// caller.  Nothing is allocated: no hipMalloc, hipFree or allocating library
// routine; the pressure iterate is cleared by a kernel on every call.
#include "../../include/fs2d_hip.h"
#include "fs2d_internal.hpp"

#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

using namespace fs2d;

namespace {

constexpr int kNoAllocMaxRows = 1026;  // 1024 x 1024 interior plus the ring
constexpr size_t kNoAllocField = static_cast<size_t>(kNoAllocMaxRows) * kNoAllocMaxRows;

// Divergence, pressure and the pressure's Jacobi partner, rows x rows floats
// each (row pitch = rows), packed from the start of the array.
static __device__ __align__(256) float scratch[3 * kNoAllocField];

constexpr int kMaxDevices = 64;
float* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each device
std::mutex g_scratch_mu;

// Errors are reported like fs2d_hip_create's: fs2d_hip_last_error(NULL).
int fail_na(int code, const std::string& msg) {
    set_handle_free_error(msg);
    return code;
}

// The array's device address on the current device, looked up once per device
// (the array's name is a host-side stand-in and never a device address).
int scratch_of_current_device(float** out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail_na(FS2D_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= kMaxDevices) return fail_na(FS2D_ERR_STATE, "fs2d_hip_project_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail_na(FS2D_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<float*>(ptr);
    }
    *out = g_scratch[dev];
    return FS2D_OK;
}

}  // namespace

extern "C" int fs2d_hip_project_noalloc(int rows, const fs2d_params* params, float* u, float* v, float* pressure,
                                        void* stream) {
    if (rows < 3 || rows > kNoAllocMaxRows)
        return fail_na(FS2D_ERR_ARGUMENT, "fs2d_hip_project_noalloc: rows must be in [3, " +
                                              std::to_string(kNoAllocMaxRows) + "] (the size of its scratch array)");
    if (!u || !v) return fail_na(FS2D_ERR_ARGUMENT, "fs2d_hip_project_noalloc: NULL u or v");
    fs2d_params prm;
    fs2d_default_params(&prm);
    if (params) prm = *params;
    const int K = prm.pressure_iterations;
    if (K < 0 || K > kImplicitMaxSweeps)
        return fail_na(FS2D_ERR_ARGUMENT, "fs2d_hip_project_noalloc: pressure_iterations must be in [0, " +
                                              std::to_string(kImplicitMaxSweeps) + "]");
    const bool gs = prm.sweep_order == FS2D_ORDER_GS_LEX;
    if (!gs && prm.sweep_order != FS2D_ORDER_JACOBI)
        return fail_na(FS2D_ERR_ARGUMENT, "fs2d_hip_project_noalloc: unknown sweep_order");
    const int gsb = prm.gs_block ? prm.gs_block : 8, lag = prm.gs_lag ? prm.gs_lag : 4;
    const int tw = prm.gs_tile_cols ? prm.gs_tile_cols : 64;
    if (gs && ((gsb != 2 && gsb != 4 && gsb != 8) || (lag != 1 && lag != 4 && lag != 8) ||
               (tw != 64 && tw != 128) || (tw == 128 && gsb == 2)))
        return fail_na(FS2D_ERR_ARGUMENT,
                       "fs2d_hip_project_noalloc: Gauss-Seidel takes gs_block 2, 4 or 8, gs_lag 1, 4 or 8 and "
                       "gs_tile_cols 64 or 128 (128 with gs_block 4 or 8)");
    const int depth = prm.jacobi_block ? prm.jacobi_block : 4;
    if (!gs && (depth < 1 || depth > 8))
        return fail_na(FS2D_ERR_ARGUMENT, "fs2d_hip_project_noalloc: jacobi_block must be in [0, 8]");
    const int jk = prm.jacobi_kernel ? prm.jacobi_kernel : JACOBI_WAVE;
    if (jk != JACOBI_WAVE && jk != JACOBI_LDS)
        return fail_na(FS2D_ERR_ARGUMENT, "fs2d_hip_project_noalloc: jacobi_kernel must be 0, 1 or 2");

    float* base = nullptr;
    int rc = scratch_of_current_device(&base);
    if (rc) return rc;
    const int R = rows, P = rows;
    const size_t field = static_cast<size_t>(R) * R;
    float* d = base;
    float* p = base + field;
    float* t = base + 2 * field;
    hipStream_t s = static_cast<hipStream_t>(stream);
    auto hip = [&](hipError_t e, const char* what) {
        return e == hipSuccess ? FS2D_OK
                               : fail_na(FS2D_ERR_HIP, std::string("fs2d_hip_project_noalloc: ") + what + ": " +
                                                           hipGetErrorString(e));
    };

    // Scalars as src/fluid_solver_cpu.cpp:183,190 computes them.
    const float cells = static_cast<float>(static_cast<size_t>(R) * static_cast<size_t>(R));
    const float h = 1.0f / std::sqrt(cells);
    const float neg_half_h = -0.5f * h;
    // divergence, and the pressure iterate cleared (ring included, :186)
    if ((rc = hip(launch_divergence(s, d, u, v, R, P, neg_half_h, K > 0 ? p : nullptr), "divergence"))) return rc;
    const float* pf = p;
    if (K > 0) {
        const SweepArgs A = sweep_args(p, d, R, P, SWEEP_PRESSURE, 0.f, 0.f, BND_CONT, K, 0);
        if (gs) {
            if ((rc = hip(launch_gs_solve_implicit(s, R, K, tw, gsb, lag, A), "Gauss-Seidel sweeps"))) return rc;
        } else {
            if ((rc = hip(launch_jacobi_solve(s, A, t, depth, jk), "Jacobi sweeps"))) return rc;
            if (jacobi_launches(K, depth) & 1) pf = t;
        }
    }
    if ((rc = hip(launch_gradient(s, u, v, pf, R, P, h, K <= 0 ? 1 : 0, nullptr, nullptr), "gradient"))) return rc;
    if (pressure) {
        if (K > 0)
            rc = hip(hipMemcpyAsync(pressure, pf, sizeof(float) * field, hipMemcpyDeviceToDevice, s), "pressure copy");
        else
            rc = hip(launch_fill(s, pressure, 0u, field), "pressure fill");
        if (rc) return rc;
    }
    return hip(hipGetLastError(), "launch");
}
