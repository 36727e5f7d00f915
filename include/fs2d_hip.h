/*
 * fs2d_hip.h -- C ABI of the MI355X stable-fluids solver (libfs2d_hip.so).
 *
 * Replaces the step of mworchel/fluid-2d behind its solver interface:
 *   fluid_solver::solve          src/fluid_solver.hpp:16-24
 *   fluid_solver_cpu::solve      src/fluid_solver_cpu.cpp:6-31   (the arithmetic it reproduces)
 *   fluid_solver_gpu::solve      src/fluid_solver_gpu.cu:222-258 (the CUDA solver it sits beside)
 *
 * Plain C: no C++ or HIP types cross this boundary.  Every function returns an
 * int status (FS2D_OK = 0); on failure fs2d_hip_last_error() describes it.
 *
 * Fields are rows x cols float32, row-major, index i*cols + j, including the
 * one-cell boundary ring (src/grid.hpp:35-44).  Only square grids are
 * accepted: the reference's grid::cols() returns rows (src/grid.hpp:20-22), so
 * its solver is defined only for rows == cols.  rows >= 3.
 *
 * Threading: a handle owns one HIP stream and all of its device memory; use a
 * handle from one host thread at a time (fluid_solver_gpu shares its
 * temporaries between calls in the same way).
 */
#ifndef FS2D_HIP_H
#define FS2D_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define FS2D_ABI_VERSION 3

enum fs2d_status {
    FS2D_OK = 0,
    FS2D_ERR_ARGUMENT = 1, /* invalid argument (sizes, NULL pointers, parameters) */
    FS2D_ERR_HIP = 2,      /* a HIP runtime call failed (allocation, copy, launch) */
    FS2D_ERR_STATE = 3     /* call not valid in the handle's current state */
};

/* Order of the diffuse and pressure relaxation sweeps. */
enum fs2d_sweep_order {
    /* In-place lexicographic Gauss-Seidel, the reference CPU solver's order
     * (src/fluid_solver_cpu.cpp:104-113, :196-204); the iterates, and hence
     * the step, equal the reference's bit for bit.  Default. */
    FS2D_ORDER_GS_LEX = 0,
    /* Jacobi (every neighbour from the previous iterate), temporally blocked
     * (fs2d_jacobi_kernel).  A different relaxation than the reference's: it
     * agrees with the reference only where the solve has converged. */
    FS2D_ORDER_JACOBI = 1
};

/* Kernel that runs the Jacobi sweeps (identical results). */
enum fs2d_jacobi_kernel {
    FS2D_JACOBI_WAVE = 1, /* one wave per tile, rows in registers, lanes exchange by DPP (default) */
    FS2D_JACOBI_LDS = 2   /* 256-thread blocks, halo exchange through LDS                         */
};

/* How fs2d_hip_step_n and fs2d_hip_solve issue a step's launches. */
enum fs2d_launch_mode {
    FS2D_LAUNCH_GRAPH = 0, /* each step state captured once as a hipGraph, then replayed (default) */
    FS2D_LAUNCH_EAGER = 1  /* every launch issued one by one                                     */
};

/* Boundary rule of a field (src/fluid_solver_cpu.cpp:33-83). */
enum fs2d_field {
    FS2D_FIELD_DENSITY = 0,    /* continuous ring                    */
    FS2D_FIELD_VELOCITY_U = 1, /* left/right edges negated (u)       */
    FS2D_FIELD_VELOCITY_V = 2  /* top/bottom edges negated (v)       */
};

typedef struct fs2d_params {
    int diffuse_iterations;  /* K_d; the reference's literal is 20 (src/fluid_solver_cpu.cpp:17,23-24) */
    int pressure_iterations; /* K_p; the reference's literal is 20 (src/fluid_solver_cpu.cpp:25,30)    */
    int sweep_order;         /* enum fs2d_sweep_order                                                   */
    int jacobi_block;        /* Jacobi sweeps fused per launch (temporal block depth); 0 = autotuned    */
    int gs_tile_cols;        /* Gauss-Seidel tile width in cells: 64 or 128 (128 takes gs_block 1, 4, 8) */
    int gs_block;            /* Gauss-Seidel sweeps per tile (temporal block): 1, 2, 4 or 8              */
    int gs_lag;              /* steps each sweep trails the previous one in a tile: 1, 4 or 8
                              * Each of these three left at 0 is chosen per solve kind and sweep count
                              * on the first solve, by timing the candidates (64/128 columns, 4/8
                              * sweeps, lag 4/8) on scratch fields; cached per process.  All choices
                              * give identical results.                                             */
    int jacobi_kernel;       /* enum fs2d_jacobi_kernel; 0 = default (register strips)                   */
    int launch_mode;         /* enum fs2d_launch_mode; 0 = graph replay                                  */
} fs2d_params;

typedef struct fs2d_handle fs2d_handle;

/* Reference defaults: K_d = K_p = 20, Gauss-Seidel order. */
void fs2d_default_params(fs2d_params* params);

/* Allocate a solver for rows x cols grids on the current HIP device.
 * Mirrors fluid_solver_gpu::fluid_solver_gpu(rows, cols)
 * (src/fluid_solver_gpu.cuh:14, src/fluid_solver_gpu.cu:209-218).
 * params may be NULL (reference defaults). */
int fs2d_hip_create(int rows, int cols, const fs2d_params* params, fs2d_handle** out);
int fs2d_hip_destroy(fs2d_handle* h);

/* One step with host arrays, synchronous: the semantics of
 * fluid_solver::solve (src/fluid_solver.hpp:16-24).  density, u and v are
 * read and updated in place, ring included; the three sources are read-only
 * and only their interior is used (src/fluid_solver_cpu.cpp:88-90). */
int fs2d_hip_solve(fs2d_handle* h, float* density, const float* density_source,
                   float diffusion_rate, float* u, float* v, const float* u_source,
                   const float* v_source, float viscosity, float dt);

/* ---- resident API: state stays in HBM between steps -------------------- */
/* Copy density, u, v (host, rows*cols floats each) into the handle. */
int fs2d_hip_upload(fs2d_handle* h, const float* density, const float* u, const float* v);
/* Set the sources applied by every following step; a NULL pointer means an
 * all-zero source.  (simulation::update clears its sources after each call,
 * src/simulation.cpp:62-64: set NULLs to reproduce that.) */
int fs2d_hip_set_sources(fs2d_handle* h, const float* density_source, const float* u_source,
                         const float* v_source);
/* Enqueue n steps on the handle's stream; returns without waiting. */
int fs2d_hip_step_n(fs2d_handle* h, int n, float diffusion_rate, float viscosity, float dt);
/* Wait for the handle's stream, then copy the state out (any pointer may be NULL). */
int fs2d_hip_download(fs2d_handle* h, float* density, float* u, float* v);
int fs2d_hip_sync(fs2d_handle* h);
/* Number of step graphs the handle holds (captured, instantiated, replayable). */
int fs2d_hip_graph_count(fs2d_handle* h, int* count);

/* ---- device-pointer variants (no host staging) --------------------------
 * Arrays are rows*cols floats on the handle's device, row-major, unpadded.
 * The copies are enqueued on the handle's stream. */
int fs2d_hip_upload_device(fs2d_handle* h, const float* density, const float* u, const float* v);
int fs2d_hip_download_device(fs2d_handle* h, float* density, float* u, float* v);

/* ---- the reference's stages, on the resident state ----------------------
 * Each runs one private stage of fluid_solver_cpu on the handle's fields, with
 * the same boundary rule the step uses for that field.  For tests and
 * tooling; fs2d_hip_step_n is the fused path. */
/* src/fluid_solver_cpu.cpp:85-93: field += dt * source (interior). */
int fs2d_hip_stage_add_sources(fs2d_handle* h, int field, float dt);
/* src/fluid_solver_cpu.cpp:95-114 with `iterations` sweeps. */
int fs2d_hip_stage_diffuse(fs2d_handle* h, int field, float rate, float dt, int iterations);
/* src/fluid_solver_cpu.cpp:127-152: density scatter along (u, v). */
int fs2d_hip_stage_advect_density(fs2d_handle* h, float dt);
/* src/fluid_solver_cpu.cpp:26-29,153-174: u and v gathered along a copy of (u, v). */
int fs2d_hip_stage_advect_velocity(fs2d_handle* h, float dt);
/* src/fluid_solver_cpu.cpp:179-215 with `iterations` pressure sweeps. */
int fs2d_hip_stage_project(fs2d_handle* h, int iterations);
/* src/fluid_solver_cpu.cpp:33-83 on one field. */
int fs2d_hip_stage_set_boundary(fs2d_handle* h, int field);
/* Copy the last projection's pressure (rows*cols floats, ring included) to host. */
int fs2d_hip_download_pressure(fs2d_handle* h, float* p);

/* ---- measurement ------------------------------------------------------- */
/* The handle's HIP stream (a hipStream_t), for callers that record events. */
void* fs2d_hip_stream(fs2d_handle* h);
/* Enqueue only the relaxation sweeps of one pressure solve (K_p sweeps on the
 * current velocity's divergence, stored first; as a step's second projection
 * runs it, the zero fill included)
 * `reps` times and return the mean time of one solve in milliseconds,
 * measured with HIP events on the handle's stream.  The state's velocity is
 * not modified. */
int fs2d_hip_time_pressure_sweeps(fs2d_handle* h, int reps, float* ms_per_solve);
/* Enqueue the three diffuse solves of one step (density, u, v: K_d sweeps
 * each, shared level launches in the Gauss-Seidel order as the step runs
 * them) `reps` times on copies of the state, and return the mean time of one
 * step's diffuses in milliseconds, measured with HIP events on the handle's
 * stream.  The state (density, velocity, pressure) is not modified. */
int fs2d_hip_time_diffuse_sweeps(fs2d_handle* h, float diff, float visc, float dt, int reps, float* ms_per_solve);
/* The Gauss-Seidel step's joint level launches -- the three diffuses and the
 * first projection's pressure solve pipelined behind them (divergence of the
 * diffused velocity computed on load) -- `reps` times on copies of the state:
 * mean milliseconds per run (HIP events, the pressure iterate's zero fill
 * included).  Gauss-Seidel units only. */
int fs2d_hip_time_joint_sweeps(fs2d_handle* h, float diff, float visc, float dt, int reps, float* ms_per_solve);
/* The Gauss-Seidel launch policy of the diffuse (kind 0, K_d sweeps) or
 * pressure (kind 1, K_p sweeps) solves: tile columns, sweeps per unit (1: the
 * one-sweep tile kernel), lag, and level launches per solve.  Chooses it first
 * if no solve of that kind has run (see gs_tile_cols, gs_block, gs_lag).
 * Kind 2: the step's joint launches (the diffuses with the first projection's
 * pressure solve pipelined behind them, on the diffuse's policy). */
int fs2d_hip_gs_policy(fs2d_handle* h, int kind, int* tile_cols, int* sweeps_per_unit, int* lag, int* levels);
/* Host only (no device call): the levels by which a Gauss-Seidel step's first
 * pressure solve trails its diffuses when both run in the same level launches
 * (tile_cols x 64 tiles, sweeps_per_unit sweeps per unit, both solves one
 * policy).  Every pressure unit then runs after the diffuse units that write
 * the velocity its divergence reads. */
int fs2d_gs_pipeline_offset(int rows, int diffuse_iterations, int pressure_iterations, int tile_cols,
                            int sweeps_per_unit, int* offset);
/* Host only (no device call): the division a diffuse solve by denom = 1 + 4a
 * uses (src/fluid_solver_cpu.cpp:107-108 divides by it in IEEE float).
 * inv = RN(1 / denom), inv_lo = RN(1 / denom - inv): the pair of the
 * two-product quotient with one Markstein correction, correctly rounded for
 * every divisor.  two_hi <= 1 / denom and two_lo = RN(1 / denom - two_hi) > 0:
 * the pair of the two-product quotient alone.  *checked = 1 when that
 * quotient, RN(x two_hi + RN(x two_lo)), was found correctly rounded for every
 * finite dividend from lo up (the Gauss-Seidel unit kernels then use it; an
 * infinite dividend comes out infinite too), else 0.  lo = the smallest
 * nonzero |dividend| the FMA-based quotients take (smaller ones, zero and NaN,
 * and for the corrected quotient those above 2^100 and inf, take the IEEE
 * division).  The first call for a divisor tries 2^23 dividends (about 4-18
 * ms); the answer is cached. */
int fs2d_division_plan(float denom, float* inv, float* inv_lo, float* two_hi, float* two_lo, float* lo,
                       int* checked);
/* On-device copy bandwidth: a float4 copy kernel over `bytes` (read + write
 * counted), timed with HIP events over `reps` copies; the achievable HBM
 * figure beside the 8 TB/s peak. */
int fs2d_hip_copy_bandwidth(long long bytes, int reps, float* gb_per_s);
/* Number of kernel launches one step enqueues (for per-launch averages). */
int fs2d_hip_launches_per_step(fs2d_handle* h, int* launches);

/* ---- projection without allocation -------------------------------------
 * No counterpart in the reference; for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  The projection of fs2d_hip_stage_project
 * (src/fluid_solver_cpu.cpp:179-215: divergence, params->pressure_iterations
 * sweeps from a zero pressure, gradient subtraction, opposite rings) on
 * caller-owned device arrays u and v (rows x rows floats each, row-major,
 * ring included, updated in place), enqueued on `stream` (a hipStream_t; NULL
 * = the null stream), returning without waiting.  pressure (may be NULL)
 * receives the final pressure, rows x rows floats.  Results equal
 * fs2d_hip_stage_project's bit for bit.
 *
 * Allocates nothing: its scratch (divergence, pressure, the pressure's Jacobi
 * partner) is one file-scope device array sized for rows <= 1026 (12.6 MB per
 * GPU, held until the process ends); larger grids are rejected.  Calls that
 * overlap on one GPU share that array and race: serialise them.  Orders: the
 * Jacobi order (jacobi_block 0 = 4 sweeps per launch, not tuned) or
 * Gauss-Seidel in units of gs_block 2, 4 or 8 sweeps (pressure_iterations <=
 * 4096).  Errors: fs2d_hip_last_error(NULL). */
int fs2d_hip_project_noalloc(int rows, const fs2d_params* params, float* u, float* v, float* pressure,
                             void* stream);

/* Last error of the handle (or, when h is NULL, of the last failed create or
 * handle-free call). */
const char* fs2d_hip_last_error(const fs2d_handle* h);
int fs2d_hip_abi_version(void);

/* ---- batched simulations -------------------------------------------------
 * Many independent simulations of different sizes advanced together: every
 * stage of the step is one launch over a tile work list spanning all of
 * them, and the step is replayed as a hipGraph.  No counterpart in the
 * reference (one fluid_solver per grid, src/simulation.cpp:17-26); each
 * simulation's results equal those of an fs2d_handle in the Jacobi order
 * with the same parameters, bit for bit.  Simulations are indexed 0..count-1
 * in the order given. */
typedef struct fs2d_batch_sim {
    int rows;             /* R = n + 2 (square grid, ring included), 3 .. 8194            */
    int iterations;       /* Jacobi sweeps of each diffuse and pressure solve, >= 2      */
    float dt;
    float diffusion_rate; /* density diffusion                                            */
    float viscosity;      /* velocity diffusion                                           */
} fs2d_batch_sim;

typedef struct fs2d_batch fs2d_batch;

/* params may be NULL; only sweep_order (must be FS2D_ORDER_JACOBI),
 * jacobi_block (0 = autotuned; otherwise it must divide every simulation's
 * iterations at least twice) and launch_mode are read.  Fields start at zero. */
int fs2d_hip_batch_create(int count, const fs2d_batch_sim* sims, const fs2d_params* params, fs2d_batch** out);
int fs2d_hip_batch_destroy(fs2d_batch* b);
/* Host arrays of rows*rows floats for simulation `sim`. */
int fs2d_hip_batch_upload(fs2d_batch* b, int sim, const float* density, const float* u, const float* v);
/* NULL = all-zero source. */
int fs2d_hip_batch_set_sources(fs2d_batch* b, int sim, const float* density_source, const float* u_source,
                               const float* v_source);
/* Enqueue n steps of every simulation; returns without waiting. */
int fs2d_hip_batch_step_n(fs2d_batch* b, int n);
/* Wait for the batch's stream, then copy one simulation out (any pointer may be NULL). */
int fs2d_hip_batch_download(fs2d_batch* b, int sim, float* density, float* u, float* v);
int fs2d_hip_batch_sync(fs2d_batch* b);
/* The configuration in use: Jacobi sweeps per launch, rows per Jacobi tile,
 * kernel launches per step, captured step graphs, total interior cells. */
int fs2d_hip_batch_info(fs2d_batch* b, int* jacobi_depth, int* tile_rows, int* launches_per_step, int* graphs,
                        long long* interior_cells);
/* The persistent launches' grid, as the autotuner chose it on the first step:
 * blocks (of 4 waves) per compute unit for the Jacobi launches and for the
 * single-pass stages (0: one block per work entry), and the compute units. */
int fs2d_hip_batch_occupancy(fs2d_batch* b, int* jacobi_blocks_per_cu, int* stage_blocks_per_cu,
                             int* compute_units);
/* Fix those two figures instead of tuning them (before the first step only;
 * 0..64, 0 = one block per work entry).  Results do not depend on them. */
int fs2d_hip_batch_set_occupancy(fs2d_batch* b, int jacobi_blocks_per_cu, int stage_blocks_per_cu);
const char* fs2d_hip_batch_last_error(const fs2d_batch* b);

#ifdef __cplusplus
}
#endif
#endif
