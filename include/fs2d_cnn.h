/* fs2d_cnn.h -- the learned pressure initial guess (C ABI of libfs2d_hip.so).
 *
 * No counterpart in the reference: fluid_solver_cpu::project always starts the
 * pressure iterate from zero (src/fluid_solver_cpu.cpp:186).  This is new
 * behaviour outside parity mode: an encoder-decoder CNN maps (divergence,
 * previous pressure) to a pressure field that seeds the Jacobi projection,
 * and plain Jacobi from the previous pressure is the fallback.
 *
 * Network: 4 levels of 16/32/64/128 channels, two 3x3 convolutions + ReLU per
 * level, 2x2 average-pool down, bilinear x2 up, skip connections by channel
 * concatenation, a 3x3 output convolution, and a residual on the normalised
 * previous pressure.  Convolutions are BF16 implicit GEMMs on MFMA with FP32
 * accumulation; weights are kept in FP32 (master copy, Adam moments) and used
 * in BF16.  Inputs are normalised per sample by the RMS of the divergence and
 * of the previous pressure; the output is scaled back by the latter.
 *
 * Grid sides given to the network must be multiples of 32 (after three
 * poolings a side still holds whole 16-pixel MFMA row groups).  Every call returns FS2D_OK (0) or an error code;
 * fs2d_net_last_error() describes the last failure. */
#ifndef FS2D_CNN_H
#define FS2D_CNN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fs2d_net fs2d_net;

/* A network with He-initialised weights drawn from a SplitMix64 stream of `seed`. */
int fs2d_net_create(uint64_t seed, fs2d_net** out);
int fs2d_net_destroy(fs2d_net* net);
/* Number of FP32 parameters (weights and biases of all layers). */
int fs2d_net_param_count(fs2d_net* net, int64_t* count);
/* Copy all parameters out / in (host arrays of param_count floats, in layer
 * order: for each layer its weights [9][cout][cin] then its biases [cout]).
 * set_params also resets the Adam state. */
int fs2d_net_get_params(fs2d_net* net, float* host);
int fs2d_net_set_params(fs2d_net* net, const float* host);
/* The gradients of the last training step's loss (same layout as the parameters). */
int fs2d_net_get_grads(fs2d_net* net, float* host);

/* One Adam step on a batch of n samples of h x w cells (host arrays):
 * divergence[n][h][w], prev_pressure[n][h][w] and target_pressure[n][h][w],
 * float32.  Returns the batch's mean squared error of the normalised
 * prediction in *loss (before the update).  lr: Adam step size (beta1 0.9,
 * beta2 0.999, eps 1e-8). */
int fs2d_net_train_step(fs2d_net* net, int n, int h, int w, const float* divergence, const float* prev_pressure,
                        const float* target_pressure, float lr, float* loss);

/* Prediction for n samples (host arrays as above; out[n][h][w]). */
int fs2d_net_predict_host(fs2d_net* net, int n, int h, int w, const float* divergence, const float* prev_pressure,
                          float* out);

/* Mean time of one forward pass over n samples of h x w (the activations of
 * that shape must exist: run a prediction or training step of it first),
 * measured with HIP events on the network's stream over `reps` passes. */
int fs2d_net_time_forward(fs2d_net* net, int n, int h, int w, int reps, float* ms);
/* Multiply-add FLOPs (2 per MAC, padded channels included: what the MFMA
 * units execute) of one forward pass over n samples of h x w. */
int fs2d_net_forward_flops(fs2d_net* net, int n, int h, int w, double* flops);

/* The network's stream (a hipStream_t). */
void* fs2d_net_stream(fs2d_net* net);
const char* fs2d_net_last_error(const fs2d_net* net);

/* ---- the seeded projection on an fs2d_handle ------------------------------
 * fs2d_hip_project_seeded: the projection of the handle's current velocity
 * (src/fluid_solver_cpu.cpp:179-215) with the pressure iterate seeded instead
 * of zero:
 *   seed = 0: zero (the reference's start), 1: the previous pressure (the
 *   handle's pressure field as the last projection left it), 2: the network's
 *   prediction from (divergence, previous pressure) (net required), 3:
 *   multigrid from zero -- V(2,2) cycles of weighted Jacobi (omega 0.8) on
 *   grids of side n, n/2, ... down to a side of 4 or an odd side (n_c^2
 *   sweeps on the coarsest, 64 to 4096), residuals restricted by 2 x 2 sums,
 *   corrections interpolated bilinearly, until tol or 30 cycles (net
 *   ignored).
 * Jacobi sweeps run in launches of the handle's depth until the compact
 * Poisson residual ||d - L p||_2 <= tol * ||d||_2 (checked every `check`
 * sweeps, rounded up to whole launches), max_sweeps is reached or the
 * residual is not finite (a NaN or infinite iterate stays so under the
 * sweeps); then the gradient is applied.  With seed = 2 and fallback != 0, a
 * seeded solve that misses tol -- a non-finite network guess at once, without
 * sweeps -- restarts from the previous pressure (seed 1) with the same
 * budget.  Reports the sweeps run, the final relative residual
 * and whether the fallback ran (sweeps_run counts the fine grid's sweeps,
 * V-cycle smoothing included).  Requires the Jacobi order and a grid side
 * n = rows - 2 that is a multiple of 32 for seed 2. */
struct fs2d_handle;
int fs2d_hip_project_seeded(struct fs2d_handle* h, fs2d_net* net, int seed, float tol, int max_sweeps, int check,
                            int fallback, int* sweeps_run, double* rel_residual, int* fell_back);
/* Relative compact Poisson residual ||d - L p||_2 / ||d||_2 of the handle's
 * stored divergence and pressure fields (FP64 sums): those of the last
 * fs2d_hip_stage_project, seeded projection or step (whose last projection
 * stores its divergence; a Gauss-Seidel step's first projection computes it
 * where the pressure solve loads it and stores none). */
int fs2d_hip_pressure_residual(struct fs2d_handle* h, double* rel_residual);

#ifdef __cplusplus
}
#endif
#endif
