/*
 * fs2d_oracle.h -- CPU oracle for the stable-fluids step (TEST INFRASTRUCTURE ONLY).
 *
 * This is the checker the HIP solver is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product path (libfs2d_hip.so) never links or calls it.
 *
 * It restates src/fluid_solver_cpu.cpp of mworchel/fluid-2d in plain C with the
 * same per-operation evaluation order, so that compiled with
 * -ffp-contract=off on x86-64 (SSE, no x87) it reproduces the reference bit
 * for bit.  Two extensions that the reference does not have:
 *   - iteration counts are parameters (the reference hard-codes 20);
 *   - a Jacobi sweep order (FS2D_ORDER_JACOBI) beside the reference's in-place
 *     lexicographic Gauss-Seidel order (FS2D_ORDER_GS_LEX).
 * plus FP64 twins of the divergence / projection for norm checks.
 *
 * Layout: every field is R x R float, row-major, index i*R + j, the one-cell
 * boundary ring included (src/grid.hpp:35-44).  R = n + 2.
 */
#ifndef FS2D_ORACLE_H
#define FS2D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { FSO_BND_CONTINUOUS = 0, FSO_BND_OPPOSITE_H = 1, FSO_BND_OPPOSITE_V = 2 };
enum { FSO_ORDER_GS_LEX = 0, FSO_ORDER_JACOBI = 1 };

/* src/fluid_solver_cpu.cpp:33-83 */
void fso_set_boundary(float* x, int R, int mode);
/* src/fluid_solver_cpu.cpp:85-93 */
void fso_add_sources(float* x, const float* s, int R, float dt);
/* src/fluid_solver_cpu.cpp:95-114 (order selects GS-lex or Jacobi) */
void fso_diffuse(float* x, int R, int mode, float rate, float dt, int iters, int order);
/* src/fluid_solver_cpu.cpp:116-177; trace=1 is the density scatter, 0 the gather */
void fso_advect(float* x, const float* u, const float* v, int R, int mode, float dt, int trace);
/* src/fluid_solver_cpu.cpp:179-215; p_out (may be NULL) receives the final pressure */
void fso_project(float* u, float* v, int R, int iters, int order, float* p_out);
/* src/fluid_solver_cpu.cpp:6-31 with K_d / K_p as parameters */
void fso_solve(int R, float* rho, const float* rho_src, float diff,
               float* u, float* v, const float* u_src, const float* v_src,
               float visc, float dt, int kd, int kp, int order);

/* FP64 twins ------------------------------------------------------------- */
/* ||div||_2 over the interior with the reference's wide stencil
 * (src/fluid_solver_cpu.cpp:188-193), evaluated in double. */
double fso_divergence_norm_f64(const float* u, const float* v, int R);
/* The projection of src/fluid_solver_cpu.cpp:179-215 carried out in double
 * (same order, same sweep count); u, v are updated in place in double. */
void fso_project_f64(double* u, double* v, int R, int iters, int order);
/* Relative compact Poisson residual ||d - A p||_2 / ||d||_2 in double, with the
 * continuous ring of p, for a given pressure p and divergence d. */
double fso_poisson_residual_f64(const float* d, const float* p, int R);
/* Divergence field as the reference computes it (float, interior only). */
void fso_divergence(float* d, const float* u, const float* v, int R);

/* Visual output (the renderers of the reference's GPU build) -------------- */
/* src/density_grid_renderer.cu:10-29: every cell of the rows x cols grid to an
 * RGBA8 pixel, channel c = (uint8) max(0, min(255, m_c * d)) with C99
 * fminf/fmaxf (a NaN product clamps to 255), alpha 255. */
void fso_render_density(const float* grid, int rows, int cols, float mr, float mg, float mb, unsigned char* rgba);

/* One SFML vertex: position, RGBA color, texture coordinates (20 bytes). */
typedef struct fso_vertex {
    float x, y;
    unsigned char r, g, b, a;
    float tx, ty;
} fso_vertex;

/* src/velocity_grid_renderer.cu:8-44: every cell (i, j) to a line (start, end);
 * start = end = (j * hs, i * vs); where i % 8 == 0 and j % 8 == 0 the end moves
 * by 250000 * (u, v) / sqrtf(rows * cols); colors white, texture coordinates 0
 * (the reference clears the vertex buffer first). */
void fso_render_velocity(const float* u, const float* v, int rows, int cols, float hs, float vs, fso_vertex* lines);

#ifdef __cplusplus
}
#endif
#endif
