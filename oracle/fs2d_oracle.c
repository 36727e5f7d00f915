/*
 * fs2d_oracle.c -- CPU oracle for the stable-fluids step (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this file's library, and only as the checker.  The HIP product path never
 * links it and has no CPU fallback.
 *
 * A restatement of src/fluid_solver_cpu.cpp (mworchel/fluid-2d) in C.  Every
 * expression keeps the reference's evaluation order; build with
 * -ffp-contract=off and without fast-math (oracle/Makefile does) so that x86-64
 * SSE arithmetic rounds exactly like the reference build.  Pins: the compiled
 * reference (oracle/_ref, bit-for-bit), committed golden fixtures, closed forms
 * and invariants -- see tests/test_oracle.py.
 */
#include "fs2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define AT(g, i, j) (g)[(size_t)(i) * (size_t)R + (size_t)(j)]

/* std::min / std::max of <algorithm>: min(a,b) = (b < a) ? b : a,
 * max(a,b) = (a < b) ? b : a.  The NaN behaviour of the gather clamp follows
 * from exactly these forms (src/fluid_solver_cpu.cpp:158-159). */
static float std_min(float a, float b) { return (b < a) ? b : a; }
static float std_max(float a, float b) { return (a < b) ? b : a; }

static float* clone(const float* x, int R) {
    size_t n = (size_t)R * (size_t)R;
    float* c = (float*)malloc(n * sizeof(float));
    memcpy(c, x, n * sizeof(float));
    return c;
}

/* ---- boundary ring: src/fluid_solver_cpu.cpp:33-83 ----------------------
 * The loops are kept in the reference's order (columns then rows for the
 * continuous and horizontal modes, rows then columns for the vertical mode);
 * the corners are 0.5f * (edge_a + edge_b) of the freshly written edges. */
static void corners(float* g, int R) {
    AT(g, 0, 0) = 0.5f * (AT(g, 0, 1) + AT(g, 1, 0));
    AT(g, 0, R - 1) = 0.5f * (AT(g, 0, R - 2) + AT(g, 1, R - 1));
    AT(g, R - 1, 0) = 0.5f * (AT(g, R - 1, 1) + AT(g, R - 2, 0));
    AT(g, R - 1, R - 1) = 0.5f * (AT(g, R - 1, R - 2) + AT(g, R - 2, R - 1));
}

void fso_set_boundary(float* g, int R, int mode) {
    if (mode == FSO_BND_OPPOSITE_V) {
        /* src/fluid_solver_cpu.cpp:67-83: top/bottom negated, then left/right copied */
        for (int j = 0; j < R; ++j) {
            AT(g, 0, j) = -AT(g, 1, j);
            AT(g, R - 1, j) = -AT(g, R - 2, j);
        }
        for (int i = 0; i < R; ++i) {
            AT(g, i, 0) = AT(g, i, 1);
            AT(g, i, R - 1) = AT(g, i, R - 2);
        }
    } else {
        /* :33-48 (continuous) and :50-65 (u: left/right negated) */
        const int neg = (mode == FSO_BND_OPPOSITE_H);
        for (int i = 0; i < R; ++i) {
            AT(g, i, 0) = neg ? -AT(g, i, 1) : AT(g, i, 1);
            AT(g, i, R - 1) = neg ? -AT(g, i, R - 2) : AT(g, i, R - 2);
        }
        for (int j = 0; j < R; ++j) {
            AT(g, 0, j) = AT(g, 1, j);
            AT(g, R - 1, j) = AT(g, R - 2, j);
        }
    }
    corners(g, R);
}

/* ---- src/fluid_solver_cpu.cpp:85-93 ------------------------------------- */
void fso_add_sources(float* x, const float* s, int R, float dt) {
    for (int i = 1; i < R - 1; ++i)
        for (int j = 1; j < R - 1; ++j)
            AT(x, i, j) += dt * AT(s, i, j);
}

/* ---- src/fluid_solver_cpu.cpp:95-114 ------------------------------------
 * a = (dt * float(R*R)) * rate; each sweep writes
 *   (x0 + a * (((N + S) + W) + E)) / (1 + 4a)
 * GS-lex reads N and W from the current sweep (in place); the Jacobi order
 * reads every neighbour, ring included, from the previous iterate. */
void fso_diffuse(float* x, int R, int mode, float rate, float dt, int iters, int order) {
    const float a = dt * (float)((size_t)R * (size_t)R) * rate;
    float* x0 = clone(x, R);
    float* prev = (order == FSO_ORDER_JACOBI) ? clone(x, R) : NULL;
    for (int k = 0; k < iters; ++k) {
        const float* src = x;
        if (prev) {
            memcpy(prev, x, (size_t)R * (size_t)R * sizeof(float));
            src = prev;
        }
        for (int i = 1; i < R - 1; ++i)
            for (int j = 1; j < R - 1; ++j)
                AT(x, i, j) = (AT(x0, i, j) + a * (AT(src, i - 1, j) + AT(src, i + 1, j) +
                                                    AT(src, i, j - 1) + AT(src, i, j + 1))) /
                              (1.f + 4.f * a);
        fso_set_boundary(x, R, mode);
    }
    free(prev);
    free(x0);
}

/* ---- src/fluid_solver_cpu.cpp:116-177 ----------------------------------- */
void fso_advect(float* g, const float* u, const float* v, int R, int mode, float dt, int trace) {
    float* g0 = clone(g, R);
    const float dt0 = sqrtf((float)((size_t)R * (size_t)R)) * dt;
    const float fR = (float)R;
    if (trace) {
        /* forward scatter (:127-152): each destination accumulates its
         * contributions in source row-major order, starting from +0. */
        memset(g, 0, (size_t)R * (size_t)R * sizeof(float));
        for (int i = 1; i < R - 1; ++i) {
            for (int j = 1; j < R - 1; ++j) {
                float x = (float)j + dt0 * AT(u, i, j);
                float y = (float)i + dt0 * AT(v, i, j);
                if (x < 0.5f || x > fR - 1.5f || y < 0.5f || y > fR - 1.5f) continue;
                /* The reference converts a NaN coordinate to size_t here
                 * (undefined behaviour, an out-of-bounds write); the oracle
                 * and the HIP solver both define it as "source skipped". */
                if (x != x || y != y) continue;
                size_t j0 = (size_t)x, i0 = (size_t)y;
                size_t j1 = j0 + 1, i1 = i0 + 1;
                float s0 = x - (float)j0;
                float s1 = 1 - s0;
                float s2 = y - (float)i0;
                float s3 = 1 - s2;
                AT(g, i0, j0) += s1 * s3 * AT(g0, i, j);
                AT(g, i1, j0) += s1 * s2 * AT(g0, i, j);
                AT(g, i0, j1) += s0 * s3 * AT(g0, i, j);
                AT(g, i1, j1) += s0 * s2 * AT(g0, i, j);
            }
        }
    } else {
        /* backtrace gather with the [1.5, R-1.5] clamp (:153-174) */
        for (int i = 1; i < R - 1; ++i) {
            for (int j = 1; j < R - 1; ++j) {
                float x = (float)j - dt0 * AT(u, i, j);
                float y = (float)i - dt0 * AT(v, i, j);
                x = std_max(1.5f, std_min(fR - 1.5f, x));
                y = std_max(1.5f, std_min(fR - 1.5f, y));
                size_t j0 = (size_t)x, i0 = (size_t)y;
                size_t j1 = j0 + 1, i1 = i0 + 1;
                float s0 = x - (float)j0;
                float s1 = 1 - s0;
                float s2 = y - (float)i0;
                float s3 = 1 - s2;
                AT(g, i, j) = s3 * (s1 * AT(g0, i0, j0) + s0 * AT(g0, i0, j1)) +
                              s2 * (s1 * AT(g0, i1, j0) + s0 * AT(g0, i1, j1));
            }
        }
    }
    free(g0);
    fso_set_boundary(g, R, mode);
}

/* ---- src/fluid_solver_cpu.cpp:188-193 ----------------------------------- */
void fso_divergence(float* d, const float* u, const float* v, int R) {
    const float h = 1.0f / sqrtf((float)((size_t)R * (size_t)R));
    for (int i = 1; i < R - 1; ++i)
        for (int j = 1; j < R - 1; ++j)
            AT(d, i, j) = -0.5f * h * (AT(u, i, j + 1) - AT(u, i, j - 1) +
                                       AT(v, i + 1, j) - AT(v, i - 1, j));
}

/* ---- src/fluid_solver_cpu.cpp:179-215 ----------------------------------- */
void fso_project(float* u, float* v, int R, int iters, int order, float* p_out) {
    const float h = 1.0f / sqrtf((float)((size_t)R * (size_t)R));
    const size_t n = (size_t)R * (size_t)R;
    float* d = (float*)calloc(n, sizeof(float));
    float* p = (float*)calloc(n, sizeof(float));
    float* prev = (order == FSO_ORDER_JACOBI) ? (float*)calloc(n, sizeof(float)) : NULL;
    fso_divergence(d, u, v, R);
    fso_set_boundary(d, R, FSO_BND_CONTINUOUS); /* :194, never read afterwards */
    for (int k = 0; k < iters; ++k) {
        const float* src = p;
        if (prev) {
            memcpy(prev, p, n * sizeof(float));
            src = prev;
        }
        for (int i = 1; i < R - 1; ++i)
            for (int j = 1; j < R - 1; ++j)
                AT(p, i, j) = (AT(d, i, j) + AT(src, i, j + 1) + AT(src, i, j - 1) +
                               AT(src, i + 1, j) + AT(src, i - 1, j)) / 4.0f;
        fso_set_boundary(p, R, FSO_BND_CONTINUOUS);
    }
    for (int i = 1; i < R - 1; ++i) {
        for (int j = 1; j < R - 1; ++j) {
            AT(u, i, j) -= 0.5f * (AT(p, i, j + 1) - AT(p, i, j - 1)) / h;
            AT(v, i, j) -= 0.5f * (AT(p, i + 1, j) - AT(p, i - 1, j)) / h;
        }
    }
    fso_set_boundary(u, R, FSO_BND_OPPOSITE_H);
    fso_set_boundary(v, R, FSO_BND_OPPOSITE_V);
    if (p_out) memcpy(p_out, p, n * sizeof(float));
    free(prev);
    free(p);
    free(d);
}

/* ---- src/fluid_solver_cpu.cpp:6-31 -------------------------------------- */
void fso_solve(int R, float* rho, const float* rho_src, float diff,
               float* u, float* v, const float* u_src, const float* v_src,
               float visc, float dt, int kd, int kp, int order) {
    fso_add_sources(rho, rho_src, R, dt);
    fso_diffuse(rho, R, FSO_BND_CONTINUOUS, diff, dt, kd, order);
    fso_advect(rho, u, v, R, FSO_BND_CONTINUOUS, dt, 1);

    fso_add_sources(u, u_src, R, dt);
    fso_add_sources(v, v_src, R, dt);
    fso_diffuse(u, R, FSO_BND_OPPOSITE_H, visc, dt, kd, order);
    fso_diffuse(v, R, FSO_BND_OPPOSITE_V, visc, dt, kd, order);
    fso_project(u, v, R, kp, order, NULL);
    float* tu = clone(u, R);
    float* tv = clone(v, R);
    fso_advect(u, tu, tv, R, FSO_BND_OPPOSITE_H, dt, 0);
    fso_advect(v, tu, tv, R, FSO_BND_OPPOSITE_V, dt, 0);
    free(tu);
    free(tv);
    fso_project(u, v, R, kp, order, NULL);
}

/* ---- FP64 twins ---------------------------------------------------------- */
#define ATD(g, i, j) (g)[(size_t)(i) * (size_t)R + (size_t)(j)]

double fso_divergence_norm_f64(const float* u, const float* v, int R) {
    const double h = 1.0 / sqrt((double)R * (double)R);
    double acc = 0.0;
    for (int i = 1; i < R - 1; ++i)
        for (int j = 1; j < R - 1; ++j) {
            double dv = -0.5 * h * (((double)AT(u, i, j + 1) - (double)AT(u, i, j - 1)) +
                                    ((double)AT(v, i + 1, j) - (double)AT(v, i - 1, j)));
            acc += dv * dv;
        }
    return sqrt(acc);
}

static void set_boundary_f64(double* g, int R, int mode) {
    if (mode == FSO_BND_OPPOSITE_V) {
        for (int j = 0; j < R; ++j) { ATD(g, 0, j) = -ATD(g, 1, j); ATD(g, R - 1, j) = -ATD(g, R - 2, j); }
        for (int i = 0; i < R; ++i) { ATD(g, i, 0) = ATD(g, i, 1); ATD(g, i, R - 1) = ATD(g, i, R - 2); }
    } else {
        const int neg = (mode == FSO_BND_OPPOSITE_H);
        for (int i = 0; i < R; ++i) {
            ATD(g, i, 0) = neg ? -ATD(g, i, 1) : ATD(g, i, 1);
            ATD(g, i, R - 1) = neg ? -ATD(g, i, R - 2) : ATD(g, i, R - 2);
        }
        for (int j = 0; j < R; ++j) { ATD(g, 0, j) = ATD(g, 1, j); ATD(g, R - 1, j) = ATD(g, R - 2, j); }
    }
    ATD(g, 0, 0) = 0.5 * (ATD(g, 0, 1) + ATD(g, 1, 0));
    ATD(g, 0, R - 1) = 0.5 * (ATD(g, 0, R - 2) + ATD(g, 1, R - 1));
    ATD(g, R - 1, 0) = 0.5 * (ATD(g, R - 1, 1) + ATD(g, R - 2, 0));
    ATD(g, R - 1, R - 1) = 0.5 * (ATD(g, R - 1, R - 2) + ATD(g, R - 2, R - 1));
}

void fso_project_f64(double* u, double* v, int R, int iters, int order) {
    const double h = 1.0 / sqrt((double)R * (double)R);
    const size_t n = (size_t)R * (size_t)R;
    double* d = (double*)calloc(n, sizeof(double));
    double* p = (double*)calloc(n, sizeof(double));
    double* prev = (order == FSO_ORDER_JACOBI) ? (double*)calloc(n, sizeof(double)) : NULL;
    for (int i = 1; i < R - 1; ++i)
        for (int j = 1; j < R - 1; ++j)
            ATD(d, i, j) = -0.5 * h * (ATD(u, i, j + 1) - ATD(u, i, j - 1) +
                                       ATD(v, i + 1, j) - ATD(v, i - 1, j));
    for (int k = 0; k < iters; ++k) {
        const double* src = p;
        if (prev) { memcpy(prev, p, n * sizeof(double)); src = prev; }
        for (int i = 1; i < R - 1; ++i)
            for (int j = 1; j < R - 1; ++j)
                ATD(p, i, j) = (ATD(d, i, j) + ATD(src, i, j + 1) + ATD(src, i, j - 1) +
                                ATD(src, i + 1, j) + ATD(src, i - 1, j)) / 4.0;
        set_boundary_f64(p, R, FSO_BND_CONTINUOUS);
    }
    for (int i = 1; i < R - 1; ++i)
        for (int j = 1; j < R - 1; ++j) {
            ATD(u, i, j) -= 0.5 * (ATD(p, i, j + 1) - ATD(p, i, j - 1)) / h;
            ATD(v, i, j) -= 0.5 * (ATD(p, i + 1, j) - ATD(p, i - 1, j)) / h;
        }
    set_boundary_f64(u, R, FSO_BND_OPPOSITE_H);
    set_boundary_f64(v, R, FSO_BND_OPPOSITE_V);
    free(prev);
    free(p);
    free(d);
}

double fso_poisson_residual_f64(const float* d, const float* p, int R) {
    /* A p = 4 p - (E + W + S + N), with the continuous ring (edge = adjacent
     * interior), i.e. the Neumann operator the p sweep iterates on. */
    double num = 0.0, den = 0.0;
    for (int i = 1; i < R - 1; ++i)
        for (int j = 1; j < R - 1; ++j) {
            double c = AT(p, i, j);
            double nb[4];
            nb[0] = (i == 1) ? c : (double)AT(p, i - 1, j);
            nb[1] = (i == R - 2) ? c : (double)AT(p, i + 1, j);
            nb[2] = (j == 1) ? c : (double)AT(p, i, j - 1);
            nb[3] = (j == R - 2) ? c : (double)AT(p, i, j + 1);
            double r = (double)AT(d, i, j) - (4.0 * c - (nb[0] + nb[1] + nb[2] + nb[3]));
            num += r * r;
            den += (double)AT(d, i, j) * (double)AT(d, i, j);
        }
    return den > 0.0 ? sqrt(num / den) : sqrt(num);
}

/* ---- visual output ------------------------------------------------------ */

/* src/density_grid_renderer.cu:17-26.  The reference's device max/min on
 * floats are fmaxf/fminf; static_cast<uint8_t> truncates a value in [0, 255]. */
void fso_render_density(const float* grid, int rows, int cols, float mr, float mg, float mb, unsigned char* rgba) {
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) {
            const float d = grid[(size_t)i * cols + j];
            unsigned char* px = rgba + 4 * ((size_t)i * cols + j);
            px[0] = (unsigned char)fmaxf(0.f, fminf(255.f, mr * d));
            px[1] = (unsigned char)fmaxf(0.f, fminf(255.f, mg * d));
            px[2] = (unsigned char)fmaxf(0.f, fminf(255.f, mb * d));
            px[3] = 255;
        }
}

/* src/velocity_grid_renderer.cu:17-41: the size_t operands convert to float
 * (j and i exactly; rows * cols rounds to float above 2^24 before sqrtf, as in
 * the reference). */
void fso_render_velocity(const float* u, const float* v, int rows, int cols, float hs, float vs, fso_vertex* lines) {
    const float s = sqrtf((float)((size_t)rows * (size_t)cols));
    for (int i = 0; i < rows; ++i)
        for (int j = 0; j < cols; ++j) {
            fso_vertex* ln = lines + 2 * ((size_t)i * cols + j);
            const float x = (float)j * hs, y = (float)i * vs;
            float ex = x, ey = y;
            if (i % 8 == 0 && j % 8 == 0) {
                const float vx = u[(size_t)i * cols + j], vy = v[(size_t)i * cols + j];
                ex = ex + 250000.f * vx / s;
                ey = ey + 250000.f * vy / s;
            }
            const fso_vertex a = {x, y, 255, 255, 255, 255, 0.f, 0.f};
            const fso_vertex b = {ex, ey, 255, 255, 255, 255, 0.f, 0.f};
            ln[0] = a;
            ln[1] = b;
        }
}
