// ref_shim.cpp -- C entry points into the reference's own CPU solver
// (TEST INFRASTRUCTURE ONLY).
//
// oracle/Makefile compiles this file together with
// src/fluid_solver_cpu.cpp of mworchel/fluid-2d, read in place from the
// reference checkout, into oracle/_ref/libfs2d_ref.so.  Nothing of the
// reference is copied into this repository.  The stage functions of
// fluid_solver_cpu are private members; access specifiers are not part of the
// Itanium mangled names, so re-opening them here calls the reference's own
// compiled code unchanged.
//
// Fields are R x R float, row-major, ring included (src/grid.hpp:35-44).
#define private public
#include "fluid_solver_cpu.hpp"
#undef private

#include <cstring>

namespace {

grid<float> load(const float* x, int R) {
    grid<float> g(static_cast<size_t>(R), static_cast<size_t>(R), 0.f);
    std::memcpy(g.data(), x, sizeof(float) * static_cast<size_t>(R) * R);
    return g;
}

void store(float* x, const grid<float>& g, int R) {
    std::memcpy(x, g.data(), sizeof(float) * static_cast<size_t>(R) * R);
}

using bnd_fn = void (*)(grid<float>&);

bnd_fn boundary(int mode) {
    switch (mode) {
        case 1: return &fluid_solver_cpu::set_boundary_opposite_horizontal;
        case 2: return &fluid_solver_cpu::set_boundary_opposite_vertical;
        default: return &fluid_solver_cpu::set_boundary_continuous;
    }
}

}  // namespace

extern "C" {

// src/fluid_solver.hpp:16-24 (K_d = K_p = 20 are literals in the reference)
void ref_solve(int R, float* rho, const float* rho_src, float diff, float* u, float* v,
               const float* u_src, const float* v_src, float visc, float dt) {
    grid<float> g_rho = load(rho, R), g_u = load(u, R), g_v = load(v, R);
    grid<float> g_rs = load(rho_src, R), g_us = load(u_src, R), g_vs = load(v_src, R);
    fluid_solver_cpu solver;
    solver.solve(g_rho, g_rs, diff, g_u, g_v, g_us, g_vs, visc, dt);
    store(rho, g_rho, R);
    store(u, g_u, R);
    store(v, g_v, R);
}

// src/fluid_solver_cpu.cpp:33-83
void ref_set_boundary(float* x, int R, int mode) {
    grid<float> g = load(x, R);
    boundary(mode)(g);
    store(x, g, R);
}

// src/fluid_solver_cpu.cpp:85-93
void ref_add_sources(float* x, const float* s, int R, float dt) {
    grid<float> g = load(x, R), gs = load(s, R);
    fluid_solver_cpu solver;
    solver.add_sources(g, gs, dt);
    store(x, g, R);
}

// src/fluid_solver_cpu.cpp:95-114
void ref_diffuse(float* x, int R, int mode, float rate, float dt, int iters) {
    grid<float> g = load(x, R);
    fluid_solver_cpu solver;
    solver.diffuse(g, boundary(mode), rate, dt, static_cast<size_t>(iters));
    store(x, g, R);
}

// src/fluid_solver_cpu.cpp:116-177
void ref_advect(float* x, const float* u, const float* v, int R, int mode, float dt, int trace) {
    grid<float> g = load(x, R), gu = load(u, R), gv = load(v, R);
    fluid_solver_cpu solver;
    solver.advect(g, gu, gv, boundary(mode), dt, trace != 0);
    store(x, g, R);
}

// src/fluid_solver_cpu.cpp:179-215
void ref_project(float* u, float* v, int R, int iters) {
    grid<float> gu = load(u, R), gv = load(v, R);
    fluid_solver_cpu solver;
    solver.project(gu, gv, static_cast<size_t>(iters));
    store(u, gu, R);
    store(v, gv, R);
}

}  // extern "C"
