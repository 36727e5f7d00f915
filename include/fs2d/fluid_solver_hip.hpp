// fs2d/fluid_solver_hip.hpp -- the MI355X solver behind the reference's
// fluid_solver interface (src/fluid_solver.hpp:16-24), a drop-in beside
// fluid_solver_cpu and fluid_solver_gpu (src/fluid_solver_gpu.cuh:12-71).
//
// Header-only; link with libfs2d_hip.so.  Construction mirrors
// fluid_solver_gpu(rows, cols).  Where the reference ignores CUDA errors,
// this adapter throws std::runtime_error with the library's message.
//
// Beyond the interface, the resident API keeps the state in HBM across steps:
//   upload(density, u, v); set_sources(...); step_n(n, ...); download(...)
#pragma once

#include "../fs2d_hip.h"
#include "fluid_solver.hpp"

#include <stdexcept>
#include <string>

class fluid_solver_hip : public fluid_solver {
public:
    // Reference defaults: 20 diffuse and 20 pressure sweeps, Gauss-Seidel order
    // (results identical to fluid_solver_cpu).
    fluid_solver_hip(std::size_t rows, std::size_t cols) : fluid_solver_hip(rows, cols, default_params()) {}

    fluid_solver_hip(std::size_t rows, std::size_t cols, fs2d_params const& params) {
        check(fs2d_hip_create(static_cast<int>(rows), static_cast<int>(cols), &params, &m_handle), nullptr,
              "fs2d_hip_create");
        m_rows = rows;
    }

    ~fluid_solver_hip() override { fs2d_hip_destroy(m_handle); }

    fluid_solver_hip(fluid_solver_hip const&) = delete;
    fluid_solver_hip& operator=(fluid_solver_hip const&) = delete;

    void solve(grid<float>& density_grid,
               grid<float> const& density_source_grid,
               float const diffusion_rate,
               grid<float>& horizontal_velocity_grid,
               grid<float>& vertical_velocity_grid,
               grid<float> const& horizontal_velocity_source_grid,
               grid<float> const& vertical_velocity_source_grid,
               float const viscosity,
               float const dt) override {
        grid<float> const* grids[] = {&density_grid, &density_source_grid, &horizontal_velocity_grid,
                                      &vertical_velocity_grid, &horizontal_velocity_source_grid,
                                      &vertical_velocity_source_grid};
        for (grid<float> const* g : grids)
            if (g->rows() != m_rows || g->cols() != m_rows)
                throw std::invalid_argument("fluid_solver_hip::solve: grid shape differs from the solver's");
        check(fs2d_hip_solve(m_handle, density_grid.data(), density_source_grid.data(), diffusion_rate,
                             horizontal_velocity_grid.data(), vertical_velocity_grid.data(),
                             horizontal_velocity_source_grid.data(), vertical_velocity_source_grid.data(), viscosity,
                             dt),
              m_handle, "fs2d_hip_solve");
    }

    // ---- resident API -------------------------------------------------------
    void upload(grid<float> const& density, grid<float> const& u, grid<float> const& v) {
        check(fs2d_hip_upload(m_handle, density.data(), u.data(), v.data()), m_handle, "fs2d_hip_upload");
    }
    // nullptr = zero source
    void set_sources(grid<float> const* density_source, grid<float> const* u_source, grid<float> const* v_source) {
        check(fs2d_hip_set_sources(m_handle, density_source ? density_source->data() : nullptr,
                                   u_source ? u_source->data() : nullptr, v_source ? v_source->data() : nullptr),
              m_handle, "fs2d_hip_set_sources");
    }
    void step_n(int n, float diffusion_rate, float viscosity, float dt) {
        check(fs2d_hip_step_n(m_handle, n, diffusion_rate, viscosity, dt), m_handle, "fs2d_hip_step_n");
    }
    void download(grid<float>& density, grid<float>& u, grid<float>& v) {
        check(fs2d_hip_download(m_handle, density.data(), u.data(), v.data()), m_handle, "fs2d_hip_download");
    }

    fs2d_handle* handle() { return m_handle; }

    static fs2d_params default_params() {
        fs2d_params p;
        fs2d_default_params(&p);
        return p;
    }

private:
    static void check(int rc, fs2d_handle const* h, char const* what) {
        if (rc != FS2D_OK) throw std::runtime_error(std::string(what) + ": " + fs2d_hip_last_error(h));
    }

    fs2d_handle* m_handle = nullptr;
    std::size_t m_rows = 0;
};
