// fs2d/fluid_solver.hpp -- the solver interface of the reference
// (src/fluid_solver.hpp:8-25), unchanged: one stable-fluids step on the
// caller's grids.  density, u and v are updated in place (ring included);
// the three sources are read-only.
#pragma once

#include "grid.hpp"

class fluid_solver {
public:
    virtual ~fluid_solver() = default;

    virtual void solve(grid<float>& density_grid,
                       grid<float> const& density_source_grid,
                       float const diffusion_rate,
                       grid<float>& horizontal_velocity_grid,
                       grid<float>& vertical_velocity_grid,
                       grid<float> const& horizontal_velocity_source_grid,
                       grid<float> const& vertical_velocity_source_grid,
                       float const viscosity,
                       float const dt) = 0;
};
