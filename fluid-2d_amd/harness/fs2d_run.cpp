// fs2d_run -- headless driver of the stable-fluids step through the
// reference's solver interface.
//
// It stands where src/simulation.cpp and src/app.cpp stand in the reference:
// it owns the six grids (src/simulation.cpp:9-14), holds the solver as a
// fluid_solver& and calls solve() once per step (src/simulation.cpp:53-59).
// Sources are re-applied every step, as the app's input handler does while a
// source is held (src/app.cpp:100-118); the values are passed as given, without
// the UI's W*H scaling (src/simulation.cpp:45,49-50).
//
//   fs2d_run --n 128 --steps 200 --dt 0.1 --diff 0 --visc 0 --iters 20
//            --sources disc --order gs --out DIR [--resident]
//
// Writes DIR/density.f32, DIR/u.f32, DIR/v.f32 (raw little-endian float32,
// (n+2) x (n+2), row-major, ring included) and prints one timing line.
#include "fs2d/fluid_solver_hip.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

namespace {

struct options {
    int n = 128, steps = 200, iters = 20, order = FS2D_ORDER_GS_LEX;
    float dt = 0.1f, diff = 0.f, visc = 0.f;
    std::string sources = "disc", out;
    bool resident = false;
};

void usage() {
    std::puts("usage: fs2d_run [--n N] [--steps S] [--dt DT] [--diff D] [--visc V] [--iters K]\n"
              "                [--sources disc|none] [--order gs|jacobi] [--resident] --out DIR");
}

bool parse(int argc, char** argv, options& o) {
    for (int a = 1; a < argc; ++a) {
        std::string k = argv[a];
        auto val = [&]() -> const char* { return a + 1 < argc ? argv[++a] : ""; };
        if (k == "--n") o.n = std::atoi(val());
        else if (k == "--steps") o.steps = std::atoi(val());
        else if (k == "--iters") o.iters = std::atoi(val());
        else if (k == "--dt") o.dt = std::strtof(val(), nullptr);
        else if (k == "--diff") o.diff = std::strtof(val(), nullptr);
        else if (k == "--visc") o.visc = std::strtof(val(), nullptr);
        else if (k == "--sources") o.sources = val();
        else if (k == "--order") o.order = std::string(val()) == "jacobi" ? FS2D_ORDER_JACOBI : FS2D_ORDER_GS_LEX;
        else if (k == "--resident") o.resident = true;
        else if (k == "--out") o.out = val();
        else return false;
    }
    return !o.out.empty() && o.n >= 1 && o.steps >= 0;
}

// Density 100 and upward velocity 5 (negative v: row 0 is the top) on the
// interior cells within 8 cells of the grid centre.
void disc_sources(grid<float>& rho_src, grid<float>& v_src) {
    const std::size_t R = rho_src.rows();
    const double c = (static_cast<double>(R) - 1.0) / 2.0;
    for (std::size_t i = 1; i + 1 < R; ++i)
        for (std::size_t j = 1; j + 1 < R; ++j) {
            const double di = static_cast<double>(i) - c, dj = static_cast<double>(j) - c;
            if (di * di + dj * dj <= 64.0) {
                rho_src(i, j) = 100.f;
                v_src(i, j) = -5.f;
            }
        }
}

bool write(const std::string& path, const grid<float>& g) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const std::size_t n = g.rows() * g.cols();
    const bool ok = std::fwrite(g.data(), sizeof(float), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char** argv) {
    options o;
    if (!parse(argc, argv, o)) {
        usage();
        return 2;
    }
    const std::size_t R = static_cast<std::size_t>(o.n) + 2;
    grid<float> density(R, R, 0.f), u(R, R, 0.f), v(R, R, 0.f);
    grid<float> density_src(R, R, 0.f), u_src(R, R, 0.f), v_src(R, R, 0.f);
    if (o.sources == "disc") disc_sources(density_src, v_src);

    fs2d_params prm = fluid_solver_hip::default_params();
    prm.diffuse_iterations = o.iters;
    prm.pressure_iterations = o.iters;
    prm.sweep_order = o.order;
    try {
        auto hip = std::make_unique<fluid_solver_hip>(R, R, prm);
        fluid_solver& solver = *hip;  // the reference's interface
        const auto t0 = std::chrono::steady_clock::now();
        if (o.resident) {
            hip->upload(density, u, v);
            hip->set_sources(&density_src, &u_src, &v_src);
            hip->step_n(o.steps, o.diff, o.visc, o.dt);
            hip->download(density, u, v);
        } else {
            for (int s = 0; s < o.steps; ++s)
                solver.solve(density, density_src, o.diff, u, v, u_src, v_src, o.visc, o.dt);
        }
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("fs2d_run: %d steps of %zux%zu (%s, %s) in %.3f s (%.1f steps/s)\n", o.steps, R, R,
                    o.order == FS2D_ORDER_GS_LEX ? "gauss-seidel" : "jacobi", o.resident ? "resident" : "solve()",
                    sec, sec > 0 ? o.steps / sec : 0.0);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "fs2d_run: %s\n", e.what());
        return 1;
    }
    if (!write(o.out + "/density.f32", density) || !write(o.out + "/u.f32", u) || !write(o.out + "/v.f32", v)) {
        std::fprintf(stderr, "fs2d_run: cannot write to %s\n", o.out.c_str());
        return 1;
    }
    return 0;
}
