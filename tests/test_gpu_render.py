"""Visual output (include/fs2d_render.h) on the GPU against the oracle.

The HIP conversions of the resident density (RGBA8) and velocity (line
vertices) must equal the oracle's restatement of the reference kernels
(src/density_grid_renderer.cu:10-29, src/velocity_grid_renderer.cu:8-44) byte
for byte, on random fields with negative, huge, NaN and infinite values, and on
the state a few solver steps leave.
"""
from __future__ import annotations

import numpy as np
import pytest

import fs2d_amd
import oracle
from fs2d_amd import workloads as W

pytestmark = pytest.mark.gpu


def special_field(R, rng, scale):
    x = (rng.standard_normal((R, R)) * scale).astype(np.float32)
    flat = x.reshape(-1)
    k = max(1, flat.size // 50)
    idx = rng.choice(flat.size, size=4 * k, replace=False)
    flat[idx[:k]] = np.nan
    flat[idx[k:2 * k]] = np.inf
    flat[idx[2 * k:3 * k]] = -np.inf
    flat[idx[3 * k:]] = -0.0
    return x


@pytest.mark.parametrize("R", [3, 10, 67, 130, 258, 1026])
def test_render_matches_oracle(R, oracle_lib):
    rng = np.random.default_rng(R)
    rho = special_field(R, rng, 2.0)
    u = special_field(R, rng, 1e-3)
    v = special_field(R, rng, 1e-3)
    with fs2d_amd.Solver(R) as s:
        s.upload(rho, u, v)
        for m in ((255.0, 128.0, 1.0), (100.0, 0.0, -3.0), (1e30, 1.0, 0.5)):
            got = s.render_density(*m)
            want = oracle_lib.render_density(rho, *m)
            assert np.array_equal(got, want), f"density R={R} m={m}: {np.argwhere(got != want)[:4]}"
        for hs, vs in ((1.0, 1.0), (800.0 / R, 600.0 / R)):
            got = s.render_velocity(hs, vs)
            want = oracle_lib.render_velocity(u, v, hs, vs)
            assert got.tobytes() == want.tobytes(), f"lines R={R} scale=({hs}, {vs})"


def test_render_after_steps(oracle_lib):
    """The conversion reads the state the enqueued steps leave (stream order)."""
    R = 130
    st = W.vortex_state(R, seed=3, n_vortices=16)
    rs, us, vs = W.disc_sources(R, radius=R / 16)
    with fs2d_amd.Solver(R) as s:
        s.upload(*st)
        s.set_sources(rs, us, vs)
        s.step_n(4, 1e-4, 1e-5, 0.05)
        px = s.render_density(255.0, 200.0, 150.0)
        lines = s.render_velocity(2.0, 2.0)
        rho, u, v = s.download()
    assert np.array_equal(px, oracle_lib.render_density(rho, 255.0, 200.0, 150.0))
    assert lines.tobytes() == oracle_lib.render_velocity(u, v, 2.0, 2.0).tobytes()


def test_render_timing_is_measured():
    with fs2d_amd.Solver(258) as s:
        s.upload(*W.vortex_state(258, seed=1, n_vortices=8))
        a, b = s.time_render(3)
    assert 0 < a < 100 and 0 < b < 100
