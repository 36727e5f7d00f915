"""Batched simulations (fs2d_hip_batch_*): every simulation of a batch must
come out bit-identical to the oracle's Jacobi order run on it alone, whatever
depth and tile the autotuner picks, eagerly or through replayed graphs.
"""
from __future__ import annotations

import numpy as np
import pytest

import fs2d_amd
import oracle
from conftest import bits_equal, describe
from fs2d_amd import workloads as W

pytestmark = pytest.mark.gpu

# (rows, iterations, dt, diffusion, viscosity): uneven sizes, ring-only and
# one-tile grids, rates from 0 to strong diffusion
SIMS = [(3, 8, 0.1, 1e-3, 1e-4), (10, 16, 0.05, 5e-2, 1e-3), (34, 8, 0.07, 1e-4, 0.0), (67, 24, 0.1, 0.0, 1e-5),
        (70, 16, 0.03, 1e-2, 1e-2), (130, 8, 0.1, 1e-4, 1e-5), (258, 16, 0.05, 1e-5, 1e-6)]


def start(k, R):
    if R < 20:
        rho, u, v, rs, us, vs = W.random_state(R, seed=40 + k, vel_scale=0.3)
        return (rho, u, v), (rs, us, vs)
    st = W.batch_state(R, seed=3, sim_index=k)
    src = W.disc_sources(R, radius=R / 12) if k % 2 else (None, None, None)
    return st, src


def run_and_check(sims, block, mode, steps, oracle_lib, starts=None, occupancy=None):
    starts = starts or [start(k, s[0]) for k, s in enumerate(sims)]
    want = [[x.copy() for x in st] for st, _ in starts]
    z = [np.zeros((s[0], s[0]), np.float32) for s in sims]
    with fs2d_amd.Batch(sims, jacobi_block=block, launch_mode=mode) as b:
        if occupancy is not None:
            b.set_occupancy(*occupancy)
        for k, (st, src) in enumerate(starts):
            b.upload(k, *st)
            b.set_sources(k, *src)
        done = 0
        for target in steps:
            b.step_n(target - done)
            for k, (R, K, dt, diff, visc) in enumerate(sims):
                src = [z[k] if x is None else x for x in starts[k][1]]
                for _ in range(target - done):
                    oracle_lib.solve(want[k][0], src[0], diff, want[k][1], want[k][2], src[1], src[2], visc, dt, K, K,
                                     oracle.JACOBI)
                got = b.download(k)
                for f, name in enumerate(("density", "u", "v")):
                    assert bits_equal(got[f], want[k][f]), \
                        f"sim {k} (R={R} K={K}) {name} after {target} steps, block={block}: " \
                        f"{describe(got[f], want[k][f])}"
            done = target
        info = b.info()
    return info


@pytest.mark.parametrize("block", [0, 1, 2, 4])
@pytest.mark.parametrize("mode", [fs2d_amd.LAUNCH_GRAPH, fs2d_amd.LAUNCH_EAGER])
def test_batch_matches_oracle(block, mode, oracle_lib):
    info = run_and_check(SIMS, block, mode, (1, 2, 5), oracle_lib)
    if block:
        assert info["jacobi_depth"] == block
    else:
        assert info["jacobi_depth"] in (1, 2, 4)  # 8 does not divide 8 at least twice
    assert info["tile_rows"] in (16, 32)
    assert info["jacobi_blocks_per_cu"] in (0, 1, 2, 4, 8) and info["stage_blocks_per_cu"] in (0, 2, 4, 8)
    assert info["compute_units"] > 0
    assert info["interior_cells"] == sum((R - 2) ** 2 for R, *_ in SIMS)
    if mode == fs2d_amd.LAUNCH_GRAPH:
        assert info["graphs"] >= 1
    else:
        assert info["graphs"] == 0


def test_batch_sample_of_the_batched_workload(oracle_lib):
    """Eight simulations drawn like the 256-simulation workload (sides 64..512,
    dt in [0.02, 0.1], 20 or 40 sweeps), two steps each."""
    cfg = W.batch_config(8, seed=11)
    sims = [(R, K, dt, 1e-5, 1e-6) for R, K, dt in cfg]
    info = run_and_check(sims, 0, fs2d_amd.LAUNCH_GRAPH, (1, 2), oracle_lib)
    assert info["jacobi_depth"] in (1, 2, 4)


def test_batch_single_simulation_equals_handle():
    """A batch of one equals the one-grid handle in the Jacobi order."""
    R, K = 130, 16
    st = W.batch_state(R, seed=5, sim_index=0)
    rs, us, vs = W.disc_sources(R, radius=10)
    with fs2d_amd.Batch([(R, K, 0.05, 1e-4, 1e-5)], jacobi_block=4) as b, \
            fs2d_amd.Solver(R, diffuse_iterations=K, pressure_iterations=K, order=fs2d_amd.JACOBI,
                            jacobi_block=4) as s:
        b.upload(0, *st)
        b.set_sources(0, rs, us, vs)
        s.upload(*st)
        s.set_sources(rs, us, vs)
        b.step_n(7)
        s.step_n(7, 1e-4, 1e-5, 0.05)
        gb, gs = b.download(0), s.download()
    for f in range(3):
        assert bits_equal(gb[f], gs[f]), describe(gb[f], gs[f])


def scaled_start(R, dt, reach, seed):
    """A random state whose largest displacement dt0 * max|vel| is `reach` cells."""
    rho, u, v, *_ = W.random_state(R, seed=seed, vel_scale=1.0)
    dt0 = np.sqrt(np.float32(R * R)) * np.float32(dt)
    scale = np.float32(reach / dt0 / max(np.abs(u).max(), np.abs(v).max()))
    return (rho, (u * scale).astype(np.float32), (v * scale).astype(np.float32)), (None, None, None)


def fast_cell_start(R, seed, fast):
    rho, u, v, *_ = W.random_state(R, seed=seed, vel_scale=0.002)
    u, v = u.copy(), v.copy()
    u[R // 3, R // 2], v[R // 3, R // 2] = fast
    return (rho, u, v), (None, None, None)


@pytest.mark.parametrize("mode", [fs2d_amd.LAUNCH_GRAPH, fs2d_amd.LAUNCH_EAGER])
def test_batch_scatter_radii_and_fast_cells(mode, oracle_lib):
    """The batched density scatter: simulations whose sources reach 0.7 to 3.7
    cells (every radius of the window path) beside simulations with one fast
    cell (u = 5 and 1e6: the batched sorted path), three steps, every
    simulation bit for bit against the oracle."""
    sims = [(66, 8, 0.1, 1e-4, 1e-5), (70, 8, 0.1, 0.0, 0.0), (130, 8, 0.05, 1e-4, 0.0), (98, 8, 0.1, 0.0, 1e-5),
            (258, 8, 0.1, 1e-5, 1e-6), (34, 8, 0.1, 0.0, 0.0), (130, 16, 0.1, 1e-5, 1e-6)]
    starts = [scaled_start(66, 0.1, 0.7, 1), scaled_start(70, 0.1, 1.5, 2), scaled_start(130, 0.05, 2.5, 3),
              scaled_start(98, 0.1, 3.7, 4), fast_cell_start(258, 5, (5.0, 0.0)), fast_cell_start(34, 6, (1e6, -2.0)),
              scaled_start(130, 0.1, 0.3, 7)]
    run_and_check(sims, 0, mode, (1, 3), oracle_lib, starts=starts)


def test_batch_launch_count_includes_the_scatter():
    sims = [(66, 8, 0.1, 1e-4, 1e-5), (130, 16, 0.05, 1e-5, 1e-6)]
    with fs2d_amd.Batch(sims, jacobi_block=4) as b:
        for k, (R, *_rest) in enumerate(sims):
            b.upload(k, *W.batch_state(R, seed=1, sim_index=k))
            b.set_sources(k)
        b.step_n(1)
        info = b.info()
    targets = sum((R - 1) ** 2 for R, *_ in sims)
    passes = -(-targets.bit_length() // 9)
    # add, 4 diffuse launches (16 sweeps / 4), the scatter (window, gate, keys, histograms, scans, scatters,
    # bounds, merge),
    # 2 x (divergence, 4 pressure launches, gradient, fold), gather
    assert info["launches_per_step"] == 1 + 4 + (4 + 3 * passes) + 2 * (3 + 4) + 1, info


@pytest.mark.parametrize("occupancy", [(1, 1), (2, 8), (0, 0)])
def test_batch_persistent_grids(occupancy, oracle_lib):
    """Fixed persistent grids: one block per CU (every wave and block walks
    many entries), a middle setting, and one block per entry -- the same
    bits as the oracle, and the setting is kept."""
    info = run_and_check(SIMS, 2, fs2d_amd.LAUNCH_GRAPH, (1, 3), oracle_lib, occupancy=occupancy)
    assert (info["jacobi_blocks_per_cu"], info["stage_blocks_per_cu"]) == occupancy
