"""CPU-side tests: the C ABI library loads and exports what include/fs2d_hip.h
declares, argument validation happens before any HIP call, the sweep
schedules of the kernels reproduce the oracle (host-side models), and the
workload generator is deterministic.  No GPU compute is issued here."""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np
import pytest

import fs2d_amd
import oracle
from conftest import ROOT, bits_equal
from fs2d_amd import workloads as W
from schedule_models import (gs_shear_lag_model, gs_shear_model, gs_shear_plan, gs_tile_model, gs_unit_lag_model, gs_unit_model,
                             gs_unit_plan, gs_unit_ring_model, jacobi_block_model, jacobi_wave_model)


def header_symbols():
    text = "".join(open(os.path.join(ROOT, "include", f)).read() for f in ("fs2d_hip.h", "fs2d_cnn.h", "fs2d_render.h"))
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fs2d_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = fs2d_amd.library()
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    assert set(syms) == set(fs2d_amd.EXPORTED_SYMBOLS)
    assert lib.fs2d_hip_abi_version() == 3
    assert ctypes.sizeof(fs2d_amd.Params) == 9 * 4


def test_default_params_are_the_reference_literals():
    lib = fs2d_amd.library()
    p = fs2d_amd.Params()
    lib.fs2d_default_params(ctypes.byref(p))
    assert (p.diffuse_iterations, p.pressure_iterations, p.sweep_order) == (20, 20, fs2d_amd.GS_LEX)


@pytest.mark.parametrize("rows,cols,kw,msg", [
    (10, 12, {}, "square"),
    (2, 2, {}, "rows"),
    (10, 10, {"diffuse_iterations": -1}, "iteration"),
    (10, 10, {"order": 7}, "sweep_order"),
    (10, 10, {"jacobi_block": 9}, "jacobi_block"),
    (10, 10, {"gs_tile_cols": 48}, "gs_tile_cols"),
    (10, 10, {"gs_tile_cols": 128, "gs_block": 2}, "gs_tile_cols 128"),
    (10, 10, {"gs_block": 3}, "gs_block"),
    (10, 10, {"gs_lag": 2}, "gs_lag"),
    (10, 10, {"jacobi_kernel": 3}, "jacobi_kernel"),
    (10, 10, {"launch_mode": 2}, "launch_mode"),
    (4098, 4098, {"pressure_iterations": 1000000}, "too many Gauss-Seidel sweeps"),
    (32768, 32768, {"diffuse_iterations": 20000, "gs_block": 1}, "too many Gauss-Seidel sweeps"),
])
def test_create_validates_arguments_before_touching_the_device(rows, cols, kw, msg):
    with pytest.raises(fs2d_amd.Fs2dError, match=msg):
        fs2d_amd.Solver(rows, cols, **kw)


@pytest.mark.parametrize("sims,kw,msg", [
    ([], {}, "count"),
    ([(2, 8, 0.1, 0.0, 0.0)], {}, "rows"),
    ([(10, 1, 0.1, 0.0, 0.0)], {}, "iterations"),
    ([(10, 8, 0.1, 0.0, 0.0)], {"order": fs2d_amd.GS_LEX}, "Jacobi"),
    ([(10, 8, 0.1, 0.0, 0.0), (12, 6, 0.1, 0.0, 0.0)], {"jacobi_block": 4}, "jacobi_block"),
    ([(10, 8, 0.1, 0.0, 0.0)], {"jacobi_block": 3}, "jacobi_block"),
    ([(10, 8, 0.1, 0.0, 0.0)], {"launch_mode": 5}, "launch_mode"),
])
def test_batch_create_validates_arguments_before_touching_the_device(sims, kw, msg):
    with pytest.raises(fs2d_amd.Fs2dError, match=msg):
        fs2d_amd.Batch(sims, **kw)


def test_batch_config_draws_the_stated_distributions():
    cfg = W.batch_config(256, seed=0)
    assert len(cfg) == 256
    sides = {R - 2 for R, _, _ in cfg}
    assert sides <= set(W.BATCH_SIDES) and len(sides) == len(W.BATCH_SIDES)
    assert {K for _, K, _ in cfg} == {20, 40}
    dts = np.array([dt for _, _, dt in cfg])
    assert dts.min() >= 0.02 and dts.max() <= 0.1 and dts.std() > 0.015
    assert cfg == W.batch_config(256, seed=0)  # seeded


def test_oracle_is_not_linked_into_the_product():
    """The product library must not depend on the oracle (no CPU fallback)."""
    so = open(fs2d_amd.LIB_PATH, "rb").read()
    assert b"fso_" not in so and b"libfs2d_oracle" not in so and b"ref_solve" not in so


# ------------------------------------------------ schedule models
@pytest.mark.parametrize("R,TW", [(3, 64), (10, 64), (67, 64), (70, 128)])
def test_gs_tile_schedule_reproduces_gauss_seidel(R, TW, oracle_lib):
    rho, u, v, *_ = W.random_state(R, seed=R)
    for mode in range(3):
        for K in (1, 3):
            a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
            den = np.float32(1) + np.float32(4) * a
            x = rho.copy()
            gs_tile_model(x, rho.copy(), K, 0, a, den, mode == 2, mode == 1, 0, TW)
            ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, K, oracle.GS_LEX)
            assert bits_equal(x, ref), (mode, K)
    d = oracle_lib.divergence(u, v)
    a, b = u.copy(), v.copy()
    p_ref = oracle_lib.project(a, b, 3, oracle.GS_LEX, want_p=True)
    p = np.full((R, R), 7.0, np.float32)  # garbage must never be read at sweep 0
    gs_tile_model(p, d, 3, 1, 0, 0, False, False, 1, TW)
    assert bits_equal(p, p_ref)


@pytest.mark.parametrize("R,M", [(3, 2), (10, 4), (67, 2), (70, 8)])
def test_gs_unit_schedule_reproduces_gauss_seidel(R, M, oracle_lib):
    """Temporally blocked Gauss-Seidel (k_gs_unit): M sweeps per skewed tile,
    one wave per sweep in lockstep, in place; bit-identical to GS-lex."""
    rho, u, v, *_ = W.random_state(R, seed=R + 2)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    for mode in range(3):
        for K in (1, 5, 9):
            x = rho.copy()
            gs_unit_model(x, rho.copy(), K, 0, a, den, mode == 2, mode == 1, 0, M)
            ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, K, oracle.GS_LEX)
            assert bits_equal(x, ref), (mode, K)
    d = oracle_lib.divergence(u, v)
    aa, bb = u.copy(), v.copy()
    p_ref = oracle_lib.project(aa, bb, 9, oracle.GS_LEX, want_p=True)
    p = np.full((R, R), 7.0, np.float32)
    gs_unit_model(p, d, 9, 1, 0, 0, False, False, 1, M)
    assert bits_equal(p, p_ref)


@pytest.mark.parametrize("R,M,TW,RW", [(3, 2, 4, 4), (10, 2, 4, 4), (23, 4, 8, 8), (40, 4, 8, 4), (34, 8, 16, 8)])
def test_gs_shear_schedule_reproduces_gauss_seidel(R, M, TW, RW, oracle_lib):
    """Sheared units (a candidate for the joint launches, not yet a kernel):
    tiles of RW x TW x M in (i + k, i + j + 2k), every lane active in every
    step, levels I + J + kappa; bit-identical to GS-lex, no two units of a
    level touching a cell another writes, no in-step read of a written cell.
    Small RW and TW give many tiles, edges and levels on a small grid."""
    rho, u, v, *_ = W.random_state(R, seed=R + 2)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    for mode in range(3):
        for K in (1, 5, 9):
            x = rho.copy()
            gs_shear_model(x, rho.copy(), K, 0, a, den, mode == 2, mode == 1, 0, M, TW, RW)
            ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, K, oracle.GS_LEX)
            assert bits_equal(x, ref), (mode, K)
    d = oracle_lib.divergence(u, v)
    aa, bb = u.copy(), v.copy()
    p_ref = oracle_lib.project(aa, bb, 9, oracle.GS_LEX, want_p=True)
    p = np.full((R, R), 7.0, np.float32)
    gs_shear_model(p, d, 9, 1, 0, 0, False, False, 1, M, TW, RW)
    assert bits_equal(p, p_ref)


@pytest.mark.parametrize("R,M,L,TW,RW", [(10, 2, 1, 4, 4), (23, 4, 2, 8, 8), (40, 4, 4, 8, 4), (34, 8, 4, 16, 8),
                                        (34, 8, 8, 16, 8)])
def test_gs_shear_lagged_schedule_reproduces_gauss_seidel(R, M, L, TW, RW, oracle_lib):
    """Sheared units with lagged waves (a barrier every L ticks, every operand
    of an interval read right after it, the held west halo read once at the
    unit's start): bit-identical to GS-lex, no interval race, no held halo
    overwritten before its lane's first step."""
    rho, u, v, *_ = W.random_state(R, seed=R + 3)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    for mode in range(3):
        for K in (1, 5, 9):
            x = rho.copy()
            gs_shear_lag_model(x, rho.copy(), K, 0, a, den, mode == 2, mode == 1, 0, M, L, TW, RW)
            ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, K, oracle.GS_LEX)
            assert bits_equal(x, ref), (mode, K)
    d = oracle_lib.divergence(u, v)
    aa, bb = u.copy(), v.copy()
    p_ref = oracle_lib.project(aa, bb, 9, oracle.GS_LEX, want_p=True)
    p = np.full((R, R), 7.0, np.float32)
    gs_shear_lag_model(p, d, 9, 1, 0, 0, False, False, 1, M, L, TW, RW)
    assert bits_equal(p, p_ref)


def test_gs_shear_plan_size_at_4096():
    """The plan behind the sheared-unit estimate: 4096^2, K = 80, 64 x 128 x 8
    units -- about as many units as the rectangles, about 30 % more levels."""
    rect, shear = gs_unit_plan(4098, 80, 128, 8), gs_shear_plan_counts(4098, 80, 128, 8)
    n_rect = sum(len(l) for l in rect)
    assert len(rect) == 107 and n_rect == 21450
    assert shear == (141, 22374)


def gs_shear_plan_counts(R, K, TW, M, RW=64):
    """(levels, units) of gs_shear_plan without testing every unit for an
    interior cell (an upper bound on the units; the levels are the same)."""
    n = R - 2
    GI, GK = -(-(n + K - 1) // RW), -(-K // M)
    levels, units = set(), 0
    for kap in range(GK):
        k_hi = min(kap * M + M, K) - 1
        for I in range(GI):
            for J in range((RW * I + 1) // TW, (RW * I + RW + n + 2 * k_hi) // TW + 1):
                levels.add(I + J + kap)
                units += 1
    return len(levels), units


@pytest.mark.parametrize("R,M,L", [(3, 2, 1), (10, 4, 4), (67, 2, 8), (70, 8, 8)])
def test_gs_unit_lagged_schedule_reproduces_gauss_seidel(R, M, L, oracle_lib):
    """Lagged waves (sweep w trails sweep w-1 by L steps, one barrier per L
    steps, operands of L steps read at once): bit-identical to GS-lex and no
    wave writes, within an interval, a cell another wave read in it."""
    rho, u, v, *_ = W.random_state(R, seed=R + 3)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    for mode in range(3):
        for K in (1, 9):
            x = rho.copy()
            gs_unit_lag_model(x, rho.copy(), K, 0, a, den, mode == 2, mode == 1, 0, M, L)
            ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, K, oracle.GS_LEX)
            assert bits_equal(x, ref), (mode, K)
    d = oracle_lib.divergence(u, v)
    aa, bb = u.copy(), v.copy()
    p_ref = oracle_lib.project(aa, bb, 9, oracle.GS_LEX, want_p=True)
    p = np.full((R, R), 7.0, np.float32)
    gs_unit_lag_model(p, d, 9, 1, 0, 0, False, False, 1, M, L)
    assert bits_equal(p, p_ref)


@pytest.mark.parametrize("R,M,L", [(3, 2, 1), (10, 4, 4), (67, 2, 8), (70, 8, 8), (70, 4, 4), (70, 8, 4), (131, 8, 4)])
def test_gs_unit_ring_as_data_reproduces_gauss_seidel(R, M, L, oracle_lib):
    """The kernel's final form: lagged waves, ring cells written as data by
    the edge cells (in LDS and at write-back), zero-filled pressure start."""
    rho, u, v, *_ = W.random_state(R, seed=R + 4)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    for mode in range(3):
        for K in (1, 9):
            x = rho.copy()
            gs_unit_ring_model(x, rho.copy(), K, 0, a, den, mode == 2, mode == 1, M, L)
            ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, K, oracle.GS_LEX)
            assert bits_equal(x, ref), (mode, K)
    d = oracle_lib.divergence(u, v)
    aa, bb = u.copy(), v.copy()
    p_ref = oracle_lib.project(aa, bb, 9, oracle.GS_LEX, want_p=True)
    p = np.zeros((R, R), np.float32)
    gs_unit_ring_model(p, d, 9, 1, 0, 0, False, False, M, L)
    assert bits_equal(p, p_ref)


@pytest.mark.parametrize("R,M,L", [(10, 8, 4), (131, 8, 4), (200, 4, 4), (140, 8, 8)])
def test_gs_unit_wide_tiles_reproduce_gauss_seidel(R, M, L, oracle_lib):
    """Units of 64 rows x 128 columns (the three step phases of the plain
    path): bit-identical to GS-lex, race-free within intervals and levels."""
    rho, u, v, *_ = W.random_state(R, seed=R + 6)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    for mode in range(3):
        x = rho.copy()
        gs_unit_ring_model(x, rho.copy(), 9, 0, a, den, mode == 2, mode == 1, M, L, TW=128)
        ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, 9, oracle.GS_LEX)
        assert bits_equal(x, ref), mode
    d = oracle_lib.divergence(u, v)
    aa, bb = u.copy(), v.copy()
    p_ref = oracle_lib.project(aa, bb, 9, oracle.GS_LEX, want_p=True)
    p = np.zeros((R, R), np.float32)
    gs_unit_ring_model(p, d, 9, 1, 0, 0, False, False, M, L, TW=128)
    assert bits_equal(p, p_ref)


@pytest.mark.parametrize("R,M,L", [(3, 2, 1), (10, 4, 2), (67, 2, 4), (70, 8, 4), (70, 4, 2), (131, 8, 2)])
def test_gs_unit_prefetch_schedule_reproduces_gauss_seidel(R, M, L, oracle_lib):
    """The prefetching unit: wave w + 1 trails wave w by 2L ticks, and each
    wave reads an interval's operands one interval ahead.  Bit-identical to
    GS-lex, and no other wave writes an operand between its read and use."""
    rho, u, v, *_ = W.random_state(R, seed=R + 5)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    for mode in range(3):
        for K in (1, 9):
            x = rho.copy()
            gs_unit_ring_model(x, rho.copy(), K, 0, a, den, mode == 2, mode == 1, M, L, lag=2 * L, prefetch=True)
            ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, K, oracle.GS_LEX)
            assert bits_equal(x, ref), (mode, K)
    d = oracle_lib.divergence(u, v)
    aa, bb = u.copy(), v.copy()
    p_ref = oracle_lib.project(aa, bb, 9, oracle.GS_LEX, want_p=True)
    p = np.zeros((R, R), np.float32)
    gs_unit_ring_model(p, d, 9, 1, 0, 0, False, False, M, L, lag=2 * L, prefetch=True)
    assert bits_equal(p, p_ref)


def test_gs_unit_prefetch_needs_the_longer_lag(oracle_lib):
    """With a lag of L ticks, an operand read one interval ahead has not been
    written yet: the model's race check (or the bit comparison) catches it."""
    R, M, L = 70, 4, 4
    rho, *_ = W.random_state(R, seed=9)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    x = rho.copy()
    try:
        gs_unit_ring_model(x, rho.copy(), 9, 0, a, den, False, False, M, L, lag=L, prefetch=True)
    except AssertionError:
        return
    ref = rho.copy(); oracle_lib.diffuse(ref, 0, 5e-2, 0.1, 9, oracle.GS_LEX)
    assert not bits_equal(x, ref)


@pytest.mark.parametrize("R", [3, 10, 67])
@pytest.mark.parametrize("M", [1, 2, 3])
def test_jacobi_block_schedule_reproduces_jacobi(R, M, oracle_lib):
    rho, u, v, *_ = W.random_state(R, seed=R + 1)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    for mode in range(3):
        y = jacobi_block_model(rho.copy(), rho.copy(), 5, 0, a, den, mode == 2, mode == 1, 0, M)
        ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, 5, oracle.JACOBI)
        assert bits_equal(y, ref), mode
    d = oracle_lib.divergence(u, v)
    aa, bb = u.copy(), v.copy()
    p_ref = oracle_lib.project(aa, bb, 4, oracle.JACOBI, want_p=True)
    p = jacobi_block_model(np.full((R, R), 7.0, np.float32), d, 4, 1, 0, 0, False, False, 1, M)
    assert bits_equal(p, p_ref)


@pytest.mark.parametrize("R,M,TY", [(3, 2, 8), (4, 4, 8), (10, 3, 8), (34, 4, 16), (67, 2, 32), (70, 8, 16)])
def test_jacobi_wave_schedule_reproduces_jacobi(R, M, TY, oracle_lib):
    """k_jacobi_wave's tiling: register strips of TY + 2M rows x 64 lanes,
    region shrinking by one row and lane per sweep, rings rebuilt in the
    region after every sweep; remainder launches use their own margin."""
    rho, u, v, *_ = W.random_state(R, seed=R)
    a = np.float32(0.1) * np.float32(R * R) * np.float32(5e-2)
    den = np.float32(1) + np.float32(4) * a
    for mode in range(3):
        for K in (1, 5, 9):
            y = jacobi_wave_model(rho.copy(), rho.copy(), K, 0, a, den, mode == 2, mode == 1, 0, M, TY)
            ref = rho.copy(); oracle_lib.diffuse(ref, mode, 5e-2, 0.1, K, oracle.JACOBI)
            assert bits_equal(y, ref), (mode, K)
    d = oracle_lib.divergence(u, v)
    aa, bb = u.copy(), v.copy()
    p_ref = oracle_lib.project(aa, bb, 9, oracle.JACOBI, want_p=True)
    p = jacobi_wave_model(np.full((R, R), 7.0, np.float32), d, 9, 1, 0, 0, False, False, 1, M, TY)
    assert bits_equal(p, p_ref)


# ------------------------------------------------ workloads
def test_splitmix64_known_answer():
    # first outputs of SplitMix64 seeded with 0 (Vigna's reference implementation)
    g = W.SplitMix64(0)
    assert [g.next_u64() for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    g1, g2 = W.SplitMix64(42), W.SplitMix64(42)
    arr = g1.uniform_array(5)
    assert [g2.uniform() for _ in range(5)] == list(arr)
    assert g1.next_u64() == g2.next_u64()


def test_config1_sources():
    rs, us, vs = W.disc_sources(130)
    disc = rs > 0
    assert int(disc.sum()) == 208  # cells with (i-64.5)^2 + (j-64.5)^2 <= 64
    assert np.all(vs[disc] == -5.0) and np.all(rs[disc] == 100.0) and not us.any()


def test_vortex_state_deterministic_and_divergence_free():
    a = W.vortex_state(66, seed=5, n_vortices=8)
    b = W.vortex_state(66, seed=5, n_vortices=8)
    for x, y in zip(a, b):
        assert bits_equal(x, y)
    rho, u, v = a
    speed = np.sqrt(u[1:-1, 1:-1].astype(np.float64) ** 2 + v[1:-1, 1:-1] ** 2).max()
    assert abs(speed - 2.0 / 66) < 1e-6
    # resolved vortices (radii 10-40 cells): the discrete divergence of the
    # analytic curl field is a small fraction of its derivative terms
    uu, vv = W.vortex_velocity(130, 8, W.SplitMix64(3), radius_scale=1.0)
    uu, vv = uu.astype(np.float64), vv.astype(np.float64)
    du, dv = uu[1:-1, 2:] - uu[1:-1, :-2], vv[2:, 1:-1] - vv[:-2, 1:-1]
    inner = (slice(3, -3), slice(3, -3))
    assert np.abs(du + dv)[inner].max() < 0.02 * (np.abs(du) + np.abs(dv))[inner].max()
    assert 0.0 <= rho.min() and rho.max() <= 1.0


@pytest.mark.parametrize("R,Kd,Kp,TW,M", [(34, 8, 8, 64, 4), (66, 20, 20, 64, 8), (130, 20, 13, 128, 8),
                                          (130, 13, 20, 64, 4), (200, 40, 40, 128, 8), (99, 7, 9, 64, 2),
                                          (258, 20, 20, 128, 4)])
def test_pipeline_offset_is_the_brute_force_minimum(R, Kd, Kp, TW, M):
    """The level offset the library pipelines a step's first pressure solve by
    (a closed form over the pressure plan) equals the smallest offset that a
    brute-force model allows: every pressure unit after every diffuse unit
    that writes (interior or ring) a velocity cell its divergence reads."""
    from schedule_models import gs_pipeline_min_offset
    assert fs2d_amd.gs_pipeline_offset(R, Kd, Kp, TW, M) == gs_pipeline_min_offset(R, Kd, Kp, TW, M)


def test_pipeline_offset_at_the_workload_sizes():
    """4096^2 and 8192^2 with K = 80 (128-column units of 8 sweeps): 13 levels;
    1024^2 with K = 40: 7 -- against 107, 203 and 55 levels per solve."""
    assert fs2d_amd.gs_pipeline_offset(4098, 80, 80, 128, 8) == 13
    assert fs2d_amd.gs_pipeline_offset(8194, 80, 80, 128, 8) == 13
    assert fs2d_amd.gs_pipeline_offset(1026, 40, 40, 128, 8) == 7
