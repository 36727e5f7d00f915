"""GPU parity of libfs2d_hip.so against the oracle and the reference's golden outputs.

Every test drives the HIP solver through the C ABI (fs2d_amd -> ctypes ->
libfs2d_hip.so) and compares with the CPU oracle on the same inputs, or with
the fixtures the reference's own compiled solver wrote (tests/golden/).

Tolerance: the Gauss-Seidel path and the Jacobi path (against the oracle's
Jacobi order) must be BIT-IDENTICAL -- every stage evaluates the reference's
float32 expressions in the reference's order with IEEE rounding, so any
difference is a bug.  (rel-L2 and max|diff| are printed on failure.)
"""
from __future__ import annotations

import numpy as np
import pytest

import fs2d_amd
import oracle
from conftest import bits_equal, describe, golden
from fs2d_amd import workloads as W

pytestmark = pytest.mark.gpu

GS, JAC = fs2d_amd.GS_LEX, fs2d_amd.JACOBI
SIZES = [3, 4, 6, 10, 34, 67, 70, 130, 258]


def assert_same(got, want, what):
    assert bits_equal(got, want), f"{what}: {describe(got, want)}"


def assert_same_nan(got, want, what):
    gn, wn = np.isnan(got), np.isnan(want)
    assert np.array_equal(gn, wn), f"{what}: NaN pattern differs"
    g, w = np.where(gn, 0, got).astype(np.float32), np.where(wn, 0, want).astype(np.float32)
    assert bits_equal(g, w), f"{what}: {describe(g, w)}"


def make(R, state, order=GS, kd=20, kp=20, block=0, gsb=0, lag=0, jk=0):
    rho, u, v, rs, us, vs = state
    s = fs2d_amd.Solver(R, diffuse_iterations=kd, pressure_iterations=kp, order=order, jacobi_block=block,
                        gs_block=gsb, gs_lag=lag, jacobi_kernel=jk)
    s.upload(rho, u, v)
    s.set_sources(rs, us, vs)
    return s


def oracle_order(order):
    return oracle.GS_LEX if order == GS else oracle.JACOBI


# ---------------------------------------------------------------- stages
@pytest.mark.parametrize("R", SIZES)
@pytest.mark.parametrize("field", [0, 1, 2])
def test_set_boundary(R, field, oracle_lib):
    st = W.random_state(R, seed=R)
    with make(R, st) as s:
        s.set_boundary(field)
        got = s.download()[field]
    want = st[field].copy()
    oracle_lib.set_boundary(want, field)
    assert_same(got, want, f"set_boundary R={R} field={field}")


@pytest.mark.parametrize("R", SIZES)
def test_add_sources(R, oracle_lib):
    st = W.random_state(R, seed=R)
    with make(R, st) as s:
        for f in range(3):
            s.add_sources(f, 0.1)
        got = s.download()
    for f in range(3):
        want = st[f].copy()
        oracle_lib.add_sources(want, st[3 + f], 0.1)
        assert_same(got[f], want, f"add_sources R={R} field={f}")


WAVE, LDS = fs2d_amd.JACOBI_WAVE, fs2d_amd.JACOBI_LDS
# (order, jacobi_block, gs_block, gs_lag, jacobi_kernel)
GS_VARIANTS = [(GS, 0, 1, 0, 0), (GS, 0, 2, 1, 0), (GS, 0, 2, 8, 0), (GS, 0, 4, 4, 0), (GS, 0, 8, 1, 0),
               (GS, 0, 8, 8, 0)]
JAC_VARIANTS = [(JAC, 1, 0, 0, WAVE), (JAC, 3, 0, 0, WAVE), (JAC, 4, 0, 0, WAVE), (JAC, 8, 0, 0, WAVE),
                (JAC, 1, 0, 0, LDS), (JAC, 4, 0, 0, LDS), (JAC, 8, 0, 0, LDS)]
SWEEP_VARIANTS = GS_VARIANTS + JAC_VARIANTS


@pytest.mark.parametrize("order,block,gsb,lag,jk", SWEEP_VARIANTS)
@pytest.mark.parametrize("R", [4, 10, 67, 130, 258])
@pytest.mark.parametrize("rate", [0.0, 1e-3, 5e-2])
def test_diffuse(R, rate, order, block, gsb, lag, jk, oracle_lib):
    st = W.random_state(R, seed=11 * R)
    for K in (1, 2, 7, 20):
        with make(R, st, order=order, block=block, gsb=gsb, lag=lag, jk=jk) as s:
            for f in range(3):
                s.diffuse(f, rate, 0.1, K)
            got = s.download()
        for f in range(3):
            want = st[f].copy()
            oracle_lib.diffuse(want, f, rate, 0.1, K, oracle_order(order))
            assert_same(got[f], want, f"diffuse R={R} field={f} rate={rate} K={K} order={order} "
                                      f"block={block} gs_block={gsb} lag={lag} jk={jk}")


@pytest.mark.parametrize("R", SIZES)
@pytest.mark.parametrize("vel", [0.0, 0.05, 1.0, 6.0])
def test_scatter_density(R, vel, oracle_lib):
    st = W.random_state(R, seed=5 * R + 1, vel_scale=max(vel, 1e-9))
    if vel == 0.0:
        st = (st[0], np.zeros_like(st[1]), np.zeros_like(st[2])) + st[3:]
    with make(R, st) as s:
        s.advect_density(0.1)
        got = s.download()[0]
    want = st[0].copy()
    oracle_lib.advect(want, st[1], st[2], 0, 0.1, True)
    assert_same(got, want, f"scatter R={R} vel={vel}")


@pytest.mark.parametrize("fast", [(5.0, 0.0), (-3.0, 2.5), (1e6, 0.0)])
def test_scatter_one_fast_cell(fast, oracle_lib):
    """A slow field with one fast cell at 1024^2: the largest speed, not the
    typical one, sets how far sources may move (513 cells for u = 5; a source
    at 1e6 leaves the grid).  The sorted path's work does not depend on it:
    bit for bit against the oracle, and the advect finishes in bounded time."""
    import time
    R = 1026
    rho, u, v, *_ = W.random_state(R, seed=77, vel_scale=0.002)
    u, v = u.copy(), v.copy()
    u[400, 300], v[400, 300] = fast
    st = (rho, u, v, None, None, None)
    with fs2d_amd.Solver(R) as s:
        s.upload(rho, u, v)
        s.advect_density(0.1)
        s.sync()
        s.upload(rho, u, v)
        t0 = time.perf_counter()
        s.advect_density(0.1)
        s.sync()
        elapsed = time.perf_counter() - t0
        got = s.download()[0]
    want = rho.copy()
    oracle_lib.advect(want, u, v, 0, 0.1, True)
    assert_same(got, want, f"scatter with one fast cell {fast}")
    assert elapsed < 0.05, f"advect took {elapsed * 1e3:.1f} ms"


@pytest.mark.parametrize("reach", [0.5, 1.5, 2.9, 3.9, 4.1, 5.5, 9.0])
def test_scatter_around_the_window_limit(reach, oracle_lib):
    """Every radius of the window path (1 to 4: sources reach at most `reach`
    cells) and the sorted path just past it."""
    R = 300
    rho, u, v, *_ = W.random_state(R, seed=int(reach * 10), vel_scale=1.0)
    dt = 0.1
    scale = np.float32(reach / (np.sqrt(np.float32(R * R)) * dt) / max(np.abs(u).max(), np.abs(v).max()))
    u, v = (u * scale).astype(np.float32), (v * scale).astype(np.float32)
    with fs2d_amd.Solver(R) as s:
        s.upload(rho, u, v)
        s.advect_density(dt)
        got = s.download()[0]
    want = rho.copy()
    oracle_lib.advect(want, u, v, 0, dt, True)
    assert_same(got, want, f"scatter reach={reach}")


@pytest.mark.parametrize("R", SIZES)
@pytest.mark.parametrize("vel", [0.0, 0.05, 1.0, 6.0])
def test_gather_velocity(R, vel, oracle_lib):
    st = W.random_state(R, seed=3 * R + 2, vel_scale=max(vel, 1e-9))
    if vel == 0.0:
        st = (st[0], np.zeros_like(st[1]), np.zeros_like(st[2])) + st[3:]
    with make(R, st) as s:
        s.advect_velocity(0.1)
        _, gu, gv = s.download()
    tu, tv = st[1].copy(), st[2].copy()
    wu, wv = st[1].copy(), st[2].copy()
    oracle_lib.advect(wu, tu, tv, 1, 0.1, False)
    oracle_lib.advect(wv, tu, tv, 2, 0.1, False)
    assert_same(gu, wu, f"gather u R={R} vel={vel}")
    assert_same(gv, wv, f"gather v R={R} vel={vel}")


def test_non_finite_velocity(oracle_lib):
    """NaN / inf velocities: scatter skips the source (the reference's size_t
    conversion of NaN is undefined behaviour), the gather clamps NaN to R-1.5
    through std::min/std::max exactly as the reference does."""
    R = 34
    st = list(W.random_state(R, seed=99, vel_scale=0.5))
    st[1][5, 7] = np.nan
    st[2][9, 3] = np.inf
    st[1][20, 20] = -np.inf
    st[2][12, 30] = np.nan
    with make(R, st) as s:
        s.advect_density(0.1)
        s.advect_velocity(0.1)
        got = s.download()
    want = [x.copy() for x in st[:3]]
    oracle_lib.advect(want[0], st[1], st[2], 0, 0.1, True)
    tu, tv = want[1].copy(), want[2].copy()
    oracle_lib.advect(want[1], tu, tv, 1, 0.1, False)
    oracle_lib.advect(want[2], tu, tv, 2, 0.1, False)
    for f in range(3):
        assert_same_nan(got[f], want[f], f"non-finite field={f}")


def extreme_state(R, seed, with_inf=False):
    """Fields mixing ordinary values with exact zeros (a whole zero band and
    scattered zeros), +-tiny values below 2^-100, subnormals and values above
    2^100: the dividends the FMA-corrected division hands to the IEEE one."""
    rng = np.random.default_rng(seed)
    st = list(W.random_state(R, seed=seed))
    for f in range(3):
        g = st[f]
        g[R // 3:R // 2, :] = 0.0
        for frac, mag in ((0.3, 0.0), (0.1, 1e-33), (0.05, 1e-40), (0.02, 1e32)):
            m = rng.random((R, R)) < frac
            g[m] = (mag * rng.choice([-1.0, 1.0], size=(R, R))).astype(np.float32)[m]
        if with_inf:
            g[R // 2 + 1, R // 4] = np.inf
    return tuple(st)


@pytest.mark.parametrize("order,block,gsb,lag,jk", [(GS, 0, 8, 4, 0), (GS, 0, 1, 0, 0), (JAC, 4, 0, 0, WAVE),
                                                    (JAC, 8, 0, 0, WAVE), (JAC, 3, 0, 0, LDS)])
@pytest.mark.parametrize("R", [10, 67, 130])
@pytest.mark.parametrize("rate", [5e-2, 1e4, 1e30])
@pytest.mark.parametrize("with_inf", [False, True])
def test_diffuse_extreme_values(R, rate, with_inf, order, block, gsb, lag, jk, oracle_lib):
    """Zeros, tiny, subnormal, huge and infinite values, and diffusion rates
    whose 1 + 4a exceeds 2^24 (division floor above 2^-100) and 2^100 (every
    nonzero dividend through the IEEE division)."""
    st = extreme_state(R, seed=R + int(with_inf), with_inf=with_inf)
    for K in (1, 6):
        with make(R, st, order=order, block=block, gsb=gsb, lag=lag, jk=jk) as s:
            for f in range(3):
                s.diffuse(f, rate, 0.1, K)
            got = s.download()
        for f in range(3):
            want = st[f].copy()
            oracle_lib.diffuse(want, f, rate, 0.1, K, oracle_order(order))
            assert_same_nan(got[f], want, f"extreme diffuse R={R} field={f} rate={rate} K={K} inf={with_inf} "
                                             f"order={order} block={block} jk={jk}")


def unchecked_rate(R, dt, start):
    """A rate near `start` whose divisor 1 + 4 (dt R^2 rate) the two-product
    quotient alone does not divide correctly for every dividend (the host's
    check), so the unit kernels take the Markstein correction."""
    for k in range(4000):
        rate = np.float32(start * (1 + k / 997.0))
        a = np.float32(dt) * np.float32(R * R) * rate
        if not fs2d_amd.division_plan(float(np.float32(1) + np.float32(4) * a))["checked"]:
            return float(rate)
    raise AssertionError(f"no unchecked divisor near rate {start} at R={R}")


@pytest.mark.parametrize("order,block,gsb,lag,jk", [v for v in GS_VARIANTS if v[2] > 1])
@pytest.mark.parametrize("R", [67, 130, 258])
def test_diffuse_unchecked_divisor(R, order, block, gsb, lag, jk, oracle_lib):
    """The unit kernels' Markstein-corrected quotient, for divisors the
    two-product check rejects (the other tests' divisors pass it)."""
    rate = unchecked_rate(R, 0.1, 1e-3)
    st = W.random_state(R, seed=5 * R)
    for K in (1, 7, 20):
        with make(R, st, order=order, block=block, gsb=gsb, lag=lag, jk=jk) as s:
            for f in range(3):
                s.diffuse(f, rate, 0.1, K)
            got = s.download()
        for f in range(3):
            want = st[f].copy()
            oracle_lib.diffuse(want, f, rate, 0.1, K, oracle_order(order))
            assert_same(got[f], want, f"unchecked-divisor diffuse R={R} field={f} rate={rate} K={K} gs_block={gsb} "
                                      f"lag={lag}")


@pytest.mark.parametrize("gsb,lag", [(4, 4), (8, 4), (8, 8)])
@pytest.mark.parametrize("which", ["density", "velocity"])
def test_steps_unchecked_divisor(gsb, lag, which, oracle_lib):
    """Whole steps (the diffuses and the first pressure solve in shared level
    launches) with one diffuse divisor that the two-product check rejects."""
    R = 130
    diff, visc = 1e-4, 1e-5
    if which == "density":
        diff = unchecked_rate(R, 0.05, diff)
    else:
        visc = unchecked_rate(R, 0.05, 2e-4)
    st = W.vortex_state(R, seed=3, n_vortices=16)
    rs, us, vs = W.disc_sources(R, radius=R / 16)
    want = [x.copy() for x in st]
    with fs2d_amd.Solver(R, diffuse_iterations=12, pressure_iterations=30, order=GS, gs_block=gsb, gs_lag=lag) as s:
        s.upload(*(x.copy() for x in st))
        s.set_sources(rs, us, vs)
        s.step_n(4, diff, visc, 0.05)
        for _ in range(4):
            oracle_lib.solve(want[0], rs, diff, want[1], want[2], us, vs, visc, 0.05, 12, 30, oracle.GS_LEX)
        got = s.download()
    for f, name in enumerate(("density", "u", "v")):
        assert_same(got[f], want[f], f"{name} after 4 steps, unchecked {which} divisor, gs_block={gsb} lag={lag}")


@pytest.mark.parametrize("order,block,gsb,lag,jk", GS_VARIANTS + [(JAC, 2, 0, 0, WAVE), (JAC, 4, 0, 0, WAVE),
                                                                  (JAC, 5, 0, 0, WAVE), (JAC, 8, 0, 0, WAVE),
                                                                  (JAC, 4, 0, 0, LDS)])
@pytest.mark.parametrize("R", [3, 4, 10, 67, 130, 258])
def test_project(R, order, block, gsb, lag, jk, oracle_lib):
    st = W.random_state(R, seed=7 * R, vel_scale=1.0)
    for K in (0, 1, 3, 20):
        with make(R, st, order=order, block=block, gsb=gsb, lag=lag, jk=jk) as s:
            s.project(K)
            _, gu, gv = s.download()
            gp = s.pressure()
        wu, wv = st[1].copy(), st[2].copy()
        wp = oracle_lib.project(wu, wv, K, oracle_order(order), want_p=True)
        assert_same(gp, wp, f"pressure R={R} K={K} order={order} gs_block={gsb} lag={lag} block={block} jk={jk}")
        assert_same(gu, wu, f"project u R={R} K={K} order={order}")
        assert_same(gv, wv, f"project v R={R} K={K} order={order}")


@pytest.mark.parametrize("order", [GS, JAC])
def test_projection_against_float64_oracle(order, oracle_lib):
    """The float32 projection against the same projection carried out in
    float64 (the oracle's fso_project_f64, same order and sweep count): the
    velocities agree to float32 rounding, and so does the post-projection
    divergence, computed in float64 from each result."""
    R, K = 258, 40
    rho, u, v = W.vortex_state(R, seed=23, n_vortices=16)
    with make(R, (rho, u, v, None, None, None), order=order, kp=K) as s:
        s.project(K)
        _, gu, gv = s.download()
    ud, vd = u.astype(np.float64), v.astype(np.float64)
    oracle_lib.project_f64(ud, vd, K, oracle_order(order))
    scale = max(np.abs(ud).max(), np.abs(vd).max())
    assert np.abs(gu - ud).max() < 1e-5 * scale and np.abs(gv - vd).max() < 1e-5 * scale
    n32 = oracle_lib.divergence_norm_f64(gu, gv)
    n64 = oracle_lib.divergence_norm_f64(ud.astype(np.float32), vd.astype(np.float32))
    n0 = oracle_lib.divergence_norm_f64(u, v)
    assert n64 < n0 and abs(n32 - n64) <= 1e-3 * n64, (n32, n64, n0)


# 64 x 128 tiles: units of 128 columns (and the one-sweep tile kernel at 128 columns)
WIDE_VARIANTS = [(8, 4), (8, 1), (8, 8), (4, 4), (4, 8), (1, 0)]


@pytest.mark.parametrize("gsb,lag", WIDE_VARIANTS)
@pytest.mark.parametrize("R", [10, 67, 130, 258, 300])
def test_wide_tiles_vs_oracle(R, gsb, lag, oracle_lib):
    """Diffuse (all three boundary rules), projection and whole steps with
    128-column Gauss-Seidel tiles, bit for bit against the oracle."""
    st = W.random_state(R, seed=13 * R, vel_scale=1.0)

    def solver(kd=20, kp=20):
        s = fs2d_amd.Solver(R, diffuse_iterations=kd, pressure_iterations=kp, order=GS, gs_block=gsb, gs_lag=lag,
                            gs_tile_cols=128)
        s.upload(*st[:3])
        s.set_sources(*st[3:])
        return s

    for K in (1, 7, 20):
        with solver() as s:
            for f in range(3):
                s.diffuse(f, 5e-2, 0.1, K)
            got = s.download()
        for f in range(3):
            want = st[f].copy()
            oracle_lib.diffuse(want, f, 5e-2, 0.1, K, oracle.GS_LEX)
            assert_same(got[f], want, f"wide diffuse R={R} field={f} K={K} gs_block={gsb} lag={lag}")
    with solver() as s:
        s.project(20)
        _, gu, gv = s.download()
        gp = s.pressure()
    wu, wv = st[1].copy(), st[2].copy()
    wp = oracle_lib.project(wu, wv, 20, oracle.GS_LEX, want_p=True)
    assert_same(gp, wp, f"wide pressure R={R} gs_block={gsb} lag={lag}")
    assert_same(gu, wu, f"wide project u R={R}")
    want = [x.copy() for x in st[:3]]
    with solver(kd=12, kp=30) as s:
        s.step_n(3, 1e-4, 1e-5, 0.05)
        got = s.download()
    for _ in range(3):
        oracle_lib.solve(want[0], st[3], 1e-4, want[1], want[2], st[4], st[5], 1e-5, 0.05, 12, 30, oracle.GS_LEX)
    for f, name in enumerate(("density", "u", "v")):
        assert_same(got[f], want[f], f"wide step {name} R={R} gs_block={gsb} lag={lag}")


def scatter_launches(R):
    """Kernels of a step's density scatter: the window path, and the sorted
    path's keys, histograms, scans, scatters (one per 9-bit digit of the
    (R - 1)^2 + 1 target keys), bounds and merge."""
    passes = -(-((R - 1) ** 2).bit_length() // 9)
    return 3 + 3 * passes


@pytest.mark.parametrize("gsb,tw,lag", [(8, 0, 4), (8, 64, 0), (8, 128, 8), (4, 0, 0), (2, 0, 0), (0, 0, 0)])
def test_launch_policy_and_launch_count(gsb, tw, lag):
    """The Gauss-Seidel launch policy of each solve kind keeps the fixed
    parameters and chooses the others among the candidates (64 or 128
    columns, 4 or 8 sweeps per unit, lag 4 or 8); a step's launches equal the
    level counts of the schedule model's unit plans for the chosen policy."""
    from schedule_models import gs_unit_plan
    R, K = 258, 20
    with fs2d_amd.Solver(R, diffuse_iterations=K, pressure_iterations=K, gs_block=gsb, gs_tile_cols=tw,
                         gs_lag=lag) as s:
        pol = [s.gs_policy(kind) for kind in (0, 1, 2)]
        launches = s.launches_per_step()
    for q in pol:
        assert q["tile_cols"] == tw if tw else q["tile_cols"] in ((64,) if gsb == 2 else (64, 128)), q
        assert q["sweeps_per_unit"] == gsb if gsb else q["sweeps_per_unit"] in (4, 8), q
        assert q["lag"] == lag if lag else q["lag"] in (4, 8), q
    d, p = (len(gs_unit_plan(R, K, q["tile_cols"], q["sweeps_per_unit"])) for q in pol[:2])
    assert (pol[0]["levels"], pol[1]["levels"]) == (d, p)
    # the first projection's pressure solve runs in the diffuses' level launches
    # (the joint policy, tuned on its own), `off` levels behind; the second one
    # on its own
    tw, m = pol[2]["tile_cols"], pol[2]["sweeps_per_unit"]
    off = fs2d_amd.gs_pipeline_offset(R, K, K, tw, m)
    lv = {u[0][0] + u[0][1] + u[0][2] for u in gs_unit_plan(R, K, tw, m)}
    joint = len(lv | {x + off for x in lv})
    assert pol[2]["levels"] == joint, (pol[2], joint)
    # adds (one launch), fill p, joint levels, scatter, gradient + fold, gather, divergence (filling p), p levels,
    # gradient + fold
    assert launches == 1 + 1 + joint + scatter_launches(R) + 2 + 1 + 1 + p + 2, (d, p, joint, pol)


# ------------------------------------------------ the reference's own outputs
@pytest.mark.parametrize("R", [4, 6, 10, 18, 34])
def test_golden_stages(R):
    """GPU stages against vectors written by the reference's compiled solver."""
    g = golden(f"stages_R{R}.npz")
    st = tuple(g[k] for k in ("rho", "u", "v", "rs", "us", "vs"))
    for m in range(3):
        # the fixture applies ring mode m to rho: load rho into field m
        st_m = list(st)
        st_m[m] = g["rho"]
        with make(R, tuple(st_m)) as s:
            s.set_boundary(m)
            assert_same(s.download()[m], g[f"bnd{m}"], f"golden set_boundary mode={m}")
    with make(R, st) as s:
        s.add_sources(0, 0.1)
        assert_same(s.download()[0], g["add"], "golden add_sources")
    for m, key in ((0, "rho"), (1, "u"), (2, "v")):
        for rate in (1e-3, 5e-2):
            for it in (1, 7, 20):
                with make(R, st) as s:
                    s.diffuse(m, rate, 0.1, it)
                    assert_same(s.download()[m], g[f"diff_m{m}_r{rate:g}_k{it}"], f"golden diffuse m={m} {rate} {it}")
    for tag, (uk, vk) in (("slow", ("u", "v")), ("fast", ("uf", "vf"))):
        stt = (g["rho"], g[uk], g[vk]) + st[3:]
        with make(R, stt) as s:
            s.advect_density(0.1)
            assert_same(s.download()[0], g[f"scatter_{tag}"], f"golden scatter {tag}")
        with make(R, stt) as s:
            s.advect_velocity(0.1)
            _, gu, gv = s.download()
            assert_same(gu, g[f"gather_u_{tag}"], f"golden gather u {tag}")
            assert_same(gv, g[f"gather_v_{tag}"], f"golden gather v {tag}")
    for it in (1, 7, 20):
        stt = (g["rho"], g["uf"], g["vf"]) + st[3:]
        with make(R, stt) as s:
            s.project(it)
            _, gu, gv = s.download()
            assert_same(gu, g[f"proj_u_k{it}"], f"golden project u K={it}")
            assert_same(gv, g[f"proj_v_k{it}"], f"golden project v K={it}")


@pytest.mark.parametrize("R", [4, 6, 10, 18, 34])
def test_golden_solve(R):
    """Whole steps (fluid_solver::solve) against the reference, 1 and 10 steps,
    through the synchronous host-array entry point."""
    g = golden(f"solve_R{R}.npz")
    rho, u, v = g["rho"].copy(), g["u"].copy(), g["v"].copy()
    with fs2d_amd.Solver(R) as s:
        for step in range(1, 11):
            s.solve(rho, g["rs"], 1e-3, u, v, g["us"], g["vs"], 1e-3, 0.1)
            if step in (1, 10):
                assert_same(rho, g[f"rho_s{step}"], f"golden solve rho R={R} step {step}")
                assert_same(u, g[f"u_s{step}"], f"golden solve u R={R} step {step}")
                assert_same(v, g[f"v_s{step}"], f"golden solve v R={R} step {step}")


def test_config1_against_reference():
    """Configuration 1: 128x128 + ring, zero start, disc sources every step,
    dt=0.1, viscosity = diffusion = 0, K = 20, 200 steps; resident API."""
    g = golden("config1_R130.npz")
    R = 130
    with fs2d_amd.Solver(R) as s:
        s.upload(*W.zero_state(R))
        s.set_sources(g["rs"], g["us"], g["vs"])
        done = 0
        for target in (1, 10, 200):
            s.step_n(target - done, 0.0, 0.0, 0.1)
            done = target
            rho, u, v = s.download()
            assert_same(rho, g[f"rho_s{target}"], f"config1 density step {target}")
            assert_same(u, g[f"u_s{target}"], f"config1 u step {target}")
            assert_same(v, g[f"v_s{target}"], f"config1 v step {target}")


# ------------------------------------------------ whole steps vs the oracle
@pytest.mark.parametrize("order,block,gsb,lag,jk", GS_VARIANTS + [(JAC, 4, 0, 0, WAVE), (JAC, 8, 0, 0, WAVE),
                                                                  (JAC, 4, 0, 0, LDS)])
@pytest.mark.parametrize("R", [67, 130, 258])
def test_steps_vs_oracle(R, order, block, gsb, lag, jk, oracle_lib):
    st = W.vortex_state(R, seed=R, n_vortices=16)
    rs, us, vs = W.disc_sources(R, radius=R / 16)
    rho, u, v = (x.copy() for x in st)
    want = [x.copy() for x in st]
    with fs2d_amd.Solver(R, diffuse_iterations=12, pressure_iterations=30, order=order, jacobi_block=block,
                         gs_block=gsb, gs_lag=lag, jacobi_kernel=jk) as s:
        s.upload(rho, u, v)
        s.set_sources(rs, us, vs)
        done = 0
        for target in (1, 10):
            s.step_n(target - done, 1e-4, 1e-5, 0.05)
            for _ in range(target - done):
                oracle_lib.solve(want[0], rs, 1e-4, want[1], want[2], us, vs, 1e-5, 0.05, 12, 30,
                                 oracle_order(order))
            done = target
            got = s.download()
            for f, name in enumerate(("density", "u", "v")):
                assert_same(got[f], want[f], f"{name} R={R} order={order} block={block} jk={jk} after {target} steps")


@pytest.mark.parametrize("jk", [WAVE, LDS])
def test_jacobi_autotuned_depth(jk, oracle_lib):
    """jacobi_block = 0 times the depths 1, 2, 4, 8 on the first solve of each
    kind and keeps the fastest; the state must come out exactly as the
    oracle's Jacobi order gives it, whatever depth was chosen."""
    R = 130
    st = W.vortex_state(R, seed=5, n_vortices=16)
    rs, us, vs = W.disc_sources(R, radius=8)
    want = [x.copy() for x in st]
    with fs2d_amd.Solver(R, diffuse_iterations=20, pressure_iterations=40, order=JAC, jacobi_block=0,
                         jacobi_kernel=jk) as s:
        s.upload(*st)
        s.set_sources(rs, us, vs)
        s.step_n(3, 1e-4, 1e-5, 0.05)
        got = s.download()
        launches = s.launches_per_step()
    for _ in range(3):
        oracle_lib.solve(want[0], rs, 1e-4, want[1], want[2], us, vs, 1e-5, 0.05, 20, 40, oracle.JACOBI)
    for f, name in enumerate(("density", "u", "v")):
        assert_same(got[f], want[f], f"autotuned Jacobi {name} jk={jk}")
    # 3 adds, 3 diffuses of ceil(20/m_d), scatter, 2 x (div, ceil(40/m_p), grad), gather
    sc = scatter_launches(R)
    possible = {3 + 3 * -(-20 // md) + sc + 2 * (2 + -(-40 // mp)) + 1 for md in (1, 2, 4, 8) for mp in (1, 2, 4, 8)}
    assert launches in possible


@pytest.mark.parametrize("order", [GS, JAC])
def test_solve_timers_leave_the_state(order, oracle_lib):
    """The event-timed diffuse, joint and pressure solves run on scratch fields
    (the pressure timer only changes the pressure field): the next step from
    the same state is still the oracle's step."""
    R = 130
    st = W.vortex_state(R, seed=9, n_vortices=16)
    z = np.zeros((R, R), np.float32)
    want = [x.copy() for x in st]
    with fs2d_amd.Solver(R, diffuse_iterations=20, pressure_iterations=20, order=order) as s:
        s.upload(*st)
        s.set_sources(None, None, None)
        ms_d = s.time_diffuse_sweeps(1e-4, 1e-5, 0.05, 3)
        assert [np.array_equal(a, b) for a, b in zip(s.download(), st)] == [True] * 3
        if order == GS:
            assert s.time_joint_sweeps(1e-4, 1e-5, 0.05, 3) > 0
            assert s.time_pressure_sweeps(3) > 0
            assert [np.array_equal(a, b) for a, b in zip(s.download(), st)] == [True] * 3
        s.step_n(1, 1e-4, 1e-5, 0.05)
        got = s.download()
    oracle_lib.solve(want[0], z, 1e-4, want[1], want[2], z, z, 1e-5, 0.05, 20, 20,
                     oracle.GS_LEX if order == GS else oracle.JACOBI)
    assert ms_d > 0
    for f, name in enumerate(("density", "u", "v")):
        assert_same(got[f], want[f], f"step after the diffuse timer, {name}")


def test_config2_one_step(oracle_lib):
    """Configuration 2 shape: 1024^2 interior, 64-vortex start, dt=0.05,
    viscosity 1e-6, diffusion 1e-5, K = 40; one step, Gauss-Seidel order."""
    R = 1026
    st = W.vortex_state(R, seed=2)
    zeros = np.zeros((R, R), np.float32)
    got = [x.copy() for x in st]
    with fs2d_amd.Solver(R, diffuse_iterations=40, pressure_iterations=40) as s:
        s.solve(got[0], zeros, 1e-5, got[1], got[2], zeros, zeros, 1e-6, 0.05)
    want = [x.copy() for x in st]
    oracle_lib.solve(want[0], zeros, 1e-5, want[1], want[2], zeros, zeros, 1e-6, 0.05, 40, 40, oracle.GS_LEX)
    for f, name in enumerate(("density", "u", "v")):
        assert_same(got[f], want[f], f"config2 {name}")


# ------------------------------------------------ full-size (4096^2) checks
@pytest.fixture(scope="module")
def big_state():
    R = 4098
    return W.vortex_state(R, seed=3)


@pytest.mark.parametrize("order,gsb,lag,block,K", [(GS, 1, 0, 0, 3), (GS, 4, 1, 0, 3), (GS, 8, 8, 0, 3),
                                                   (GS, 8, 4, 0, 20), (JAC, 0, 0, 4, 8), (JAC, 0, 0, 8, 16)])
def test_full_size_stages(big_state, order, gsb, lag, block, K, oracle_lib):
    """4096^2 interior: each stage against the oracle with short solves (the
    oracle cannot afford 80 sweeps here); exercises the full tile grid."""
    R = 4098
    rho, u, v = big_state
    z = np.zeros((R, R), np.float32)
    st = (rho, u, v, z, z, z)
    with make(R, st, order=order, kd=K, kp=K, gsb=gsb, lag=lag, block=block) as s:
        s.diffuse(0, 1e-5, 0.05, K)
        s.advect_density(0.05)
        s.project(K)
        got = s.download()
    want = [rho.copy(), u.copy(), v.copy()]
    oracle_lib.diffuse(want[0], 0, 1e-5, 0.05, K, oracle_order(order))
    oracle_lib.advect(want[0], want[1], want[2], 0, 0.05, True)
    oracle_lib.project(want[1], want[2], K, oracle_order(order))
    for f in range(3):
        assert_same(got[f], want[f], f"4096^2 stages field {f} order={order} block={block}")


@pytest.mark.parametrize("order", [GS, JAC])
def test_full_size_step_deterministic(big_state, order):
    """Two independent runs of 3 full steps (K = 80) are bit-identical (in the
    Jacobi order whatever depths the autotuner picks in each), and the second
    projection leaves the wide-stencil divergence below its input."""
    R = 4098
    rho, u, v = big_state
    out = []
    for _ in range(2):
        with fs2d_amd.Solver(R, diffuse_iterations=80, pressure_iterations=80, order=order) as s:
            s.upload(rho, u, v)
            s.set_sources(None, None, None)
            s.step_n(3, 1e-5, 1e-6, 0.05)
            out.append(s.download())
    for f in range(3):
        assert bits_equal(out[0][f], out[1][f]), f"run-to-run difference in field {f}"
    o = oracle.Oracle()
    assert o.divergence_norm_f64(out[0][1], out[0][2]) < o.divergence_norm_f64(u, v)


# ------------------------------------------------ errors
def test_rejects_bad_shapes():
    with pytest.raises(fs2d_amd.Fs2dError, match="square"):
        fs2d_amd.Solver(10, 12)
    with pytest.raises(fs2d_amd.Fs2dError):
        fs2d_amd.Solver(2)
    with pytest.raises(fs2d_amd.Fs2dError):
        fs2d_amd.Solver(10, diffuse_iterations=-1)
