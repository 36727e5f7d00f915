"""Pins of the CPU oracle (oracle/fs2d_oracle.c).

The oracle is trusted as the checker of the HIP solver only after these pass:

1. bit-for-bit against golden vectors written by the reference's own compiled
   solver (tests/golden/, made by tests/golden/make_golden.py);
2. bit-for-bit against the compiled reference itself on fresh inputs, where
   oracle/_ref was built in this checkout;
3. closed forms and invariants the mathematics fixes, chosen so that a
   plausible slip (a dropped term, a wrong sign or index, a transposed
   operand) fails one of them -- these also pin the Jacobi order and the FP64
   twins, which the reference does not have.
"""
from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from conftest import bits_equal, describe, golden
from fs2d_amd import workloads as W

GS, JAC = oracle.GS_LEX, oracle.JACOBI


# ------------------------------------------------ 1. golden vectors
@pytest.mark.parametrize("R", [4, 6, 10, 18, 34])
def test_stages_match_golden(R, oracle_lib):
    g = golden(f"stages_R{R}.npz")
    for m in range(3):
        x = g["rho"].copy(); oracle_lib.set_boundary(x, m)
        assert bits_equal(x, g[f"bnd{m}"]), describe(x, g[f"bnd{m}"])
    x = g["rho"].copy(); oracle_lib.add_sources(x, g["rs"], 0.1)
    assert bits_equal(x, g["add"])
    for m, key in ((0, "rho"), (1, "u"), (2, "v")):
        for rate in (1e-3, 5e-2):
            for it in (1, 7, 20):
                x = g[key].copy(); oracle_lib.diffuse(x, m, rate, 0.1, it, GS)
                want = g[f"diff_m{m}_r{rate:g}_k{it}"]
                assert bits_equal(x, want), (m, rate, it, describe(x, want))
    for tag, (uk, vk) in (("slow", ("u", "v")), ("fast", ("uf", "vf"))):
        x = g["rho"].copy(); oracle_lib.advect(x, g[uk], g[vk], 0, 0.1, True)
        assert bits_equal(x, g[f"scatter_{tag}"]), describe(x, g[f"scatter_{tag}"])
        x = g[uk].copy(); oracle_lib.advect(x, g[uk], g[vk], 1, 0.1, False)
        assert bits_equal(x, g[f"gather_u_{tag}"])
        x = g[vk].copy(); oracle_lib.advect(x, g[uk], g[vk], 2, 0.1, False)
        assert bits_equal(x, g[f"gather_v_{tag}"])
    for it in (1, 7, 20):
        a, b = g["uf"].copy(), g["vf"].copy(); oracle_lib.project(a, b, it, GS)
        assert bits_equal(a, g[f"proj_u_k{it}"]) and bits_equal(b, g[f"proj_v_k{it}"])


@pytest.mark.parametrize("R", [4, 6, 10, 18, 34])
def test_solve_matches_golden(R, oracle_lib):
    g = golden(f"solve_R{R}.npz")
    rho, u, v = g["rho"].copy(), g["u"].copy(), g["v"].copy()
    for step in range(1, 11):
        oracle_lib.solve(rho, g["rs"], 1e-3, u, v, g["us"], g["vs"], 1e-3, 0.1, 20, 20, GS)
        if step in (1, 10):
            for name, x in (("rho", rho), ("u", u), ("v", v)):
                assert bits_equal(x, g[f"{name}_s{step}"]), (name, step, describe(x, g[f"{name}_s{step}"]))


def test_config1_matches_golden(oracle_lib):
    g = golden("config1_R130.npz")
    R = 130
    rho, u, v = W.zero_state(R)
    rs, us, vs = W.disc_sources(R)
    assert bits_equal(rs, g["rs"]) and bits_equal(us, g["us"]) and bits_equal(vs, g["vs"])
    for step in range(1, 201):
        oracle_lib.solve(rho, rs, 0.0, u, v, us, vs, 0.0, 0.1, 20, 20, GS)
        if step in (1, 10, 200):
            for name, x in (("rho", rho), ("u", u), ("v", v)):
                assert bits_equal(x, g[f"{name}_s{step}"]), (name, step)


# ------------------------------------------------ 2. the compiled reference
@pytest.mark.parametrize("R", [3, 5, 33, 66, 131])
def test_against_compiled_reference(R, oracle_lib, reference_lib):
    st = W.random_state(R, seed=1000 + R, vel_scale=0.3)
    for m in range(3):
        a, b = st[0].copy(), st[0].copy()
        oracle_lib.diffuse(a, m, 2e-2, 0.07, 13, GS); reference_lib.diffuse(b, m, 2e-2, 0.07, 13)
        assert bits_equal(a, b), describe(a, b)
    for trace, m in ((True, 0), (False, 1), (False, 2)):
        a, b = st[m].copy(), st[m].copy()
        oracle_lib.advect(a, st[1], st[2], m, 0.37, trace); reference_lib.advect(b, st[1], st[2], m, 0.37, trace)
        assert bits_equal(a, b), describe(a, b)
    a = [st[1].copy(), st[2].copy()]; b = [st[1].copy(), st[2].copy()]
    oracle_lib.project(a[0], a[1], 33, GS); reference_lib.project(b[0], b[1], 33)
    assert bits_equal(a[0], b[0]) and bits_equal(a[1], b[1])
    a = [x.copy() for x in st[:3]]; b = [x.copy() for x in st[:3]]
    for _ in range(3):
        oracle_lib.solve(a[0], st[3], 1e-3, a[1], a[2], st[4], st[5], 2e-3, 0.05, 20, 20, GS)
        reference_lib.solve(b[0], st[3], 1e-3, b[1], b[2], st[4], st[5], 2e-3, 0.05)
    for x, y in zip(a, b):
        assert bits_equal(x, y), describe(x, y)


# ------------------------------------------------ 3. closed forms, invariants
def test_ring_closed_form(oracle_lib):
    """Edges copy (continuous) or negate (u: left/right, v: top/bottom) the
    adjacent interior cell; corners average their two edge neighbours."""
    R = 9
    x = W.random_state(R, seed=1)[0]
    for m, (tb, lr) in enumerate(((1, 1), (1, -1), (-1, 1))):
        y = x.copy(); oracle_lib.set_boundary(y, m)
        assert np.array_equal(y[1:-1, 1:-1], x[1:-1, 1:-1])
        assert np.array_equal(y[0, 1:-1], tb * x[1, 1:-1]) and np.array_equal(y[-1, 1:-1], tb * x[-2, 1:-1])
        assert np.array_equal(y[1:-1, 0], lr * x[1:-1, 1]) and np.array_equal(y[1:-1, -1], lr * x[1:-1, -2])
        c = x[1, 1] if m == 0 else 0.0
        assert y[0, 0] == np.float32(c) and (m != 0 or y[-1, -1] == x[-2, -2])


@pytest.mark.parametrize("order", [GS, JAC])
def test_diffuse_rate_zero_is_identity(order, oracle_lib):
    R = 12
    x = W.random_state(R, seed=2)[0]
    y = x.copy(); oracle_lib.diffuse(y, 0, 0.0, 0.1, 20, order)
    assert np.array_equal(y[1:-1, 1:-1], x[1:-1, 1:-1])
    ring = x.copy(); oracle_lib.set_boundary(ring, 0)
    assert np.array_equal(y, ring)


@pytest.mark.parametrize("order", [GS, JAC])
def test_diffuse_constant_field_invariant(order, oracle_lib):
    R = 16
    x = np.full((R, R), 0.75, np.float32)  # 0.75 * (1 + 4a) / (1 + 4a) is exact for these a
    for rate in (1e-3, 1.0 / 1024):
        y = x.copy(); oracle_lib.diffuse(y, 0, rate, 0.5, 9, order)
        assert np.allclose(y, 0.75, rtol=0, atol=2e-7)


def _diffuse_matrix(R, a, lr_sign, tb_sign):
    """(1+4a) I - a * (neighbour sum with the ring folded in as the signed
    adjacent interior cell): the fixed point of every diffuse sweep."""
    n = R - 2
    idx = lambda i, j: i * n + j
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(n):
            diag = 1 + 4 * a
            for di, dj, s in ((-1, 0, tb_sign), (1, 0, tb_sign), (0, -1, lr_sign), (0, 1, lr_sign)):
                ii, jj = i + di, j + dj
                if 0 <= ii < n and 0 <= jj < n:
                    rows.append(idx(i, j)); cols.append(idx(ii, jj)); vals.append(-a)
                else:
                    diag -= a * s
            rows.append(idx(i, j)); cols.append(idx(i, j)); vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n * n, n * n))


@pytest.mark.parametrize("mode,tb,lr", [(0, 1, 1), (1, 1, -1), (2, -1, 1)])
def test_diffuse_converges_to_direct_solve(mode, tb, lr, oracle_lib):
    """Both sweep orders converge to the direct solution of the implicit step
    (a textbook reduction): pins the stencil, the signs and the ring."""
    R = 14
    x = W.random_state(R, seed=3)[0]
    rate, dt = 0.02, 0.1
    a = float(np.float32(dt) * np.float32(R * R) * np.float32(rate))
    A = _diffuse_matrix(R, a, lr, tb)
    exact = spla.spsolve(A.tocsc(), x[1:-1, 1:-1].astype(np.float64).ravel()).reshape(R - 2, R - 2)
    for order in (GS, JAC):
        y = x.copy(); oracle_lib.diffuse(y, mode, rate, dt, 400, order)
        assert np.abs(y[1:-1, 1:-1] - exact).max() < 2e-6, order


def test_jacobi_iterate_is_matrix_iteration(oracle_lib):
    """Jacobi order = x_{k+1} = D^-1 (b + a N x_k) with N the neighbour matrix
    (the ring of x_0 is the given, stale ring for the first sweep)."""
    R = 11
    x = W.random_state(R, seed=4)[0]
    rate, dt = 0.05, 0.1
    a = np.float32(dt) * np.float32(R * R) * np.float32(rate)
    y = x.copy(); oracle_lib.diffuse(y, 0, rate, dt, 3, JAC)
    cur = x.astype(np.float64)
    b = x[1:-1, 1:-1].astype(np.float64)
    for k in range(3):
        nb = cur[:-2, 1:-1] + cur[2:, 1:-1] + cur[1:-1, :-2] + cur[1:-1, 2:]
        new = cur.copy()
        new[1:-1, 1:-1] = (b + float(a) * nb) / (1 + 4 * float(a))
        W.ring(new, 0)
        cur = new
    assert np.abs(y - cur).max() < 1e-6
    g = x.copy(); oracle_lib.diffuse(g, 0, rate, dt, 3, GS)
    assert np.abs(g - cur).max() > 1e-4  # the two orders really differ


def test_scatter_zero_velocity_identity(oracle_lib):
    R = 10
    x = W.random_state(R, seed=5)[0]
    z = np.zeros((R, R), np.float32)
    y = x.copy(); oracle_lib.advect(y, z, z, 0, 0.1, True)
    ring = x.copy(); oracle_lib.set_boundary(ring, 0)
    assert np.array_equal(y, ring)


def test_scatter_integer_shift_and_mass(oracle_lib):
    """A uniform flow of exactly 2 cells per step translates the field; all
    interior mass that stays inside [0.5, R-1.5] is conserved."""
    R = 20
    x = np.zeros((R, R), np.float32)
    x[5:9, 4:10] = W.random_state(R, seed=6)[0][5:9, 4:10]
    dt = 0.125
    dt0 = np.float32(np.sqrt(np.float32(R * R))) * np.float32(dt)
    u = np.full((R, R), 2.0 / dt0, np.float32)
    v = np.full((R, R), -1.0 / dt0, np.float32)
    y = x.copy(); oracle_lib.advect(y, u, v, 0, dt, True)
    assert np.array_equal(y[4:8, 6:12], x[5:9, 4:10])
    assert abs(float(y[1:-1, 1:-1].sum()) - float(x.sum())) < 1e-5


def test_gather_zero_velocity(oracle_lib):
    """Zero velocity: identity except row 1 and column 1, which the [1.5, R-1.5]
    clamp averages with row 2 / column 2."""
    R = 9
    f = W.random_state(R, seed=7)[0]
    z = np.zeros((R, R), np.float32)
    y = f.copy(); oracle_lib.advect(y, z, z, 0, 0.1, False)
    assert np.array_equal(y[2:-1, 2:-1], f[2:-1, 2:-1])
    h = np.float32(0.5)
    assert np.array_equal(y[3, 1], h * f[3, 1] + h * f[3, 2])
    assert np.array_equal(y[1, 3], h * f[1, 3] + h * f[2, 3])


def test_project_zero_velocity_and_telescoping(oracle_lib):
    R = 12
    z = np.zeros((R, R), np.float32)
    a, b = z.copy(), z.copy(); p = oracle_lib.project(a, b, 20, GS, want_p=True)
    assert not a.any() and not b.any() and not p.any()
    st = W.vortex_state(66, seed=1, n_vortices=8)
    d = oracle_lib.divergence(st[1], st[2])
    # with the opposite rings the wide-stencil divergence telescopes to ~0
    assert abs(float(d.astype(np.float64).sum())) < 1e-6 * float(np.abs(d).sum())


@pytest.mark.parametrize("order", [GS, JAC])
def test_pressure_residual_decreases(order, oracle_lib):
    R = 34
    _, u, v, *_ = W.random_state(R, seed=8, vel_scale=1.0)
    d = oracle_lib.divergence(u, v)
    last = np.inf
    for K in (1, 4, 16, 64):
        a, b = u.copy(), v.copy()
        p = oracle_lib.project(a, b, K, order, want_p=True)
        r = oracle_lib.poisson_residual_f64(d, p)
        assert r < last
        last = r


def test_f64_project_tracks_f32(oracle_lib):
    R = 34
    _, u, v, *_ = W.random_state(R, seed=9, vel_scale=1.0)
    a, b = u.copy(), v.copy(); oracle_lib.project(a, b, 20, GS)
    ad, bd = u.astype(np.float64), v.astype(np.float64); oracle_lib.project_f64(ad, bd, 20, GS)
    assert np.abs(a - ad).max() < 1e-5 * np.abs(ad).max()
    n32 = oracle_lib.divergence_norm_f64(a, b)
    assert n32 < oracle_lib.divergence_norm_f64(u, v)


def test_divergence_norm_f64_closed_form(oracle_lib):
    """u = c * j (so du/dj = c): every interior divergence is -0.5*h*(2c)."""
    R = 10
    u = np.tile(np.arange(R, dtype=np.float32), (R, 1)) * np.float32(0.25)
    v = np.zeros((R, R), np.float32)
    h = 1.0 / R
    want = np.sqrt((R - 2) ** 2 * (0.5 * h * 0.5) ** 2)
    assert abs(oracle_lib.divergence_norm_f64(u, v) - want) < 1e-12


def test_reference_denies_nan_scatter_is_documented(oracle_lib):
    """NaN sources are skipped by the oracle (the reference's behaviour is
    undefined there); finite results elsewhere are unaffected."""
    R = 12
    st = W.random_state(R, seed=10, vel_scale=0.2)
    u = st[1].copy(); u[4, 4] = np.nan
    y = st[0].copy(); oracle_lib.advect(y, u, st[2], 0, 0.1, True)
    assert np.isfinite(y).all()


# ------------------------------------------------ visual output (renderers)
def test_render_density_known_answers(oracle_lib):
    """Pixels of src/density_grid_renderer.cu:17-26 worked by hand: channel =
    trunc(clamp(m * d, 0, 255)), a NaN product clamps to 255 (the device
    fminf/fmaxf return the non-NaN operand), alpha 255."""
    d = np.array([[0.5, 1.0, 2.0, -1.0], [np.nan, np.inf, -np.inf, -0.0],
                  [np.float32(1) / np.float32(255), 0.1, 0.25, 100.0], [0.0, 3.0, 0.125, 1e-30]], np.float32)
    px = oracle_lib.render_density(d, 255.0, 128.0, 1.0)
    want_r = [[127, 255, 255, 0], [255, 255, 0, 0], [1, 25, 63, 255], [0, 255, 31, 0]]
    want_g = [[64, 128, 255, 0], [255, 255, 0, 0], [0, 12, 32, 255], [0, 255, 16, 0]]
    want_b = [[0, 1, 2, 0], [255, 255, 0, 0], [0, 0, 0, 100], [0, 3, 0, 0]]
    assert np.array_equal(px[:, :, 0], want_r)
    assert np.array_equal(px[:, :, 1], want_g)
    assert np.array_equal(px[:, :, 2], want_b)
    assert (px[:, :, 3] == 255).all()


def test_render_velocity_known_answers(oracle_lib):
    """Lines of src/velocity_grid_renderer.cu:17-41 on a 64 x 64 grid, where
    sqrtf(rows * cols) = 64 exactly: with u = 0.5 and v = -0.25 every 8th row
    and column ends 250000 * 0.5 / 64 = 1953.125 right and 976.5625 up of its
    start; the others are points; all white, texture coordinates 0."""
    R = 64
    u = np.full((R, R), 0.5, np.float32)
    v = np.full((R, R), -0.25, np.float32)
    L = oracle_lib.render_velocity(u, v, 2.0, 3.0)
    i, j = np.meshgrid(np.arange(R), np.arange(R), indexing="ij")
    assert np.array_equal(L["start"]["x"], (2.0 * j).astype(np.float32))
    assert np.array_equal(L["start"]["y"], (3.0 * i).astype(np.float32))
    moved = (i % 8 == 0) & (j % 8 == 0)
    assert np.array_equal(L["end"]["x"], np.where(moved, 2.0 * j + 1953.125, 2.0 * j).astype(np.float32))
    assert np.array_equal(L["end"]["y"], np.where(moved, 3.0 * i - 976.5625, 3.0 * i).astype(np.float32))
    for end in ("start", "end"):
        assert (L[end]["color"] == 255).all() and (L[end]["tex"] == 0).all()
    # a zero velocity leaves every line a point
    Z = oracle_lib.render_velocity(np.zeros((R, R), np.float32), np.zeros((R, R), np.float32), 1.5, 0.5)
    assert np.array_equal(Z["start"]["x"], Z["end"]["x"]) and np.array_equal(Z["start"]["y"], Z["end"]["y"])


def test_render_velocity_length_and_direction(oracle_lib):
    """At the sampled cells the end offset is the velocity scaled by
    250000 / sqrt(rows * cols), within float rounding of the two operations."""
    R = 130
    rng = np.random.default_rng(3)
    u = rng.standard_normal((R, R)).astype(np.float32) * 1e-3
    v = rng.standard_normal((R, R)).astype(np.float32) * 1e-3
    L = oracle_lib.render_velocity(u, v, 1.0, 1.0)
    s = np.sqrt(np.float64(R) * R)
    dx = L["end"]["x"][::8, ::8].astype(np.float64) - L["start"]["x"][::8, ::8]
    want = 250000.0 * u[::8, ::8].astype(np.float64) / s
    assert np.allclose(dx, want, rtol=1e-5, atol=2e-4)


def test_jacobi_projection_is_matrix_iteration(oracle_lib):
    """The Jacobi-order projection: p_{k+1} = (d + N p_k) / 4 on the interior
    from p_0 = 0 with the continuous ring after every sweep (N sums the four
    neighbours), then u -= 0.5 (pE - pW) / h, v -= 0.5 (pS - pN) / h with the
    opposite rings -- replayed in float64 from the oracle's own divergence.  A
    Gauss-Seidel/Jacobi mix-up, a dropped ring or a wrong sign fails it."""
    R = 13
    _, u, v, *_ = W.random_state(R, seed=21, vel_scale=1.0)
    K = 4
    a, b = u.copy(), v.copy()
    p = oracle_lib.project(a, b, K, JAC, want_p=True)
    d = oracle_lib.divergence(u, v).astype(np.float64)
    q = np.zeros((R, R))
    for _ in range(K):
        nb = q[:-2, 1:-1] + q[2:, 1:-1] + q[1:-1, :-2] + q[1:-1, 2:]
        new = q.copy()
        new[1:-1, 1:-1] = (d[1:-1, 1:-1] + nb) / 4.0
        W.ring(new, 0)
        q = new
    assert np.abs(p - q).max() < 1e-6 * np.abs(q).max()
    h = 1.0 / R
    uu, vv = u.astype(np.float64), v.astype(np.float64)
    uu[1:-1, 1:-1] -= 0.5 * (q[1:-1, 2:] - q[1:-1, :-2]) / h
    vv[1:-1, 1:-1] -= 0.5 * (q[2:, 1:-1] - q[:-2, 1:-1]) / h
    W.ring(uu, 1)
    W.ring(vv, 2)
    assert np.abs(a - uu).max() < 1e-5 * np.abs(uu).max()
    assert np.abs(b - vv).max() < 1e-5 * np.abs(vv).max()
    pg = oracle_lib.project(u.copy(), v.copy(), K, GS, want_p=True)
    assert np.abs(pg - q).max() > 1e-3 * np.abs(q).max()  # the Gauss-Seidel order differs


def test_poisson_residual_closed_forms(oracle_lib):
    """The compact residual ||d - A p|| / ||d|| (A p = 4p - sum of neighbours,
    continuous ring): exactly 1 for p = 0 and for any constant p (the
    operator's null space); below 1e-5 for the exact solution (DCT, float64,
    rounded to float32); and a hand-worked 2 x 2 interior value."""
    from fs2d_amd import pressure_net

    R = 34
    _, u, v = W.vortex_state(R, seed=22, n_vortices=8)  # opposite rings: sum(d) telescopes to ~0
    d = oracle_lib.divergence(u, v)
    assert oracle_lib.poisson_residual_f64(d, np.zeros((R, R), np.float32)) == 1.0
    assert oracle_lib.poisson_residual_f64(d, np.full((R, R), 3.25, np.float32)) == 1.0
    exact = pressure_net.poisson_dct(d.astype(np.float64)[1:-1, 1:-1])
    p = np.zeros((R, R), np.float32)
    p[1:-1, 1:-1] = exact.astype(np.float32)
    W.ring(p, 0)
    assert oracle_lib.poisson_residual_f64(d, p) < 1e-5
    # R = 4: interior cells (1,1) (1,2) (2,1) (2,2); with the continuous ring an
    # interior neighbour outside the 2 x 2 block is the cell itself.
    d4 = np.zeros((4, 4), np.float32)
    d4[1:3, 1:3] = [[1.0, 2.0], [3.0, 4.0]]
    p4 = np.zeros((4, 4), np.float32)
    p4[1:3, 1:3] = [[1.0, 0.0], [0.0, 0.0]]
    W.ring(p4, 0)
    # A p at (1,1): 4*1 - (1 + 0 + 1 + 0) = 2; at (1,2) and (2,1): 0 - 1 = -1; at (2,2): 0
    r = np.array([1.0 - 2.0, 2.0 + 1.0, 3.0 + 1.0, 4.0])
    want = np.sqrt((r ** 2).sum() / 30.0)
    assert abs(oracle_lib.poisson_residual_f64(d4, p4) - want) < 1e-15
