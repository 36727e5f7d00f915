"""The spectral model behind the bound on a local seed for the pressure solve
(fs2d_amd.pressure_net: jacobi_residual_model, sweeps_to_reach,
local_seed_bound), pinned against Jacobi iterated directly (float64) and
against the oracle's Jacobi-order projection (src/fluid_solver_cpu.cpp:196-204
relaxed in Jacobi order, float32)."""
import numpy as np
import pytest

from fs2d_amd import pressure_net as PN
from fs2d_amd import workloads as W


def jacobi64(d, p0, k):
    """k sweeps of p <- (d + pN + pS + pW + pE) / 4 on the interior with the
    continuous ring (ring cell = its interior neighbour), float64."""
    p = p0.astype(np.float64).copy()
    for _ in range(k):
        q = np.pad(p, 1, mode="edge")
        p = 0.25 * (d + q[:-2, 1:-1] + q[2:, 1:-1] + q[1:-1, :-2] + q[1:-1, 2:])
    return p


def fields(n, seed):
    rng = np.random.default_rng(seed)
    d = rng.standard_normal((n, n))
    p0 = rng.standard_normal((n, n))
    return d, p0


@pytest.mark.parametrize("n,seed", [(8, 0), (17, 1), (32, 2)])
def test_model_equals_direct_iteration(n, seed):
    """Residual after k sweeps from an arbitrary start, by the DCT model and by
    iterating: equal to rounding (a wrong eigenvalue, DCT type, sign or the
    missing mean term would differ at once)."""
    d, p0 = fields(n, seed)
    ks = [0, 1, 2, 7, 40]
    model = PN.jacobi_residual_model(d, p0, ks)
    direct = [PN.compact_residual(d, jacobi64(d, p0, k)) for k in ks]
    np.testing.assert_allclose(model, direct, rtol=1e-10, atol=1e-12)


def test_model_matches_the_oracle_projection(oracle_lib):
    """The oracle's Jacobi-order pressure from zero after K sweeps (float32) has
    the modelled residual to float32 accuracy."""
    R, K = 66, 50
    _, u, v = W.vortex_state(R, seed=5, n_vortices=8)
    d32 = oracle_lib.divergence(u, v)
    p = oracle_lib.project(u.copy(), v.copy(), K, 1, want_p=True)
    d = d32.astype(np.float64)[1:-1, 1:-1]
    got = PN.compact_residual(d, p.astype(np.float64)[1:-1, 1:-1])
    want = float(PN.jacobi_residual_model(d, np.zeros_like(d), [K])[0])
    assert abs(got - want) < 1e-5 * max(want, 1.0), (got, want)


def test_sweeps_to_reach_is_the_brute_force_minimum():
    n = 12
    d, p0 = fields(n, 3)
    res = [PN.compact_residual(d, jacobi64(d, p0, k)) for k in range(200)]
    for target in (res[5] * 1.0000001, res[37] * 1.0000001, res[150] * 1.0000001):
        want = next(k for k, r in enumerate(res) if r <= target)
        assert PN.sweeps_to_reach(d, p0, target) == want
    assert PN.sweeps_to_reach(d, p0, res[0] * 1.0000001) == 0
    # below the mean's floor nothing reaches it
    floor = abs(d.mean()) * n / np.linalg.norm(d)
    assert PN.sweeps_to_reach(d, p0, 0.5 * floor) == -1


def test_local_seed_cutoff_direction():
    """An error made of one DCT mode of wavelength 2n/|(k,l)|: windows shorter
    than that wavelength leave it (the sweeps of the previous pressure),
    longer ones resolve it (the floor at once)."""
    from scipy.fft import idctn

    n = 64
    rng = np.random.default_rng(4)
    d = rng.standard_normal((n, n))
    exact = PN.poisson_dct(d)
    e = np.zeros((n, n))
    e[3, 4] = 50.0  # |(3, 4)| = 5: wavelength 2 * 64 / 5 = 25.6 cells
    p_prev = exact - idctn(e, type=2, norm="ortho")
    b = PN.local_seed_bound(d, p_prev, 300, wavelengths=(16, 40))
    assert b["sweeps_from_previous"] > 0
    assert b["windows"][16]["sweeps"] == b["sweeps_from_previous"]
    assert b["windows"][40]["sweeps"] == 0 and b["windows"][40]["speedup_bound"] is None
