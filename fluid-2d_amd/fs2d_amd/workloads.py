"""Synthetic inputs of the stable-fluids workloads, generated on the host.

All randomness comes from one self-contained generator (SplitMix64 with
Box-Muller in float64), so the arrays are identical on every machine and for
every consumer (HIP solver, oracle, golden fixtures).  Fields are float32
(R, R), row-major, ring included; R = n + 2 (src/grid.hpp:35-44).

Conventions (the reference has no harness; these are this package's):
  * i = row = y (grows downward on screen, src/grid_renderer.cpp:10),
    j = column = x; "upward" velocity is negative v.
  * Velocities are in the reference's units: one step moves a cell by
    dt0 * vel with dt0 = R * dt (src/fluid_solver_cpu.cpp:126), so a speed of
    "c cells per unit time" is vel = c / R.
  * Sources are passed to the solver as-is (no W*H scaling of
    src/simulation.cpp:45,49-50).
"""
from __future__ import annotations

import math

import numpy as np

_M64 = (1 << 64) - 1


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014) with float64 uniform / normal draws."""

    def __init__(self, seed: int):
        self.state = seed & _M64

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """float64 in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def normal(self) -> float:
        """Standard normal by Box-Muller (one of the pair)."""
        u1 = 1.0 - self.uniform()  # (0, 1]
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)

    def uniform_array(self, n: int) -> np.ndarray:
        """n float64 uniforms, vectorised SplitMix64 (same stream as uniform())."""
        idx = np.arange(1, n + 1, dtype=np.uint64)
        with np.errstate(over="ignore"):
            z = np.uint64(self.state) + idx * np.uint64(0x9E3779B97F4A7C15)
            z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
            z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
            z = z ^ (z >> np.uint64(31))
            self.state = int((np.uint64(self.state) + np.uint64(n) * np.uint64(0x9E3779B97F4A7C15)))
        self.state &= _M64
        return (z >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def stream_seed(config_seed: int, sim_index: int = 0) -> int:
    """Seed of simulation ``sim_index`` of a configuration."""
    return SplitMix64((config_seed * 0x100000001B3 + sim_index) & _M64).next_u64()


# --- boundary rings (the closed forms of src/fluid_solver_cpu.cpp:33-83) ---
def ring(x: np.ndarray, mode: int) -> np.ndarray:
    """Apply the reference's ring rule in place (0 continuous, 1 u, 2 v)."""
    tb = -1.0 if mode == 2 else 1.0
    lr = -1.0 if mode == 1 else 1.0
    x[0, 1:-1] = tb * x[1, 1:-1]
    x[-1, 1:-1] = tb * x[-2, 1:-1]
    x[1:-1, 0] = lr * x[1:-1, 1]
    x[1:-1, -1] = lr * x[1:-1, -2]
    f = np.float32(0.5)
    x[0, 0] = f * (np.float32(tb) * x[1, 1] + np.float32(lr) * x[1, 1])
    x[0, -1] = f * (np.float32(tb) * x[1, -2] + np.float32(lr) * x[1, -2])
    x[-1, 0] = f * (np.float32(tb) * x[-2, 1] + np.float32(lr) * x[-2, 1])
    x[-1, -1] = f * (np.float32(tb) * x[-2, -2] + np.float32(lr) * x[-2, -2])
    return x


# --- configuration 1: zero state, disc sources --------------------------------
def disc_sources(R: int, radius: float = 8.0, rho_value: float = 100.0,
                 v_value: float = -5.0):
    """rho_src = rho_value and v_src = v_value on interior cells within
    ``radius`` of the grid centre ((R-1)/2, (R-1)/2); u_src = 0."""
    c = (R - 1) / 2.0
    i = np.arange(R, dtype=np.float64)[:, None]
    j = np.arange(R, dtype=np.float64)[None, :]
    disc = (i - c) ** 2 + (j - c) ** 2 <= radius * radius
    disc[0, :] = disc[-1, :] = disc[:, 0] = disc[:, -1] = False
    rho_src = np.where(disc, np.float32(rho_value), np.float32(0)).astype(np.float32)
    v_src = np.where(disc, np.float32(v_value), np.float32(0)).astype(np.float32)
    u_src = np.zeros((R, R), np.float32)
    return rho_src, u_src, v_src


def zero_state(R: int):
    z = lambda: np.zeros((R, R), np.float32)
    return z(), z(), z()


# --- configurations 2-4: Gaussian vortices + smoothed noise density ----------
def vortex_velocity(R: int, n_vortices: int, rng: SplitMix64, max_speed_cells: float = 2.0,
                    radius_range=(10.0, 40.0), radius_scale: float = 1.0,
                    cutoff: float = 6.0):
    """Divergence-free velocity from psi = sum_k G_k exp(-r^2 / (2 r_k^2)).

    u = d(psi)/d(i), v = -d(psi)/d(j) (analytic, float64), centres uniform over
    the interior, r_k ~ U[radius_range] * radius_scale cells, G_k ~ N(0, 1);
    each vortex is evaluated within cutoff * r_k of its centre.  The field is
    scaled so the maximum interior speed is max_speed_cells cells per unit
    time (vel = cells / R) and the opposite rings are applied."""
    n = R - 2
    u = np.zeros((R, R), np.float64)
    v = np.zeros((R, R), np.float64)
    params = []
    for _ in range(n_vortices):
        ci = 1.0 + rng.uniform() * (n - 1)
        cj = 1.0 + rng.uniform() * (n - 1)
        rk = (radius_range[0] + rng.uniform() * (radius_range[1] - radius_range[0])) * radius_scale
        g = rng.normal()
        params.append((ci, cj, rk, g))
    for ci, cj, rk, g in params:
        w = cutoff * rk
        i0, i1 = max(1, int(math.floor(ci - w))), min(R - 2, int(math.ceil(ci + w)))
        j0, j1 = max(1, int(math.floor(cj - w))), min(R - 2, int(math.ceil(cj + w)))
        di = np.arange(i0, i1 + 1, dtype=np.float64)[:, None] - ci
        dj = np.arange(j0, j1 + 1, dtype=np.float64)[None, :] - cj
        e = g * np.exp(-(di * di + dj * dj) / (2.0 * rk * rk)) / (rk * rk)
        u[i0:i1 + 1, j0:j1 + 1] += -di * e
        v[i0:i1 + 1, j0:j1 + 1] += dj * e
    speed = np.sqrt(u[1:-1, 1:-1] ** 2 + v[1:-1, 1:-1] ** 2).max()
    scale = (max_speed_cells / R) / speed if speed > 0 else 0.0
    u32 = (u * scale).astype(np.float32)
    v32 = (v * scale).astype(np.float32)
    return ring(u32, 1), ring(v32, 2)


def smoothed_noise_density(R: int, rng: SplitMix64, sigma: float = 4.0):
    """U[0,1] noise on the interior, separable Gaussian blur (sigma cells,
    truncated at 4 sigma, reflecting edges), continuous ring."""
    from scipy.ndimage import gaussian_filter

    n = R - 2
    noise = rng.uniform_array(n * n).reshape(n, n)
    sm = gaussian_filter(noise, sigma=sigma, mode="reflect", truncate=4.0)
    rho = np.zeros((R, R), np.float32)
    rho[1:-1, 1:-1] = sm.astype(np.float32)
    return ring(rho, 0)


def vortex_state(R: int, seed: int, n_vortices: int = 64, sim_index: int = 0,
                 reference_n: int = 1024):
    """Seeded (rho, u, v) of the vortex configurations; radii are scaled by
    n / reference_n so the flow looks alike at every grid size."""
    rng = SplitMix64(stream_seed(seed, sim_index))
    u, v = vortex_velocity(R, n_vortices, rng, radius_scale=(R - 2) / reference_n)
    rho = smoothed_noise_density(R, rng)
    return rho, u, v


def random_state(R: int, seed: int, vel_scale: float = 0.05):
    """Unstructured random fields (rings included) for stress tests."""
    rng = SplitMix64(stream_seed(seed, 7))
    a = rng.uniform_array(6 * R * R).reshape(6, R, R)
    rho = a[0].astype(np.float32)
    u = ((a[1] - 0.5) * 2 * vel_scale).astype(np.float32)
    v = ((a[2] - 0.5) * 2 * vel_scale).astype(np.float32)
    rs = a[3].astype(np.float32)
    us = (a[4] - 0.5).astype(np.float32)
    vs = (a[5] - 0.5).astype(np.float32)
    return rho, u, v, rs, us, vs


BATCH_SIDES = (64, 96, 128, 192, 256, 384, 512)


def batch_config(count: int, seed: int, sides=BATCH_SIDES, iterations=(20, 40), dt_range=(0.02, 0.1)):
    """The batched workload's simulations: side n uniform over `sides`, dt
    uniform over dt_range, Jacobi count uniform over `iterations`, each drawn
    from the simulation's own stream.  Returns [(R, K, dt)]."""
    out = []
    for s in range(count):
        rng = SplitMix64(stream_seed(seed, 1000 + s))
        n = sides[min(int(rng.uniform() * len(sides)), len(sides) - 1)]
        dt = dt_range[0] + rng.uniform() * (dt_range[1] - dt_range[0])
        K = iterations[min(int(rng.uniform() * len(iterations)), len(iterations) - 1)]
        out.append((n + 2, K, float(np.float32(dt))))
    return out


def batch_state(R: int, seed: int, sim_index: int):
    """A batched simulation's own seeded 16-vortex start (radii 10-40 cells
    scaled by n / 256)."""
    return vortex_state(R, seed, n_vortices=16, sim_index=sim_index, reference_n=256)
