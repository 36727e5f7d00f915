"""The learned pressure initial guess (include/fs2d_cnn.h), from Python.

``PressureNet`` wraps an fs2d_net: training steps, prediction, parameters.
``poisson_dct`` gives the converged pressure of the reference's pressure
equation (src/fluid_solver_cpu.cpp:196-204: 4 p - (pN + pS + pW + pE) = d on
the interior, continuous ring) exactly, by the DCT-II that diagonalises it;
it labels the training pairs.  ``make_pairs`` runs the GPU solver to collect
(divergence, previous converged pressure, converged pressure) crops, and
``seeded_projection`` measures a projection seeded from zero, from the
previous pressure, or from the network.  The network itself runs only in the
HIP library: there is no CPU path.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import JACOBI, Fs2dError, Solver, library
from . import workloads as W

_i, _f, _p = ctypes.c_int, ctypes.c_float, ctypes.c_void_p

# parameters written by python -m fs2d_amd.train_pressure_net
DEFAULT_PARAMS = __import__("os").path.join(__import__("os").path.dirname(__import__("os").path.abspath(__file__)),
                                            "pressure_net_params.npz")

LAYERS = [(16, 16), (16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (64, 128), (128, 128),
          (192, 64), (64, 64), (96, 32), (32, 32), (48, 16), (16, 16), (16, 16)]  # (cin, cout), padded


def _bind():
    L = library()
    if getattr(L, "_fs2d_net_bound", False):
        return L
    sig = {
        "fs2d_net_create": (_i, [ctypes.c_uint64, ctypes.POINTER(_p)]),
        "fs2d_net_destroy": (_i, [_p]),
        "fs2d_net_param_count": (_i, [_p, ctypes.POINTER(ctypes.c_int64)]),
        "fs2d_net_get_params": (_i, [_p, _p]),
        "fs2d_net_set_params": (_i, [_p, _p]),
        "fs2d_net_get_grads": (_i, [_p, _p]),
        "fs2d_net_train_step": (_i, [_p, _i, _i, _i, _p, _p, _p, _f, ctypes.POINTER(_f)]),
        "fs2d_net_predict_host": (_i, [_p, _i, _i, _i, _p, _p, _p]),
        "fs2d_net_stream": (_p, [_p]),
        "fs2d_net_time_forward": (_i, [_p, _i, _i, _i, _i, ctypes.POINTER(_f)]),
        "fs2d_net_forward_flops": (_i, [_p, _i, _i, _i, ctypes.POINTER(ctypes.c_double)]),
        "fs2d_net_last_error": (ctypes.c_char_p, [_p]),
        "fs2d_hip_project_seeded": (_i, [_p, _p, _i, _f, _i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_double),
                                         ctypes.POINTER(_i)]),
        "fs2d_hip_pressure_residual": (_i, [_p, ctypes.POINTER(ctypes.c_double)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    L._fs2d_net_bound = True
    return L


def _f32(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


class PressureNet:
    """An fs2d_net: 4-level 16/32/64/128 U-Net, BF16 MFMA convolutions."""

    def __init__(self, seed: int = 0):
        self._L = _bind()
        h = _p()
        rc = self._L.fs2d_net_create(seed, ctypes.byref(h))
        if rc != 0:
            raise Fs2dError(f"fs2d_net_create failed ({rc}): {self._L.fs2d_net_last_error(None).decode()}")
        self._h = h
        n = ctypes.c_int64()
        self._ck(self._L.fs2d_net_param_count(h, ctypes.byref(n)))
        self.param_count = int(n.value)

    def _ck(self, rc: int) -> None:
        if rc != 0:
            raise Fs2dError(f"fs2d_net error {rc}: {self._L.fs2d_net_last_error(self._h).decode()}")

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            self._L.fs2d_net_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def params(self) -> np.ndarray:
        out = np.empty(self.param_count, np.float32)
        self._ck(self._L.fs2d_net_get_params(self._h, out.ctypes.data))
        return out

    def set_params(self, p: np.ndarray) -> None:
        p = _f32(p, (self.param_count,))
        self._ck(self._L.fs2d_net_set_params(self._h, p.ctypes.data))

    def grads(self) -> np.ndarray:
        out = np.empty(self.param_count, np.float32)
        self._ck(self._L.fs2d_net_get_grads(self._h, out.ctypes.data))
        return out

    def train_step(self, d, prev, target, lr: float) -> float:
        n, h, w = np.shape(d)
        d, prev, target = (_f32(a, (n, h, w)) for a in (d, prev, target))
        loss = _f()
        self._ck(self._L.fs2d_net_train_step(self._h, n, h, w, d.ctypes.data, prev.ctypes.data, target.ctypes.data,
                                             lr, ctypes.byref(loss)))
        return float(loss.value)

    def predict(self, d, prev) -> np.ndarray:
        n, h, w = np.shape(d)
        d, prev = _f32(d, (n, h, w)), _f32(prev, (n, h, w))
        out = np.empty((n, h, w), np.float32)
        self._ck(self._L.fs2d_net_predict_host(self._h, n, h, w, d.ctypes.data, prev.ctypes.data, out.ctypes.data))
        return out

    def time_forward(self, n: int, h: int, w: int, reps: int = 10) -> float:
        """Milliseconds per forward pass (HIP events); runs a prediction of
        that shape first so its activations exist."""
        z = np.zeros((n, h, w), np.float32)
        self.predict(z + 1e-3, z + 1e-3)
        ms = _f()
        self._ck(self._L.fs2d_net_time_forward(self._h, n, h, w, reps, ctypes.byref(ms)))
        return float(ms.value)

    def forward_flops(self, n: int, h: int, w: int) -> float:
        f = ctypes.c_double()
        self._ck(self._L.fs2d_net_forward_flops(self._h, n, h, w, ctypes.byref(f)))
        return float(f.value)

    def save(self, path: str) -> None:
        np.savez(path, params=self.params())

    def load(self, path: str) -> None:
        with np.load(path, allow_pickle=False) as z:
            self.set_params(z["params"])


def split_params(p: np.ndarray):
    """[(weights [9][cout][cin], bias [cout])] per layer, from the flat vector."""
    out, off = [], 0
    for cin, cout in LAYERS:
        wn = 9 * cout * cin
        out.append((p[off:off + wn].reshape(9, cout, cin), p[off + wn:off + wn + cout]))
        off += wn + cout
    return out


# ---------------------------------------------------------------- labels
def divergence(u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """The reference's divergence (src/fluid_solver_cpu.cpp:188-193) of the
    interior, FP64: d = (-0.5 h) (uE - uW + vS - vN), h = 1 / R."""
    R = u.shape[0]
    h = 1.0 / np.sqrt(float(R * R))
    u64, v64 = u.astype(np.float64), v.astype(np.float64)
    return (-0.5 * h) * (u64[1:-1, 2:] - u64[1:-1, :-2] + v64[2:, 1:-1] - v64[:-2, 1:-1])


def poisson_dct(d: np.ndarray) -> np.ndarray:
    """Interior p with 4 p - (pN + pS + pW + pE) = d - mean(d), ring = adjacent
    interior (Neumann), mean(p) = 0: the DCT-II diagonalises the operator with
    eigenvalues (2 - 2 cos(pi k / n)) + (2 - 2 cos(pi l / n))."""
    from scipy.fft import dctn, idctn

    n = d.shape[0]
    dh = dctn(d - d.mean(), type=2, norm="ortho")
    lam = 2.0 - 2.0 * np.cos(np.pi * np.arange(n) / n)
    den = lam[:, None] + lam[None, :]
    den[0, 0] = 1.0
    ph = dh / den
    ph[0, 0] = 0.0
    return idctn(ph, type=2, norm="ortho")


def jacobi_spectrum(n: int):
    """Eigenvalues on the DCT-II modes (k, l) of the pressure equation with the
    continuous ring: Jacobi's iteration matrix (pN + pS + pW + pE) / 4 has
    mu = (cos(pi k / n) + cos(pi l / n)) / 2 and the operator L = 4 - 4 mu."""
    c = np.cos(np.pi * np.arange(n) / n)
    mu = 0.5 * (c[:, None] + c[None, :])
    return mu, 4.0 - 4.0 * mu


def jacobi_residual_model(d: np.ndarray, p0: np.ndarray, sweeps, keep: np.ndarray | None = None):
    """Relative compact residuals ||d - L p_k|| / ||d|| of Jacobi from p0 after
    each count in `sweeps`, exactly (float64, no rounding): the error
    e = p* - p0 (p* = poisson_dct(d)) decays as mu^k per DCT mode and the
    residual is L e plus d's mean, which no p can cancel (L p sums to zero
    under the continuous ring).  `keep` (bool per mode) restricts the starting
    error to those modes: the start then equals p* in every other mode."""
    from scipy.fft import dctn

    n = d.shape[0]
    mu, lam = jacobi_spectrum(n)
    eh = dctn(poisson_dct(d) - p0, type=2, norm="ortho")
    if keep is not None:
        eh = np.where(keep, eh, 0.0)
    a = (lam * eh) ** 2
    a[0, 0] = 0.0
    mean2 = float(n * n) * float(np.mean(d)) ** 2
    m2 = mu * mu
    nd = float(np.linalg.norm(d))
    out = []
    for k in np.atleast_1d(sweeps):
        out.append(np.sqrt(mean2 + float(np.sum(a * m2 ** int(k)))) / nd)
    return np.asarray(out)


def sweeps_to_reach(d: np.ndarray, p0: np.ndarray, target: float, keep: np.ndarray | None = None,
                    limit: int = 1 << 22) -> int:
    """Fewest Jacobi sweeps from p0 whose modelled relative residual is at most
    `target` (the residual never grows: |mu| <= 1), or -1 past `limit`."""
    from scipy.fft import dctn

    n = d.shape[0]
    mu, lam = jacobi_spectrum(n)
    eh = dctn(poisson_dct(d) - p0, type=2, norm="ortho")
    if keep is not None:
        eh = np.where(keep, eh, 0.0)
    a = ((lam * eh) ** 2).ravel()
    a[0] = 0.0
    nz = a > 0
    a, lm = a[nz], np.log(np.maximum((mu * mu).ravel()[nz], 1e-300))
    budget = (target * float(np.linalg.norm(d))) ** 2 - float(n * n) * float(np.mean(d)) ** 2

    def ok(k: int) -> bool:
        return float(np.sum(a * np.exp(lm * k))) <= budget

    if budget < 0:
        return -1
    if ok(0):
        return 0
    hi = 1
    while not ok(hi):
        hi *= 2
        if hi > limit:
            return -1
    lo = hi // 2
    while hi - lo > 1:
        mid = (lo + hi) // 2
        if ok(mid):
            hi = mid
        else:
            lo = mid
    return hi


# Receptive field of the network in cells: two 3x3 convolutions per level
# (16/32/64/128 channels), 2 x 2 pools down, bilinear x2 up, two convolutions per
# decoder level and the output convolution give each output cell a 112 x 112
# window of its inputs.
RECEPTIVE_FIELD = 112


def local_seed_bound(d: np.ndarray, p_prev: np.ndarray, budget: int, wavelengths=(RECEPTIVE_FIELD,)):
    """What any seed computed from a window of the inputs can gain over plain
    Jacobi from the previous pressure, on one projection.  The target is the
    residual Jacobi reaches from p_prev in `budget` sweeps; for each window
    side W, the idealised local seed equals the exact pressure p* in every
    DCT mode of wavelength 2 n / |(k, l)| shorter than W and p_prev in the
    others (the modes no window of side W resolves), and the result is the
    sweeps it still needs to reach the target.  Returns a dict of the target,
    the sweeps from p_prev and, per W, the sweeps and their ratio (the best
    speedup in sweeps, None where the seed alone meets the target; a
    network's own forward time only lowers it)."""
    n = d.shape[0]
    target = float(jacobi_residual_model(d, p_prev, [budget])[0])
    base = sweeps_to_reach(d, p_prev, target)
    k = np.arange(n, dtype=np.float64)
    rad = np.hypot(k[:, None], k[None, :])
    out = {"n": n, "budget_sweeps": budget, "target_rel_residual": target, "sweeps_from_previous": base,
           "windows": {}}
    for w in wavelengths:
        keep = rad <= 2.0 * n / float(w)  # wavelength >= w: unresolved, left at p_prev's error
        sw = sweeps_to_reach(d, p_prev, target, keep=keep)
        out["windows"][int(w)] = {"sweeps": sw, "speedup_bound": (base / sw) if sw > 0 else None}
    return out


def compact_residual(d: np.ndarray, p: np.ndarray) -> float:
    """||d - L p|| / ||d|| with p's continuous ring (interior arrays)."""
    q = np.pad(p, 1, mode="edge")
    Lp = 4 * p - (q[:-2, 1:-1] + q[2:, 1:-1] + q[1:-1, :-2] + q[1:-1, 2:])
    return float(np.linalg.norm(d - Lp) / np.linalg.norm(d))


def projection_inputs(solver: Solver, k_from: int, k_to: int, K: int, diff: float, visc: float, dt: float):
    """Advance the solver (Jacobi order, no sources) from step k_from to k_to
    stage by stage, as fluid_solver_cpu::solve orders them
    (src/fluid_solver_cpu.cpp:16-30), yielding (step, projection, d) with d
    the divergence of the velocity each projection receives."""
    for k in range(k_from + 1, k_to + 1):
        solver.diffuse(0, diff, dt, K)
        solver.advect_density(dt)
        solver.diffuse(1, visc, dt, K)
        solver.diffuse(2, visc, dt, K)
        _, u, v = solver.download()
        yield k, 1, divergence(u, v)
        solver.project(K)
        solver.advect_velocity(dt)
        _, u, v = solver.download()
        yield k, 2, divergence(u, v)
        solver.project(K)


def make_pairs(R: int, seeds, steps, crop: int, crops_per_pair: int, K: int = 40, solve_sweeps: int = 20000,
               rng_seed: int = 0, log=None):
    """Training crops from seeded 64-vortex runs of the GPU solver (Jacobi
    order, K sweeps, dt 0.05, diffusion 1e-5, viscosity 1e-6).  Between
    samples the solver takes plain steps; at each sampled step k, steps k-1
    and k run stage by stage with every projection solved for `solve_sweeps`
    sweeps (from zero, then from the previous pressure), as a seeded solve to
    a tight tolerance would leave them.  A pair is (divergence entering a
    projection of step k, the handle's pressure from the projection before
    it, the exact pressure of that divergence from poisson_dct), and
    `crops_per_pair` random crop x crop windows of it are kept."""
    rng = np.random.default_rng(rng_seed)
    n = R - 2
    D, PV, T = [], [], []
    for seed in seeds:
        st = W.vortex_state(R, seed=seed)
        with Solver(R, diffuse_iterations=K, pressure_iterations=K, order=JACOBI) as s:
            s.upload(*st)
            s.set_sources(None, None, None)
            done = 0
            for k in steps:
                if k - 2 > done:
                    s.step_n(k - 2 - done, 1e-5, 1e-6, 0.05)
                for kk in (k - 1, k):
                    for proj in (1, 2):
                        if proj == 1:
                            s.diffuse(0, 1e-5, 0.05, K)
                            s.advect_density(0.05)
                            s.diffuse(1, 1e-6, 0.05, K)
                            s.diffuse(2, 1e-6, 0.05, K)
                        else:
                            s.advect_velocity(0.05)
                        if kk == k:
                            _, u, v = s.download()
                            d = divergence(u, v)
                            prev = s.pressure()[1:-1, 1:-1].astype(np.float64)
                            p = poisson_dct(d)
                            for _ in range(crops_per_pair):
                                i, j = rng.integers(0, n - crop + 1, size=2)
                                D.append(d[i:i + crop, j:j + crop].astype(np.float32))
                                PV.append(prev[i:i + crop, j:j + crop].astype(np.float32))
                                T.append(p[i:i + crop, j:j + crop].astype(np.float32))
                        seeded_projection(s, None, 0 if (kk == k - 1 and proj == 1) else 1, 0.0, solve_sweeps,
                                          solve_sweeps)
                done = k
        if log:
            log(f"seed {seed}: {len(D)} crops")
    return (np.asarray(D, np.float32), np.asarray(PV, np.float32), np.asarray(T, np.float32))


def train(net: PressureNet, data, epochs: int, batch: int = 16, lr: float = 1e-3, seed: int = 0, log=None):
    """Adam over shuffled batches; returns the per-epoch mean losses."""
    D, PV, T = data
    rng = np.random.default_rng(seed)
    losses = []
    for ep in range(epochs):
        order = rng.permutation(len(D))
        tot, cnt = 0.0, 0
        for b0 in range(0, len(order) - batch + 1, batch):
            idx = order[b0:b0 + batch]
            tot += net.train_step(D[idx], PV[idx], T[idx], lr)
            cnt += 1
        losses.append(tot / max(cnt, 1))
        if log:
            log(f"epoch {ep}: loss {losses[-1]:.4e}")
    return losses


def seeded_projection(solver: Solver, net, seed: int, tol: float, max_sweeps: int, check: int = 64,
                      fallback: bool = True):
    """One projection of the solver's current velocity seeded from zero (0),
    the previous pressure (1), the network (2) or multigrid V-cycles (3);
    returns (sweeps, relative residual, fell back, seconds)."""
    L = _bind()
    sw, fb = _i(), _i()
    rel = ctypes.c_double()
    solver.sync()
    t0 = time.perf_counter()
    rc = L.fs2d_hip_project_seeded(solver._h, net.handle if net is not None else None, seed, tol, max_sweeps, check,
                                   1 if fallback else 0, ctypes.byref(sw), ctypes.byref(rel), ctypes.byref(fb))
    sec = time.perf_counter() - t0
    if rc != 0:
        raise Fs2dError(f"fs2d_hip_project_seeded failed ({rc}): {L.fs2d_hip_last_error(solver._h).decode()}")
    return int(sw.value), float(rel.value), bool(fb.value), sec


def pressure_residual(solver: Solver) -> float:
    L = _bind()
    rel = ctypes.c_double()
    rc = L.fs2d_hip_pressure_residual(solver._h, ctypes.byref(rel))
    if rc != 0:
        raise Fs2dError(f"fs2d_hip_pressure_residual failed ({rc}): {L.fs2d_hip_last_error(solver._h).decode()}")
    return float(rel.value)


def projection_race(R: int, net, seed: int, start_step: int, steps: int, tol: float = 1e-4, max_sweeps: int = 400000,
                    check: int = 256, K: int = 40, methods=(1, 2)):
    """Config-5 comparison: from the seeded 64-vortex state at `start_step`
    (plain steps), one warm-up step whose projections are solved from zero to
    `tol`, then `steps` steps whose every projection is solved to `tol` from
    the previous pressure (method 1) or from the network's guess (method 2,
    with the fallback).  Each method runs on its own handle from the same
    state.  Returns {method: {"sweeps", "seconds", "fallbacks", "projections",
    "max_rel"}} with per-projection means."""
    st = W.vortex_state(R, seed=seed)
    out = {}
    for m in methods:
        with Solver(R, diffuse_iterations=K, pressure_iterations=K, order=JACOBI) as s:
            s.upload(*st)
            s.set_sources(None, None, None)
            s.step_n(start_step, 1e-5, 1e-6, 0.05)
            rec = {"sweeps": 0, "seconds": 0.0, "fallbacks": 0, "projections": 0, "max_rel": 0.0}
            for k in range(steps + 1):
                seed_k = 0 if k == 0 else m
                for proj in (1, 2):
                    if proj == 1:
                        s.diffuse(0, 1e-5, 0.05, K)
                        s.advect_density(0.05)
                        s.diffuse(1, 1e-6, 0.05, K)
                        s.diffuse(2, 1e-6, 0.05, K)
                    else:
                        s.advect_velocity(0.05)
                    sw, rel, fb, sec = seeded_projection(s, net if seed_k == 2 else None, seed_k, tol, max_sweeps,
                                                         check)
                    if k > 0:
                        rec["sweeps"] += sw
                        rec["seconds"] += sec
                        rec["fallbacks"] += int(fb)
                        rec["projections"] += 1
                        rec["max_rel"] = max(rec["max_rel"], rel)
            n = max(rec["projections"], 1)
            out[m] = {"sweeps": rec["sweeps"] / n, "seconds": rec["seconds"] / n, "fallbacks": rec["fallbacks"],
                      "projections": rec["projections"], "max_rel": rec["max_rel"]}
    return out


def equal_residual_race(R: int, net, seed: int = 7, start_step: int = 100, budget: int = 16384,
                        warm_sweeps: int = 20000, K: int = 40, check: int = 1024,
                        windows=(RECEPTIVE_FIELD, 2 * RECEPTIVE_FIELD, 4 * RECEPTIVE_FIELD, 8 * RECEPTIVE_FIELD)):
    """The first projection of step start_step + 2 of a seeded 64-vortex flow,
    after one step whose projections were solved for warm_sweeps sweeps (so
    the previous pressure is a converged one), from identical states: the
    relative residual plain Jacobi reaches in `budget` sweeps from zero and
    from the previous pressure, and the sweeps and time the network's seed
    needs to reach the previous-pressure start's residual.  Every run checks
    the residual every `check` sweeps (a host synchronisation each).  Also the
    bound of local_seed_bound on the same projection for each window side in
    `windows`, and the same ideal local correction applied to the pressure
    of one multigrid V-cycle instead (how many V-cycles it could replace)."""
    st = W.vortex_state(R, seed=seed)

    def state():
        s = Solver(R, diffuse_iterations=K, pressure_iterations=K, order=JACOBI)
        s.upload(*st)
        s.set_sources(None, None, None)
        s.step_n(start_step, 1e-5, 1e-6, 0.05)
        for first in (True, False):
            if first:
                s.diffuse(0, 1e-5, 0.05, K); s.advect_density(0.05); s.diffuse(1, 1e-6, 0.05, K)
                s.diffuse(2, 1e-6, 0.05, K)
            else:
                s.advect_velocity(0.05)
            seeded_projection(s, None, 0 if first else 1, 0.0, warm_sweeps, warm_sweeps)
        s.diffuse(0, 1e-5, 0.05, K); s.advect_density(0.05); s.diffuse(1, 1e-6, 0.05, K); s.diffuse(2, 1e-6, 0.05, K)
        return s

    out = {}
    for name, sd in (("zero", 0), ("previous_pressure", 1)):
        with state() as s:
            sw, rel, _, sec = seeded_projection(s, None, sd, 0.0, budget, check)
        out[name] = {"sweeps": sw, "rel_residual": rel, "ms": sec * 1e3}
    target = out["previous_pressure"]["rel_residual"]
    with state() as s:
        sw, rel, fb, sec = seeded_projection(s, net, 2, target, 4 * budget, check, fallback=False)
    out["network"] = {"sweeps": sw, "rel_residual": rel, "ms": sec * 1e3, "reached": rel <= target}
    with state() as s:
        sw, rel, fb, sec = seeded_projection(s, None, 3, target, 4 * budget, check, fallback=False)
    out["multigrid"] = {"sweeps": sw, "rel_residual": rel, "ms": sec * 1e3, "reached": rel <= target}
    # V-cycles alone towards 1e-4: float32 iterates floor the compact residual
    # (~1e-5 at 256^2, ~9e-5 at 1024^2, growing as n^2), which this reports
    with state() as s:
        sw, rel, fb, sec = seeded_projection(s, None, 3, 1e-4, 0, check, fallback=False)
    out["multigrid_to_1e-4"] = {"sweeps": sw, "rel_residual": rel, "ms": sec * 1e3, "reached": rel <= 1e-4}
    out["target_rel_residual"] = target
    # the same projection in the DCT model: its residuals from zero and from the
    # previous pressure after the budget (to set beside the measured ones), and
    # what an ideal seed resolving every wavelength below a window side gains
    with state() as s:
        _, u, v = s.download()
        d = divergence(u, v)
        prev = s.pressure()[1:-1, 1:-1].astype(np.float64)
    bound = local_seed_bound(d, prev, budget, wavelengths=windows)
    bound["model_from_zero"] = float(jacobi_residual_model(d, np.zeros_like(d), [budget])[0])
    out["local_seed_bound"] = bound
    # a local network refining the multigrid seed: one V-cycle's pressure, the
    # residual it leaves, what an ideal correction of every wavelength below W
    # would leave, and the measured factor of a second V-cycle -- so the
    # V-cycles such a correction could replace at most
    with state() as s:
        _, u, v = s.download()
        d = divergence(u, v)
        _, rel1, _, sec1 = seeded_projection(s, None, 3, target, 0, check, fallback=False)
        p_mg = s.pressure()[1:-1, 1:-1].astype(np.float64)
    with state() as s:
        _, rel2, _, sec2 = seeded_projection(s, None, 3, 0.5 * rel1, 0, check, fallback=False)
    n = d.shape[0]
    k = np.arange(n, dtype=np.float64)
    rad = np.hypot(k[:, None], k[None, :])
    r_mg = float(jacobi_residual_model(d, p_mg, [0])[0])
    refine = {"measured_one_cycle": rel1, "model_one_cycle": r_mg, "measured_two_cycles": rel2,
              "cycle_factor": rel1 / rel2 if rel2 > 0 else None, "ms_one_cycle": sec1 * 1e3,
              "ms_two_cycles": sec2 * 1e3, "windows": {}}
    for w in windows:
        r_w = float(jacobi_residual_model(d, p_mg, [0], keep=rad <= 2.0 * n / float(w))[0])
        saved = (np.log(r_mg / r_w) / np.log(rel1 / rel2)) if (rel2 > 0 and rel1 > rel2 and r_w > 0) else None
        refine["windows"][int(w)] = {"residual_after_ideal_correction": r_w, "v_cycles_replaced_at_most": saved}
    out["multigrid_local_refinement"] = refine
    return out


if __name__ == "__main__":
    # python -m fs2d_amd.pressure_net [side ...]: the trained network's forward
    # time (HIP events, 20 passes) and FLOPs at each square side (default 1024)
    import json
    import sys

    net = PressureNet(seed=0)
    net.load(DEFAULT_PARAMS)
    for side in (int(a) for a in (sys.argv[1:] or ["1024"])):
        ms = net.time_forward(1, side, side, reps=20)
        print(json.dumps({"side": side, "forward_ms": round(ms, 4), "flops": net.forward_flops(1, side, side)}),
              flush=True)
    net.close()
