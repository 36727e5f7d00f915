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


def compact_residual(d: np.ndarray, p: np.ndarray) -> float:
    """||d - L p|| / ||d|| with p's continuous ring (interior arrays)."""
    q = np.pad(p, 1, mode="edge")
    Lp = 4 * p - (q[:-2, 1:-1] + q[2:, 1:-1] + q[1:-1, :-2] + q[1:-1, 2:])
    return float(np.linalg.norm(d - Lp) / np.linalg.norm(d))


def make_pairs(R: int, seeds, steps, crop: int, crops_per_state: int, rng_seed: int = 0):
    """Training crops from seeded 64-vortex runs of the GPU solver (Jacobi
    order, K = 40): at each step in `steps` the divergence of the velocity is
    taken, its pressure solved exactly (poisson_dct), and random crop x crop
    windows of (divergence, previous step's pressure, pressure) kept."""
    rng = np.random.default_rng(rng_seed)
    n = R - 2
    D, PV, T = [], [], []
    for seed in seeds:
        st = W.vortex_state(R, seed=seed)
        with Solver(R, diffuse_iterations=40, pressure_iterations=40, order=JACOBI) as s:
            s.upload(*st)
            s.set_sources(None, None, None)
            done, prev = 0, None
            for k in sorted(set(steps) | {max(steps[0] - 1, 0)}):
                s.step_n(k - done, 1e-5, 1e-6, 0.05)
                done = k
                _, u, v = s.download()
                d = divergence(u, v)
                p = poisson_dct(d)
                if prev is not None and k in steps:
                    for _ in range(crops_per_state):
                        i, j = rng.integers(0, n - crop + 1, size=2)
                        D.append(d[i:i + crop, j:j + crop])
                        PV.append(prev[i:i + crop, j:j + crop])
                        T.append(p[i:i + crop, j:j + crop])
                prev = p
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
    the previous pressure (1) or the network (2); returns (sweeps, relative
    residual, fell back, seconds)."""
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
