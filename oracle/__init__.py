"""CPU oracle for the stable-fluids step -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this package, and only as the checker; the HIP product path never calls
it and has no CPU fallback.

Two libraries are bound here:

* ``libfs2d_oracle.so`` (``oracle/fs2d_oracle.c``): a C restatement of
  ``src/fluid_solver_cpu.cpp`` with the reference's evaluation order, iteration
  counts as parameters and a Jacobi sweep order beside the reference's
  Gauss-Seidel order, plus FP64 twins.
* ``_ref/libfs2d_ref.so``: the reference's own ``src/fluid_solver_cpu.cpp``
  compiled from the reference checkout by ``oracle/Makefile`` (present only
  where it was built; never committed).

Arrays are float32, shape (R, R), C-contiguous, ring included.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "libfs2d_oracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libfs2d_ref.so")

BND_CONTINUOUS, BND_OPPOSITE_H, BND_OPPOSITE_V = 0, 1, 2
GS_LEX, JACOBI = 0, 1

_f = ctypes.c_float
_i = ctypes.c_int
_p = ctypes.c_void_p


# One SFML vertex (position, RGBA color, texture coordinates) and a line of two.
VERTEX = np.dtype([("x", "<f4"), ("y", "<f4"), ("color", "u1", (4,)), ("tex", "<f4", (2,))])
LINE = np.dtype([("start", VERTEX), ("end", VERTEX)])


def _ptr(a: np.ndarray) -> int:
    assert a.flags.c_contiguous
    return a.ctypes.data


def _check(a: np.ndarray, R: int, dtype=np.float32) -> None:
    if a.dtype != dtype or a.shape != (R, R) or not a.flags.c_contiguous:
        raise ValueError(f"expected C-contiguous {dtype.__name__} ({R},{R}), got {a.dtype} {a.shape}")


class _Lib:
    def __init__(self, path: str):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run __graft_entry__.build())")
        self.lib = ctypes.CDLL(path)


class Oracle(_Lib):
    """Binding of oracle/fs2d_oracle.c."""

    def __init__(self, path: str = ORACLE_LIB):
        super().__init__(path)
        L = self.lib
        L.fso_set_boundary.argtypes = [_p, _i, _i]
        L.fso_add_sources.argtypes = [_p, _p, _i, _f]
        L.fso_diffuse.argtypes = [_p, _i, _i, _f, _f, _i, _i]
        L.fso_advect.argtypes = [_p, _p, _p, _i, _i, _f, _i]
        L.fso_project.argtypes = [_p, _p, _i, _i, _i, _p]
        L.fso_solve.argtypes = [_i, _p, _p, _f, _p, _p, _p, _p, _f, _f, _i, _i, _i]
        L.fso_divergence_norm_f64.argtypes = [_p, _p, _i]
        L.fso_divergence_norm_f64.restype = ctypes.c_double
        L.fso_project_f64.argtypes = [_p, _p, _i, _i, _i]
        L.fso_poisson_residual_f64.argtypes = [_p, _p, _i]
        L.fso_poisson_residual_f64.restype = ctypes.c_double
        L.fso_divergence.argtypes = [_p, _p, _p, _i]
        L.fso_render_density.argtypes = [_p, _i, _i, _f, _f, _f, _p]
        L.fso_render_velocity.argtypes = [_p, _p, _i, _i, _f, _f, _p]

    # each call works in place on its array arguments
    def set_boundary(self, x, mode):
        R = x.shape[0]; _check(x, R)
        self.lib.fso_set_boundary(_ptr(x), R, mode)

    def add_sources(self, x, s, dt):
        R = x.shape[0]; _check(x, R); _check(s, R)
        self.lib.fso_add_sources(_ptr(x), _ptr(s), R, dt)

    def diffuse(self, x, mode, rate, dt, iters, order=GS_LEX):
        R = x.shape[0]; _check(x, R)
        self.lib.fso_diffuse(_ptr(x), R, mode, rate, dt, iters, order)

    def advect(self, x, u, v, mode, dt, trace):
        R = x.shape[0]
        for a in (x, u, v):
            _check(a, R)
        self.lib.fso_advect(_ptr(x), _ptr(u), _ptr(v), R, mode, dt, int(trace))

    def project(self, u, v, iters, order=GS_LEX, want_p=False):
        R = u.shape[0]; _check(u, R); _check(v, R)
        p = np.zeros((R, R), np.float32) if want_p else None
        self.lib.fso_project(_ptr(u), _ptr(v), R, iters, order, _ptr(p) if want_p else None)
        return p

    def solve(self, rho, rho_src, diff, u, v, u_src, v_src, visc, dt, kd=20, kp=20, order=GS_LEX):
        R = rho.shape[0]
        for a in (rho, rho_src, u, v, u_src, v_src):
            _check(a, R)
        self.lib.fso_solve(R, _ptr(rho), _ptr(rho_src), diff, _ptr(u), _ptr(v),
                           _ptr(u_src), _ptr(v_src), visc, dt, kd, kp, order)

    def divergence(self, u, v):
        R = u.shape[0]
        d = np.zeros((R, R), np.float32)
        self.lib.fso_divergence(_ptr(d), _ptr(u), _ptr(v), R)
        return d

    def divergence_norm_f64(self, u, v) -> float:
        R = u.shape[0]; _check(u, R); _check(v, R)
        return float(self.lib.fso_divergence_norm_f64(_ptr(u), _ptr(v), R))

    def project_f64(self, u, v, iters, order=GS_LEX):
        R = u.shape[0]
        _check(u, R, np.float64); _check(v, R, np.float64)
        self.lib.fso_project_f64(_ptr(u), _ptr(v), R, iters, order)

    def poisson_residual_f64(self, d, p) -> float:
        R = d.shape[0]; _check(d, R); _check(p, R)
        return float(self.lib.fso_poisson_residual_f64(_ptr(d), _ptr(p), R))

    def render_density(self, grid, mr, mg, mb):
        """(R, R, 4) uint8 RGBA of a density grid (src/density_grid_renderer.cu:10-29)."""
        R = grid.shape[0]; _check(grid, R)
        out = np.zeros((R, R, 4), np.uint8)
        self.lib.fso_render_density(_ptr(grid), R, R, mr, mg, mb, _ptr(out))
        return out

    def render_velocity(self, u, v, hs, vs):
        """(R, R) array of LINE records (start and end vertex) of a velocity grid
        (src/velocity_grid_renderer.cu:8-44)."""
        R = u.shape[0]; _check(u, R); _check(v, R)
        out = np.zeros((R, R), LINE)
        self.lib.fso_render_velocity(_ptr(u), _ptr(v), R, R, hs, vs, _ptr(out))
        return out


class Reference(_Lib):
    """Binding of the compiled reference (oracle/_ref/libfs2d_ref.so).

    Iteration counts are the reference's literals (20)."""

    def __init__(self, path: str = REF_LIB):
        super().__init__(path)
        L = self.lib
        L.ref_solve.argtypes = [_i, _p, _p, _f, _p, _p, _p, _p, _f, _f]
        L.ref_set_boundary.argtypes = [_p, _i, _i]
        L.ref_add_sources.argtypes = [_p, _p, _i, _f]
        L.ref_diffuse.argtypes = [_p, _i, _i, _f, _f, _i]
        L.ref_advect.argtypes = [_p, _p, _p, _i, _i, _f, _i]
        L.ref_project.argtypes = [_p, _p, _i, _i]

    def set_boundary(self, x, mode):
        R = x.shape[0]; _check(x, R)
        self.lib.ref_set_boundary(_ptr(x), R, mode)

    def add_sources(self, x, s, dt):
        R = x.shape[0]; _check(x, R); _check(s, R)
        self.lib.ref_add_sources(_ptr(x), _ptr(s), R, dt)

    def diffuse(self, x, mode, rate, dt, iters):
        R = x.shape[0]; _check(x, R)
        self.lib.ref_diffuse(_ptr(x), R, mode, rate, dt, iters)

    def advect(self, x, u, v, mode, dt, trace):
        R = x.shape[0]
        for a in (x, u, v):
            _check(a, R)
        self.lib.ref_advect(_ptr(x), _ptr(u), _ptr(v), R, mode, dt, int(trace))

    def project(self, u, v, iters):
        R = u.shape[0]; _check(u, R); _check(v, R)
        self.lib.ref_project(_ptr(u), _ptr(v), R, iters)

    def solve(self, rho, rho_src, diff, u, v, u_src, v_src, visc, dt):
        R = rho.shape[0]
        for a in (rho, rho_src, u, v, u_src, v_src):
            _check(a, R)
        self.lib.ref_solve(R, _ptr(rho), _ptr(rho_src), diff, _ptr(u), _ptr(v),
                           _ptr(u_src), _ptr(v_src), visc, dt)


def reference_available() -> bool:
    return os.path.exists(REF_LIB)
