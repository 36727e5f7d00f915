"""Python binding of libfs2d_hip.so, the MI355X stable-fluids solver.

A thin ctypes layer over the C ABI of ``include/fs2d_hip.h``; every call goes
to the HIP library.  There is no CPU fallback: if the library is missing or
has not been built, importing the solver raises.

Arrays are numpy float32 (R, R), C-contiguous, ring included (the layout of
the reference's ``grid<float>``, src/grid.hpp:35-44).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "lib", "libfs2d_hip.so")

GS_LEX, JACOBI = 0, 1
JACOBI_WAVE, JACOBI_LDS = 1, 2  # fs2d_params.jacobi_kernel
LAUNCH_GRAPH, LAUNCH_EAGER = 0, 1  # fs2d_params.launch_mode
DENSITY, VELOCITY_U, VELOCITY_V = 0, 1, 2

_f, _i, _p = ctypes.c_float, ctypes.c_int, ctypes.c_void_p


class Fs2dError(RuntimeError):
    pass


class Params(ctypes.Structure):
    _fields_ = [("diffuse_iterations", _i), ("pressure_iterations", _i), ("sweep_order", _i),
                ("jacobi_block", _i), ("gs_tile_cols", _i), ("gs_block", _i), ("gs_lag", _i),
                ("jacobi_kernel", _i), ("launch_mode", _i)]


class BatchSim(ctypes.Structure):
    _fields_ = [("rows", _i), ("iterations", _i), ("dt", _f), ("diffusion_rate", _f), ("viscosity", _f)]


_lib = None


class ColorMultipliers(ctypes.Structure):
    """fs2d_color_multipliers (include/fs2d_render.h)."""
    _fields_ = [("r", _f), ("g", _f), ("b", _f)]


# fs2d_vertex / fs2d_line (include/fs2d_render.h): an SFML vertex and a line of two.
VERTEX = np.dtype([("x", "<f4"), ("y", "<f4"), ("color", "u1", (4,)), ("tex", "<f4", (2,))])
LINE = np.dtype([("start", VERTEX), ("end", VERTEX)])


def library() -> ctypes.CDLL:
    """Load libfs2d_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Fs2dError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                        "(make -C fluid-2d_amd); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    H = _p
    sig = {
        "fs2d_default_params": (None, [ctypes.POINTER(Params)]),
        "fs2d_hip_create": (_i, [_i, _i, ctypes.POINTER(Params), ctypes.POINTER(H)]),
        "fs2d_hip_destroy": (_i, [H]),
        "fs2d_hip_solve": (_i, [H, _p, _p, _f, _p, _p, _p, _p, _f, _f]),
        "fs2d_hip_upload": (_i, [H, _p, _p, _p]),
        "fs2d_hip_set_sources": (_i, [H, _p, _p, _p]),
        "fs2d_hip_step_n": (_i, [H, _i, _f, _f, _f]),
        "fs2d_hip_download": (_i, [H, _p, _p, _p]),
        "fs2d_hip_sync": (_i, [H]),
        "fs2d_hip_graph_count": (_i, [H, ctypes.POINTER(_i)]),
        "fs2d_hip_copy_bandwidth": (_i, [ctypes.c_longlong, _i, ctypes.POINTER(_f)]),
        "fs2d_hip_batch_create": (_i, [_i, ctypes.POINTER(BatchSim), ctypes.POINTER(Params), ctypes.POINTER(H)]),
        "fs2d_hip_batch_destroy": (_i, [H]),
        "fs2d_hip_batch_upload": (_i, [H, _i, _p, _p, _p]),
        "fs2d_hip_batch_set_sources": (_i, [H, _i, _p, _p, _p]),
        "fs2d_hip_batch_step_n": (_i, [H, _i]),
        "fs2d_hip_batch_download": (_i, [H, _i, _p, _p, _p]),
        "fs2d_hip_batch_sync": (_i, [H]),
        "fs2d_hip_batch_info": (_i, [H, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                                     ctypes.POINTER(_i), ctypes.POINTER(ctypes.c_longlong)]),
        "fs2d_hip_batch_last_error": (ctypes.c_char_p, [H]),
        "fs2d_hip_batch_occupancy": (_i, [H, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
        "fs2d_hip_batch_set_occupancy": (_i, [H, _i, _i]),
        "fs2d_hip_upload_device": (_i, [H, _p, _p, _p]),
        "fs2d_hip_download_device": (_i, [H, _p, _p, _p]),
        "fs2d_hip_stage_add_sources": (_i, [H, _i, _f]),
        "fs2d_hip_stage_diffuse": (_i, [H, _i, _f, _f, _i]),
        "fs2d_hip_stage_advect_density": (_i, [H, _f]),
        "fs2d_hip_stage_advect_velocity": (_i, [H, _f]),
        "fs2d_hip_stage_project": (_i, [H, _i]),
        "fs2d_hip_stage_set_boundary": (_i, [H, _i]),
        "fs2d_hip_download_pressure": (_i, [H, _p]),
        "fs2d_hip_stream": (_p, [H]),
        "fs2d_hip_time_pressure_sweeps": (_i, [H, _i, ctypes.POINTER(_f)]),
        "fs2d_hip_time_diffuse_sweeps": (_i, [H, _f, _f, _f, _i, ctypes.POINTER(_f)]),
        "fs2d_hip_time_joint_sweeps": (_i, [H, _f, _f, _f, _i, ctypes.POINTER(_f)]),
        "fs2d_gs_pipeline_offset": (_i, [_i, _i, _i, _i, _i, ctypes.POINTER(_i)]),
        "fs2d_division_plan": (_i, [_f] + [ctypes.POINTER(_f)] * 5 + [ctypes.POINTER(_i)]),
        "fs2d_hip_gs_policy": (_i, [H, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
        "fs2d_hip_launches_per_step": (_i, [H, ctypes.POINTER(_i)]),
        "fs2d_hip_last_error": (ctypes.c_char_p, [H]),
        "fs2d_hip_abi_version": (_i, []),
        "fs2d_hip_project_noalloc": (_i, [_i, ctypes.POINTER(Params), _p, _p, _p, _p]),
        "fs2d_hip_render_density": (_i, [H, ctypes.POINTER(ColorMultipliers), _p]),
        "fs2d_hip_render_density_device": (_i, [H, ctypes.POINTER(ColorMultipliers), _p]),
        "fs2d_hip_render_velocity": (_i, [H, _f, _f, _p]),
        "fs2d_hip_render_velocity_device": (_i, [H, _f, _f, _p]),
        "fs2d_hip_time_render": (_i, [H, _i, ctypes.POINTER(_f), ctypes.POINTER(_f)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _arr(a: np.ndarray, R: int) -> int:
    if not isinstance(a, np.ndarray) or a.dtype != np.float32 or a.shape != (R, R) or not a.flags.c_contiguous:
        raise ValueError(f"expected C-contiguous float32 array of shape ({R},{R})")
    return a.ctypes.data


class Solver:
    """One simulation on the current HIP device (wraps an fs2d_handle).

    The constructor mirrors ``fluid_solver_gpu(rows, cols)``
    (src/fluid_solver_gpu.cuh:14); ``solve`` mirrors ``fluid_solver::solve``
    (src/fluid_solver.hpp:16-24)."""

    def __init__(self, rows: int, cols: int | None = None, diffuse_iterations: int = 20,
                 pressure_iterations: int = 20, order: int = GS_LEX, jacobi_block: int = 0,
                 gs_tile_cols: int = 0, gs_block: int = 0, gs_lag: int = 0, jacobi_kernel: int = 0,
                 launch_mode: int = LAUNCH_GRAPH):
        L = library()
        cols = rows if cols is None else cols
        prm = Params()
        L.fs2d_default_params(ctypes.byref(prm))
        prm.diffuse_iterations = diffuse_iterations
        prm.pressure_iterations = pressure_iterations
        prm.sweep_order = order
        prm.jacobi_block = jacobi_block
        prm.gs_tile_cols = gs_tile_cols
        prm.gs_block = gs_block
        prm.gs_lag = gs_lag
        prm.jacobi_kernel = jacobi_kernel
        prm.launch_mode = launch_mode
        h = _p()
        rc = L.fs2d_hip_create(rows, cols, ctypes.byref(prm), ctypes.byref(h))
        if rc != 0:
            raise Fs2dError(f"fs2d_hip_create failed ({rc}): {L.fs2d_hip_last_error(None).decode()}")
        self._L, self._h, self.R = L, h, rows

    def _ck(self, rc: int) -> None:
        if rc != 0:
            raise Fs2dError(f"fs2d error {rc}: {self._L.fs2d_hip_last_error(self._h).decode()}")

    def close(self) -> None:
        if self._h:
            self._L.fs2d_hip_destroy(self._h)
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

    # -- the reference interface ------------------------------------------------
    def solve(self, rho, rho_src, diffusion, u, v, u_src, v_src, viscosity, dt) -> None:
        R = self.R
        self._ck(self._L.fs2d_hip_solve(self._h, _arr(rho, R), _arr(rho_src, R), diffusion, _arr(u, R),
                                        _arr(v, R), _arr(u_src, R), _arr(v_src, R), viscosity, dt))

    # -- resident API -------------------------------------------------------------
    def upload(self, rho, u, v) -> None:
        R = self.R
        self._ck(self._L.fs2d_hip_upload(self._h, _arr(rho, R), _arr(u, R), _arr(v, R)))

    def set_sources(self, rho_src=None, u_src=None, v_src=None) -> None:
        R = self.R
        ptr = lambda a: None if a is None else _arr(a, R)
        self._ck(self._L.fs2d_hip_set_sources(self._h, ptr(rho_src), ptr(u_src), ptr(v_src)))

    def step_n(self, n: int, diffusion: float, viscosity: float, dt: float) -> None:
        self._ck(self._L.fs2d_hip_step_n(self._h, n, diffusion, viscosity, dt))

    def graph_count(self) -> int:
        n = _i()
        self._ck(self._L.fs2d_hip_graph_count(self._h, ctypes.byref(n)))
        return int(n.value)

    def sync(self) -> None:
        self._ck(self._L.fs2d_hip_sync(self._h))

    def download(self):
        R = self.R
        out = [np.empty((R, R), np.float32) for _ in range(3)]
        self._ck(self._L.fs2d_hip_download(self._h, *(a.ctypes.data for a in out)))
        return tuple(out)

    def pressure(self) -> np.ndarray:
        p = np.empty((self.R, self.R), np.float32)
        self._ck(self._L.fs2d_hip_download_pressure(self._h, p.ctypes.data))
        return p

    # -- stages -------------------------------------------------------------------
    def add_sources(self, field: int, dt: float) -> None:
        self._ck(self._L.fs2d_hip_stage_add_sources(self._h, field, dt))

    def diffuse(self, field: int, rate: float, dt: float, iterations: int) -> None:
        self._ck(self._L.fs2d_hip_stage_diffuse(self._h, field, rate, dt, iterations))

    def advect_density(self, dt: float) -> None:
        self._ck(self._L.fs2d_hip_stage_advect_density(self._h, dt))

    def advect_velocity(self, dt: float) -> None:
        self._ck(self._L.fs2d_hip_stage_advect_velocity(self._h, dt))

    def project(self, iterations: int) -> None:
        self._ck(self._L.fs2d_hip_stage_project(self._h, iterations))

    def set_boundary(self, field: int) -> None:
        self._ck(self._L.fs2d_hip_stage_set_boundary(self._h, field))

    # -- measurement --------------------------------------------------------------
    def stream(self) -> int:
        return int(self._L.fs2d_hip_stream(self._h) or 0)

    def time_pressure_sweeps(self, reps: int) -> float:
        ms = _f()
        self._ck(self._L.fs2d_hip_time_pressure_sweeps(self._h, reps, ctypes.byref(ms)))
        return float(ms.value)

    def time_joint_sweeps(self, diffusion: float, viscosity: float, dt: float, reps: int) -> float:
        """Milliseconds of a Gauss-Seidel step's joint level launches (three
        diffuses and the first pressure solve, pipelined), on copies of the state."""
        ms = _f()
        self._ck(self._L.fs2d_hip_time_joint_sweeps(self._h, diffusion, viscosity, dt, reps, ctypes.byref(ms)))
        return float(ms.value)

    def gs_policy(self, kind: int) -> dict:
        """Gauss-Seidel launch policy of the diffuse (0) or pressure (1) solves."""
        tw, m, lag, lv = _i(), _i(), _i(), _i()
        self._ck(self._L.fs2d_hip_gs_policy(self._h, kind, ctypes.byref(tw), ctypes.byref(m), ctypes.byref(lag),
                                            ctypes.byref(lv)))
        return {"tile_cols": tw.value, "sweeps_per_unit": m.value, "lag": lag.value, "levels": lv.value}

    def time_diffuse_sweeps(self, diffusion: float, viscosity: float, dt: float, reps: int) -> float:
        """Milliseconds of one step's three diffuse solves (HIP events), on
        copies of the state."""
        ms = _f()
        self._ck(self._L.fs2d_hip_time_diffuse_sweeps(self._h, diffusion, viscosity, dt, reps, ctypes.byref(ms)))
        return float(ms.value)

    def render_density(self, r: float, g: float, b: float) -> np.ndarray:
        """The density as (R, R, 4) uint8 RGBA (density_grid_renderer's kernel,
        src/density_grid_renderer.cu:10-29)."""
        out = np.empty((self.R, self.R, 4), np.uint8)
        m = ColorMultipliers(r, g, b)
        self._ck(self._L.fs2d_hip_render_density(self._h, ctypes.byref(m), out.ctypes.data))
        return out

    def render_velocity(self, horizontal_scale: float, vertical_scale: float) -> np.ndarray:
        """The velocity as an (R, R) array of LINE records (velocity_grid_renderer's
        kernel, src/velocity_grid_renderer.cu:8-44)."""
        out = np.empty((self.R, self.R), LINE)
        self._ck(self._L.fs2d_hip_render_velocity(self._h, horizontal_scale, vertical_scale, out.ctypes.data))
        return out

    def time_render(self, reps: int = 10):
        """Mean ms of the density and the velocity conversion (device output)."""
        a, b = _f(), _f()
        self._ck(self._L.fs2d_hip_time_render(self._h, reps, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def launches_per_step(self) -> int:
        n = _i()
        self._ck(self._L.fs2d_hip_launches_per_step(self._h, ctypes.byref(n)))
        return int(n.value)


EXPORTED_SYMBOLS = (
    "fs2d_default_params", "fs2d_hip_create", "fs2d_hip_destroy", "fs2d_hip_solve", "fs2d_hip_upload",
    "fs2d_hip_set_sources", "fs2d_hip_step_n", "fs2d_hip_download", "fs2d_hip_sync",
    "fs2d_hip_graph_count", "fs2d_hip_copy_bandwidth", "fs2d_hip_batch_create", "fs2d_hip_batch_destroy", "fs2d_hip_batch_upload",
    "fs2d_hip_batch_set_sources", "fs2d_hip_batch_step_n", "fs2d_hip_batch_download", "fs2d_hip_batch_sync",
    "fs2d_hip_batch_info", "fs2d_hip_batch_last_error", "fs2d_hip_batch_occupancy", "fs2d_hip_batch_set_occupancy",
    "fs2d_net_create", "fs2d_net_destroy", "fs2d_net_param_count", "fs2d_net_get_params", "fs2d_net_set_params",
    "fs2d_net_get_grads", "fs2d_net_train_step", "fs2d_net_predict_host", "fs2d_net_stream", "fs2d_net_last_error",
    "fs2d_net_time_forward", "fs2d_net_forward_flops",
    "fs2d_hip_project_seeded", "fs2d_hip_pressure_residual",
    "fs2d_hip_upload_device", "fs2d_hip_download_device", "fs2d_hip_stage_add_sources",
    "fs2d_hip_stage_diffuse", "fs2d_hip_stage_advect_density", "fs2d_hip_stage_advect_velocity",
    "fs2d_hip_stage_project", "fs2d_hip_stage_set_boundary", "fs2d_hip_download_pressure",
    "fs2d_hip_stream", "fs2d_hip_time_pressure_sweeps", "fs2d_hip_time_diffuse_sweeps", "fs2d_hip_time_joint_sweeps", "fs2d_hip_gs_policy", "fs2d_gs_pipeline_offset", "fs2d_division_plan", "fs2d_hip_launches_per_step",
    "fs2d_hip_last_error", "fs2d_hip_abi_version",
    "fs2d_hip_render_density", "fs2d_hip_render_density_device", "fs2d_hip_render_velocity",
    "fs2d_hip_render_velocity_device", "fs2d_hip_time_render", "fs2d_hip_project_noalloc",
)


def project_noalloc(rows: int, u_ptr: int, v_ptr: int, pressure_ptr: int | None = None, stream: int | None = None,
                    pressure_iterations: int = 20, order: int = GS_LEX, jacobi_block: int = 0, gs_block: int = 0,
                    gs_lag: int = 0, gs_tile_cols: int = 0, jacobi_kernel: int = 0) -> None:
    """fs2d_hip_project_noalloc: the projection on caller-owned device arrays
    (addresses of rows x rows float32, row-major) enqueued on `stream` (a
    hipStream_t handle; None = the null stream); allocates nothing."""
    L = library()
    prm = Params()
    L.fs2d_default_params(ctypes.byref(prm))
    prm.pressure_iterations = pressure_iterations
    prm.sweep_order = order
    prm.jacobi_block = jacobi_block
    prm.gs_block = gs_block
    prm.gs_lag = gs_lag
    prm.gs_tile_cols = gs_tile_cols
    prm.jacobi_kernel = jacobi_kernel
    rc = L.fs2d_hip_project_noalloc(rows, ctypes.byref(prm), u_ptr, v_ptr, pressure_ptr, stream)
    if rc != 0:
        raise Fs2dError(f"fs2d_hip_project_noalloc failed ({rc}): {L.fs2d_hip_last_error(None).decode()}")


class Batch:
    """Many independent simulations advanced together (fs2d_hip_batch_*): each
    stage is one launch over all of them.  ``sims`` is a list of
    (rows, iterations, dt, diffusion_rate, viscosity); the Jacobi order is
    the only one batched."""

    def __init__(self, sims, jacobi_block: int = 0, launch_mode: int = LAUNCH_GRAPH, order: int = JACOBI):
        L = library()
        arr = (BatchSim * len(sims))()
        for k, (rows, iters, dt, diff, visc) in enumerate(sims):
            arr[k] = BatchSim(rows, iters, dt, diff, visc)
        prm = Params()
        L.fs2d_default_params(ctypes.byref(prm))
        prm.sweep_order = order
        prm.jacobi_block = jacobi_block
        prm.launch_mode = launch_mode
        h = _p()
        rc = L.fs2d_hip_batch_create(len(sims), arr, ctypes.byref(prm), ctypes.byref(h))
        if rc != 0:
            raise Fs2dError(f"fs2d_hip_batch_create failed ({rc}): {L.fs2d_hip_batch_last_error(None).decode()}")
        self._L, self._h = L, h
        self.rows = [int(x[0]) for x in sims]

    def _ck(self, rc: int) -> None:
        if rc != 0:
            raise Fs2dError(f"fs2d batch error {rc}: {self._L.fs2d_hip_batch_last_error(self._h).decode()}")

    def close(self) -> None:
        if self._h:
            self._L.fs2d_hip_batch_destroy(self._h)
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

    def upload(self, sim: int, rho, u, v) -> None:
        R = self.rows[sim]
        self._ck(self._L.fs2d_hip_batch_upload(self._h, sim, _arr(rho, R), _arr(u, R), _arr(v, R)))

    def set_sources(self, sim: int, rho_src=None, u_src=None, v_src=None) -> None:
        R = self.rows[sim]
        args = [None if a is None else _arr(a, R) for a in (rho_src, u_src, v_src)]
        self._ck(self._L.fs2d_hip_batch_set_sources(self._h, sim, *args))

    def step_n(self, n: int) -> None:
        self._ck(self._L.fs2d_hip_batch_step_n(self._h, n))

    def sync(self) -> None:
        self._ck(self._L.fs2d_hip_batch_sync(self._h))

    def download(self, sim: int):
        R = self.rows[sim]
        out = [np.empty((R, R), np.float32) for _ in range(3)]
        self._ck(self._L.fs2d_hip_batch_download(self._h, sim, *(o.ctypes.data for o in out)))
        return out

    def set_occupancy(self, jacobi_blocks_per_cu: int, stage_blocks_per_cu: int) -> None:
        """Fix the persistent launches' blocks per CU (before the first step)."""
        self._ck(self._L.fs2d_hip_batch_set_occupancy(self._h, jacobi_blocks_per_cu, stage_blocks_per_cu))

    def info(self) -> dict:
        d, t, l, g = _i(), _i(), _i(), _i()
        c = ctypes.c_longlong()
        self._ck(self._L.fs2d_hip_batch_info(self._h, ctypes.byref(d), ctypes.byref(t), ctypes.byref(l),
                                            ctypes.byref(g), ctypes.byref(c)))
        jb, sb, cu = _i(), _i(), _i()
        self._ck(self._L.fs2d_hip_batch_occupancy(self._h, ctypes.byref(jb), ctypes.byref(sb), ctypes.byref(cu)))
        return {"jacobi_depth": d.value, "tile_rows": t.value, "launches_per_step": l.value, "graphs": g.value,
                "interior_cells": c.value, "jacobi_blocks_per_cu": jb.value, "stage_blocks_per_cu": sb.value,
                "compute_units": cu.value}


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20) -> float:
    """GB/s of an on-device float4 copy (read + write counted)."""
    L = library()
    g = _f()
    rc = L.fs2d_hip_copy_bandwidth(nbytes, reps, ctypes.byref(g))
    if rc != 0:
        raise Fs2dError(f"fs2d_hip_copy_bandwidth failed ({rc}): {L.fs2d_hip_last_error(None).decode()}")
    return float(g.value)


def division_plan(denom: float) -> dict:
    """The division of a diffuse solve by denom = 1 + 4a (host computation, no
    device call): inv = RN(1/denom) and inv_lo = RN(1/denom - inv), the pair of
    the corrected quotient; two_hi <= 1/denom and two_lo = RN(1/denom - two_hi)
    > 0, the pair of the two-product quotient alone; lo, the smallest nonzero
    |dividend| the FMA-based quotients take; checked, whether the two-product
    quotient alone is correctly rounded from lo up (what the Gauss-Seidel unit
    kernels then use)."""
    L = library()
    out = [_f() for _ in range(5)]
    chk = _i()
    rc = L.fs2d_division_plan(float(denom), *(ctypes.byref(o) for o in out), ctypes.byref(chk))
    if rc != 0:
        raise Fs2dError(f"fs2d_division_plan failed ({rc}): {L.fs2d_hip_last_error(None).decode()}")
    inv, inv_lo, two_hi, two_lo, lo = (o.value for o in out)
    return {"inv": inv, "inv_lo": inv_lo, "two_hi": two_hi, "two_lo": two_lo, "lo": lo, "checked": bool(chk.value)}


def gs_pipeline_offset(rows: int, diffuse_iterations: int, pressure_iterations: int, tile_cols: int,
                       sweeps_per_unit: int) -> int:
    """Levels a Gauss-Seidel step's first pressure solve trails its diffuses by
    when pipelined (host computation, no device call)."""
    L = library()
    off = _i()
    rc = L.fs2d_gs_pipeline_offset(rows, diffuse_iterations, pressure_iterations, tile_cols, sweeps_per_unit,
                                   ctypes.byref(off))
    if rc != 0:
        raise Fs2dError(f"fs2d_gs_pipeline_offset failed ({rc}): {L.fs2d_hip_last_error(None).decode()}")
    return off.value
