"""Write the golden fixtures of the stable-fluids step.

Every expected output here comes from the reference's own CPU solver
(src/fluid_solver_cpu.cpp of mworchel/fluid-2d), compiled from the reference
checkout by ``make -C oracle ref`` into oracle/_ref/libfs2d_ref.so.  Only the
input/output arrays are committed (``*.npz``, float32, ring included); the
reference itself never enters the repository.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "fluid-2d_amd"))

import oracle  # noqa: E402
from fs2d_amd import workloads as W  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))

STAGE_SIZES = (4, 6, 10, 18, 34)
SOLVE_SIZES = (4, 6, 10, 18, 34)


def stages(ref: oracle.Reference, R: int) -> dict:
    rho, u, v, rs, us, vs = W.random_state(R, seed=100 + R, vel_scale=0.05)
    _, uf, vf, _, _, _ = W.random_state(R, seed=200 + R, vel_scale=1.0)  # fast flow
    out = {"rho": rho, "u": u, "v": v, "rs": rs, "us": us, "vs": vs, "uf": uf, "vf": vf}
    for m in range(3):
        x = rho.copy(); ref.set_boundary(x, m); out[f"bnd{m}"] = x
    x = rho.copy(); ref.add_sources(x, rs, 0.1); out["add"] = x
    for m, field in ((0, rho), (1, u), (2, v)):
        for rate in (1e-3, 5e-2):
            for it in (1, 7, 20):
                x = field.copy(); ref.diffuse(x, m, rate, 0.1, it)
                out[f"diff_m{m}_r{rate:g}_k{it}"] = x
    for tag, (uu, vv) in (("slow", (u, v)), ("fast", (uf, vf))):
        x = rho.copy(); ref.advect(x, uu, vv, 0, 0.1, 1); out[f"scatter_{tag}"] = x
        x = uu.copy(); ref.advect(x, uu, vv, 1, 0.1, 0); out[f"gather_u_{tag}"] = x
        x = vv.copy(); ref.advect(x, uu, vv, 2, 0.1, 0); out[f"gather_v_{tag}"] = x
    for it in (1, 7, 20):
        a, b = uf.copy(), vf.copy(); ref.project(a, b, it)
        out[f"proj_u_k{it}"], out[f"proj_v_k{it}"] = a, b
    return out


def solves(ref: oracle.Reference, R: int) -> dict:
    rho, u, v, rs, us, vs = W.random_state(R, seed=300 + R, vel_scale=0.05)
    out = {"rho": rho, "u": u, "v": v, "rs": rs, "us": us, "vs": vs}
    a = [rho.copy(), u.copy(), v.copy()]
    for step in range(1, 11):
        ref.solve(a[0], rs, 1e-3, a[1], a[2], us, vs, 1e-3, 0.1)
        if step in (1, 10):
            out[f"rho_s{step}"], out[f"u_s{step}"], out[f"v_s{step}"] = (x.copy() for x in a)
    return out


def config1(ref: oracle.Reference) -> dict:
    """128x128 interior + ring, zero start, disc sources, dt=0.1, nu=kappa=0."""
    R = 130
    rho, u, v = W.zero_state(R)
    rs, us, vs = W.disc_sources(R)
    out = {"rs": rs, "us": us, "vs": vs}
    for step in range(1, 201):
        ref.solve(rho, rs, 0.0, u, v, us, vs, 0.0, 0.1)
        if step in (1, 10, 200):
            out[f"rho_s{step}"], out[f"u_s{step}"], out[f"v_s{step}"] = rho.copy(), u.copy(), v.copy()
    return out


def main() -> None:
    ref = oracle.Reference()
    manifest = {"source": "reference src/fluid_solver_cpu.cpp compiled by oracle/Makefile (ref)",
                "layout": "float32 (R,R) row-major, ring included", "files": {}}
    for R in STAGE_SIZES:
        name = f"stages_R{R}.npz"
        np.savez_compressed(os.path.join(OUT, name), **stages(ref, R))
        manifest["files"][name] = {"R": R, "dt": 0.1, "rates": [1e-3, 5e-2], "iters": [1, 7, 20]}
    for R in SOLVE_SIZES:
        name = f"solve_R{R}.npz"
        np.savez_compressed(os.path.join(OUT, name), **solves(ref, R))
        manifest["files"][name] = {"R": R, "dt": 0.1, "diffusion": 1e-3, "viscosity": 1e-3,
                                   "iterations": 20, "steps": [1, 10]}
    np.savez_compressed(os.path.join(OUT, "config1_R130.npz"), **config1(ref))
    manifest["files"]["config1_R130.npz"] = {"R": 130, "dt": 0.1, "diffusion": 0.0, "viscosity": 0.0,
                                            "iterations": 20, "steps": [1, 10, 200],
                                            "sources": "disc r=8 at centre, rho 100, v -5"}
    with open(os.path.join(OUT, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
