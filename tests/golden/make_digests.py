"""Write sha256 digests of whole steps at the large workload sizes.

The solver is bit-exact, so a digest of each output field (float32 bytes,
(R, R) row-major, ring included) is a complete fixture, and a few hundred bytes
stand in for arrays of up to 268 MB that could not be committed.

Expected outputs come from:

* the reference's own CPU solver (src/fluid_solver_cpu.cpp of mworchel/fluid-2d,
  compiled by ``make -C oracle ref`` into oracle/_ref/libfs2d_ref.so), with its
  private stages composed in the order of fluid_solver_cpu::solve
  (src/fluid_solver_cpu.cpp:16-30) so the iteration counts can be 40 and 80
  instead of the literal 20 -- the Gauss-Seidel workloads;
* the oracle's Jacobi order (oracle/fs2d_oracle.c, pinned by tests/test_oracle.py)
  for the batched workload, which the reference has no counterpart for.

Each record also holds the digests of the inputs, so a consumer that regenerates
them (fs2d_amd.workloads, seeded) can first check it built the same arrays.

Run from the repository root (about 15 minutes on 4 cores; needs oracle/_ref):
    python tests/golden/make_digests.py [name ...]
"""
from __future__ import annotations

import hashlib
import json
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "fluid-2d_amd"))

import oracle  # noqa: E402
from fs2d_amd import workloads as W  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "digests.json")

DIFFUSION, VISCOSITY, DT = 1e-5, 1e-6, 0.05


def digest(a: np.ndarray) -> dict:
    a = np.ascontiguousarray(a, dtype=np.float32)
    with np.errstate(invalid="ignore", over="ignore"):
        return {"sha256": hashlib.sha256(a.tobytes()).hexdigest(),
                "sum": float(a.astype(np.float64).sum()),
                "max_abs": float(np.abs(a).max())}


def reference_step(ref: oracle.Reference, rho, u, v, K: int) -> None:
    """fluid_solver_cpu::solve (src/fluid_solver_cpu.cpp:16-30) with zero
    sources and K sweeps in every diffuse and projection, through the
    reference's own compiled stages.  Adding a zero source is x + dt * 0 = x
    (+0 added; -0 cells become +0), so it is kept to follow the reference."""
    R = rho.shape[0]
    z = np.zeros((R, R), np.float32)
    ref.add_sources(rho, z, DT)
    ref.diffuse(rho, 0, DIFFUSION, DT, K)
    ref.advect(rho, u, v, 0, DT, 1)
    ref.add_sources(u, z, DT)
    ref.add_sources(v, z, DT)
    ref.diffuse(u, 1, VISCOSITY, DT, K)
    ref.diffuse(v, 2, VISCOSITY, DT, K)
    ref.project(u, v, K)
    tu, tv = u.copy(), v.copy()
    ref.advect(u, tu, tv, 1, DT, 0)
    ref.advect(v, tu, tv, 2, DT, 0)
    ref.project(u, v, K)


def vortex_record(name: str, R: int, seed: int, K: int, checkpoints) -> dict:
    ref = oracle.Reference()
    rho, u, v = W.vortex_state(R, seed=seed)
    rec = {"R": R, "seed": seed, "K": K, "dt": DT, "diffusion": DIFFUSION, "viscosity": VISCOSITY,
           "order": "gauss-seidel (reference)", "state": "workloads.vortex_state(R, seed)",
           "sources": "zero", "inputs": {"rho": digest(rho), "u": digest(u), "v": digest(v)}, "steps": {}}
    t0 = time.time()
    for step in range(1, max(checkpoints) + 1):
        reference_step(ref, rho, u, v, K)
        if step in checkpoints:
            rec["steps"][str(step)] = {"rho": digest(rho), "u": digest(u), "v": digest(v)}
            print(f"{name}: step {step} done ({time.time() - t0:.0f} s)", flush=True)
    return rec


def batch_record(name: str, count: int, seed: int, steps: int) -> dict:
    """The batched workload: workloads.batch_config(count, seed) simulations with
    workloads.batch_state starts, no sources, the oracle's Jacobi order."""
    o = oracle.Oracle()
    cfg = W.batch_config(count, seed=seed)
    sims = []
    t0 = time.time()
    for k, (R, K, dt) in enumerate(cfg):
        rho, u, v = W.batch_state(R, seed=seed, sim_index=k)
        inputs = {"rho": digest(rho), "u": digest(u), "v": digest(v)}
        z = np.zeros((R, R), np.float32)
        for _ in range(steps):
            o.solve(rho, z, DIFFUSION, u, v, z, z, VISCOSITY, dt, K, K, oracle.JACOBI)
        sims.append({"R": R, "K": K, "dt": dt, "inputs": inputs,
                     "out": {"rho": digest(rho), "u": digest(u), "v": digest(v)}})
    print(f"{name}: {count} simulations done ({time.time() - t0:.0f} s)", flush=True)
    return {"count": count, "seed": seed, "steps": steps, "diffusion": DIFFUSION, "viscosity": VISCOSITY,
            "order": "jacobi (oracle)", "state": "workloads.batch_state(R, seed, k)", "sources": "zero",
            "sims": sims}


JOBS = {
    "config2_R1026": (vortex_record, ("config2_R1026", 1026, 2, 40, (1, 10, 500))),
    "config3_R4098": (vortex_record, ("config3_R4098", 4098, 3, 80, (1, 3))),
    "config3_R8194": (vortex_record, ("config3_R8194", 8194, 3, 80, (1,))),
    "config4_batch256": (batch_record, ("config4_batch256", 256, 0, 2)),
}


def _run(name: str):
    fn, args = JOBS[name]
    return name, fn(*args)


def main(names) -> None:
    names = names or list(JOBS)
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    data.setdefault("format", "sha256 of float32 (R,R) row-major bytes, ring included; sum in float64; max |x|")
    data.setdefault("records", {})
    with ProcessPoolExecutor(max_workers=min(4, len(names))) as ex:
        for name, rec in ex.map(_run, names):
            data["records"][name] = rec
            with open(OUT, "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:])
