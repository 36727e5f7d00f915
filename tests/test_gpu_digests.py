"""Whole steps at the workloads' own sizes against digests of the reference.

tests/golden/digests.json holds sha256 digests (plus sum and max |x|) of
density, u and v written by the reference's own CPU solver (compiled from
src/fluid_solver_cpu.cpp, its stages composed in the order of
fluid_solver_cpu::solve with 40 or 80 sweeps), and, for the batched workload,
by the oracle's Jacobi order; tests/golden/make_digests.py wrote them.  The
solver is bit-exact, so equal digests are the whole comparison.  Each test
first checks that the seeded inputs it regenerates (fs2d_amd.workloads) are
the ones the digests were made from, then runs the default kernels.
"""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import fs2d_amd
from conftest import GOLDEN
from fs2d_amd import workloads as W

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "digests.json")) as _f:
    DIGESTS = json.load(_f)["records"]


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float32).tobytes()).hexdigest()


def check_fields(fields, want, what):
    for name, a in zip(("rho", "u", "v"), fields):
        w = want[name]
        if sha(a) != w["sha256"]:
            s = float(a.astype(np.float64).sum())
            m = float(np.abs(a).max())
            pytest.fail(f"{what} {name}: digest differs (sum {s!r} vs {w['sum']!r}, max|x| {m!r} vs {w['max_abs']!r})")


def check_inputs(fields, rec, what):
    for name, a in zip(("rho", "u", "v"), fields):
        assert sha(a) == rec["inputs"][name]["sha256"], \
            f"{what}: regenerated input {name} differs from the one the digest was made from"


@pytest.mark.parametrize("name", ["config2_R1026", "config3_R4098", "config3_R8194"])
def test_whole_steps_against_reference_digests(name):
    """The Gauss-Seidel step (reference order, default kernels, graph replay)
    after every recorded step count: 1, 10, 500 at 1024^2 (K = 40); 1 and 3 at
    4096^2 and 1 at 8192^2 (K = 80)."""
    rec = DIGESTS[name]
    R, K = rec["R"], rec["K"]
    state = W.vortex_state(R, seed=rec["seed"])
    check_inputs(state, rec, name)
    with fs2d_amd.Solver(R, diffuse_iterations=K, pressure_iterations=K) as s:
        s.upload(*state)
        s.set_sources(None, None, None)
        done = 0
        for step in sorted(int(k) for k in rec["steps"]):
            s.step_n(step - done, rec["diffusion"], rec["viscosity"], rec["dt"])
            done = step
            check_fields(s.download(), rec["steps"][str(step)], f"{name} after {step} steps")


@pytest.mark.parametrize("occupancy", [None, (1, 1)])
def test_batched_workload_against_oracle_digests(occupancy):
    """The full batched workload (256 simulations of sides 64..512, uneven dt and
    sweep counts), two steps in the Jacobi order, every simulation: with the
    tuned persistent grids, and with one block per CU (every block and wave
    walks hundreds of work entries)."""
    rec = DIGESTS["config4_batch256"]
    cfg = W.batch_config(rec["count"], seed=rec["seed"])
    assert [(R, K, dt) for R, K, dt in cfg] == [(x["R"], x["K"], x["dt"]) for x in rec["sims"]]
    sims = [(R, K, dt, rec["diffusion"], rec["viscosity"]) for R, K, dt in cfg]
    with fs2d_amd.Batch(sims) as b:
        if occupancy:
            b.set_occupancy(*occupancy)
        for k, (R, K, dt) in enumerate(cfg):
            st = W.batch_state(R, seed=rec["seed"], sim_index=k)
            check_inputs(st, rec["sims"][k], f"simulation {k}")
            b.upload(k, *st)
            b.set_sources(k)
        b.step_n(rec["steps"])
        for k in range(len(cfg)):
            check_fields(b.download(k), rec["sims"][k]["out"], f"simulation {k} (R={cfg[k][0]}, K={cfg[k][1]})")
