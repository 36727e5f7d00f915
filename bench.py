"""Benchmark of the MI355X stable-fluids step (one GPU).

Workload (the bandwidth-roofline configuration): 4096 x 4096 interior cells
plus the boundary ring (R = 4098), float32, dt = 0.05, viscosity 1e-6,
diffusion 1e-5, 80 relaxation sweeps for each diffuse and pressure solve,
seeded 64-vortex start with smoothed-noise density, no sources.  A "step" is
one fluid_solver::solve (src/fluid_solver_cpu.cpp:6-31) on the resident
state; inputs are in HBM before the timed region starts.

The headline is the Gauss-Seidel order, whose results equal the reference
solver's bit for bit.  Reported beside it: the Jacobi order (a different
relaxation, agreeing with the reference only where the solves converge) with
its autotuned temporal-block depth; both orders at 8192 x 8192; and the
batched workload, 256 independent simulations (sides 64..512, dt in
[0.02, 0.1], 20 or 40 Jacobi sweeps, seeded 16-vortex starts) advanced by one
launch per stage.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 4096]
                    [--order gs|jacobi] [--no-cpu] [--no-jacobi] [--no-8192]
                    [--no-batch]

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "fluid-2d_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
# wave64 VALU instructions per second the chip can issue: 256 CUs x 4 SIMDs,
# one wave64 instruction per 4 cycles per SIMD (16 lanes), 2.4 GHz
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4


def finite(o):
    """The bench record with every non-finite float as null (strict JSON)."""
    if isinstance(o, dict):
        return {k: finite(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [finite(v) for v in o]
    if isinstance(o, float) and not math.isfinite(o):
        return None
    return o


def profile_value(order: int, n: int, K: int, key: str):
    """A figure from the committed rocprofv3 summary of this configuration
    (profiles/{gs,jacobi}_solves_{n}_traffic.txt, K = 80): HBM bytes per solve
    from FETCH_SIZE / WRITE_SIZE calibrated against a plain copy, or VALU
    instructions per launch; None where no summary covers it."""
    path = os.path.join(ROOT, "profiles", f"{'gs' if order == 0 else 'jacobi'}_solves_{n}_traffic.txt")
    if K != 80 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith(key + ":"):
            return float(line.split(":", 1)[1])
    return None


def bytes_per_cell_step(kd: int, kp: int, a_dens: bool, a_visc: bool) -> int:
    """Algorithmic HBM bytes per interior cell per step: every stage reads and
    writes each distinct field once (float32).  add rho 12, diffuse rho 12*Kd,
    scatter 16, add u,v 24, diffuse u,v 24*Kd, two projections of
    (divergence 12 + 12*Kp + gradient 20), velocity gather 16.  A diffuse with
    a = 0 earns no credit."""
    b = 12 + 16 + 24 + 2 * (32 + 12 * kp) + 16
    if a_dens:
        b += 12 * kd
    if a_visc:
        b += 24 * kd
    return b


LAUNCH_MODE = 0  # fs2d_amd.LAUNCH_GRAPH; --launch eager sets 1
GS_FIXED = {}  # Gauss-Seidel launch parameters fixed by --gs-fixed (empty: tuned)


def fs2d_amd_copy_bandwidth() -> float:
    """On-box copy bandwidth (1 GiB float4 copy kernel, read + write)."""
    import fs2d_amd

    return fs2d_amd.copy_bandwidth(1 << 30, 20)


def run_gpu(order: int, n: int, kd: int, kp: int, steps: int, warmup: int, state, params, time_solve=True):
    import fs2d_amd

    R = n + 2
    diff, visc, dt = params
    s = fs2d_amd.Solver(R, diffuse_iterations=kd, pressure_iterations=kp, order=order, launch_mode=LAUNCH_MODE,
                        **(GS_FIXED if order == 0 else {}))
    s.upload(*state)
    s.set_sources(None, None, None)
    s.step_n(warmup, diff, visc, dt)
    s.sync()
    t0 = time.perf_counter()
    s.step_n(steps, diff, visc, dt)
    s.sync()
    elapsed = time.perf_counter() - t0
    # the step's two relaxations, each timed with HIP events on the solver's
    # own stream around repeated solves (the kernels of the timed steps)
    out = {"sec_step": elapsed / steps, "launches": s.launches_per_step()}
    if time_solve:
        reps = max(3, min(20, int(0.5 / max(elapsed / steps / 5, 1e-4))))
        out["ms_pressure"] = s.time_pressure_sweeps(reps)
        out["ms_diffuse"] = s.time_diffuse_sweeps(diff, visc, dt, max(2, reps // 2))
        if order == 0:
            out["policy"] = {"diffuse": s.gs_policy(0), "pressure": s.gs_policy(1), "joint": s.gs_policy(2)}
            out["ms_joint"] = s.time_joint_sweeps(diff, visc, dt, max(2, reps // 2))
    s.close()
    return out


def solve_roofline(order: int, n: int, K: int, ms: float, kind: str, fields: int, policy) -> dict:
    """Roofline of one relaxation solve (the unit of work: K sweeps over
    `fields` fields of n^2 cells, 12 algorithmic bytes per cell per sweep)."""
    cells = n * n
    alg = 12 * cells * K * fields
    gbs = alg / (ms * 1e-3) / 1e9
    traffic = profile_value(order, n, K, f"traffic_bytes_per_{kind}_solve")
    if order == 0 and kind == "joint":
        name = (f"k_gs_unit_mixed<{policy['tile_cols']},{policy['sweeps_per_unit']},{policy['lag']}> with "
                f"k_gs_unit<..,diffuse> / <..,pressure> (three diffuses and the first pressure solve, pipelined)")
    elif order == 0:
        name = f"k_gs_unit<{policy['tile_cols']},{policy['sweeps_per_unit']},{policy['lag']},{kind}>"
    else:
        name = f"k_jacobi_wave ({kind})"
    r = {"kernel": name, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs <= HBM_PEAK_GBS else None,
         "traffic": int(traffic) if traffic else None,
         "hbm_frac_counters": round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
         "unit_of_work": f"one {kind} solve: {K} sweeps x {fields} field(s) x {cells} cells x 12 B = "
                         f"{alg / 1e9:.2f} GB algorithmic, {ms:.3f} ms (HIP events); traffic = HBM bytes "
                         "per solve by rocprofv3 counters"}
    if gbs > HBM_PEAK_GBS:
        r["algorithmic_over_peak"] = round(gbs / HBM_PEAK_GBS, 3)
        r["note"] = "temporal blocking: the algorithmic bytes are not moved; frac is given only by counters"
    if order == 0:
        lv = policy["levels"]
        r["latency_model"] = {"levels": lv, "us_per_level": round(ms * 1e3 / lv, 2),
                              "model": "the solve is a chain of dependent level launches: time = levels x "
                                       "(slowest unit of the level + launch gap)"}
    return r


def run_batch(count: int, steps: int, warmup: int):
    """The batched workload: `count` simulations drawn by workloads.batch_config
    (seed 0), diffusion 1e-5, viscosity 1e-6, no sources."""
    import fs2d_amd
    from fs2d_amd import workloads as W

    cfg = W.batch_config(count, seed=0)
    sims = [(R, K, dt, 1e-5, 1e-6) for R, K, dt in cfg]
    with fs2d_amd.Batch(sims, launch_mode=LAUNCH_MODE) as b:
        for k, (R, K, dt) in enumerate(cfg):
            b.upload(k, *W.batch_state(R, seed=0, sim_index=k))
            b.set_sources(k)
        b.step_n(warmup)
        b.sync()
        t0 = time.perf_counter()
        b.step_n(steps)
        b.sync()
        sec = (time.perf_counter() - t0) / steps
        info = b.info()
    cells = info["interior_cells"]
    return {"simulations": count, "interior_cells": cells, "steps": steps, "value": round(1.0 / sec, 2),
            "unit": "steps/s (all simulations)", "ms_per_step": round(sec * 1e3, 3),
            "cell_updates_per_s": round(cells / sec, 1), "jacobi_depth": info["jacobi_depth"],
            "tile_rows": info["tile_rows"], "launches_per_step": info["launches_per_step"],
            "persistent_blocks_per_cu": {"jacobi": info["jacobi_blocks_per_cu"], "stages": info["stage_blocks_per_cu"],
                                         "note": "4 waves per block; 0 = one block per work entry (tuned)"},
            "graph_replayed": info["graphs"] > 0,
            "workload": "sides uniform over {64,96,128,192,256,384,512}, dt ~ U[0.02,0.1], K in {20,40}, "
                        "16-vortex starts, diffusion 1e-5, viscosity 1e-6, Jacobi order"}


def run_learned_guess():
    """Config 5, measured at equal residual: the first projection of a step of
    the seeded 64-vortex flow, from identical states, solved by plain Jacobi
    from the previous pressure for a fixed sweep budget, and by Jacobi from the
    network's guess until it reaches the same compact residual
    ||d - L p|| / ||d||; zero start beside them; the multigrid seed (V-cycles
    from zero, then Jacobi) to the same residual, and V-cycles alone towards
    1e-4.  Plus the network's forward time and MFMA utilisation at 1024^2."""
    import fs2d_amd  # noqa: F401  (library first)
    from fs2d_amd import pressure_net as PN

    if not os.path.exists(PN.DEFAULT_PARAMS):
        return {"status": "no trained parameters; run PYTHONPATH=fluid-2d_amd python -m fs2d_amd.train_pressure_net"}
    net = PN.PressureNet(seed=0)
    net.load(PN.DEFAULT_PARAMS)
    out = {"criterion": "compact Poisson residual ||d - L p||_2 / ||d||_2",
           "target_1e-4": "below float32's floor on these states: the exact (DCT, float64) pressure rounded "
                          "to float32 has relative compact residual 1.6e-3 at 1024^2 and 7.7e-3 at 2048^2, so "
                          "the comparison is at the residual plain Jacobi reaches in the budget"}
    fwd_ms = {R: net.time_forward(1, R - 2, R - 2, reps=10) for R in (1026, 2050)}  # counted against its seed
    for R, budget in ((1026, 16384), (2050, 16384)):
        r = PN.equal_residual_race(R, net, budget=budget)
        fwd = fwd_ms[R]
        plain, learned = r["previous_pressure"], r["network"]
        plain_rel = plain["rel_residual"]
        out[f"n{R - 2}"] = {
            "budget_sweeps": budget, "target_rel_residual": round(r["target_rel_residual"], 6),
            "zero_start": {"rel_residual": round(r["zero"]["rel_residual"], 6), "ms": round(r["zero"]["ms"], 3)},
            "plain_jacobi_from_previous": {"sweeps": plain["sweeps"], "ms": round(plain["ms"], 3)},
            "learned_guess": {"sweeps": learned["sweeps"], "ms": round(learned["ms"], 3),
                              "reached": learned["reached"], "rel_residual": round(learned["rel_residual"], 6)},
            "speedup": round(plain["ms"] / max(learned["ms"], 1e-9), 3)}
        b = r["local_seed_bound"]
        per_sweep = plain["ms"] / max(plain["sweeps"], 1)
        out[f"n{R - 2}"]["local_seed_bound"] = {
            "model": "DCT modes of the pressure equation, float64: Jacobi damps mode (k,l) by "
                     "(cos(pi k/n)+cos(pi l/n))/2 per sweep; the ideal seed of window side W equals the exact "
                     "pressure in every mode of wavelength < W and the previous pressure elsewhere",
            "model_residual_from_previous": round(b["target_rel_residual"], 6),
            "measured_residual_from_previous": round(plain_rel, 6),
            "model_residual_from_zero": round(b["model_from_zero"], 6),
            "measured_residual_from_zero": round(r["zero"]["rel_residual"], 6),
            "network_receptive_field": PN.RECEPTIVE_FIELD,
            "windows": {str(w): {"sweeps": x["sweeps"],
                                 "sweep_speedup_bound": (round(x["speedup_bound"], 3) if x["speedup_bound"]
                                                         else "seed alone meets the target"),
                                 "time_speedup_bound_with_forward": round(
                                     plain["ms"] / (x["sweeps"] * per_sweep + fwd), 3)}
                        for w, x in b["windows"].items()}}
        mr = r["multigrid_local_refinement"]
        per_cycle = mr["ms_two_cycles"] - mr["ms_one_cycle"]
        refine = {
            "question": "a network refining one V-cycle's pressure: the V-cycles an ideal correction of every "
                        "wavelength below W could replace, against its forward time",
            "one_cycle_residual": {"measured": round(mr["measured_one_cycle"], 6),
                                   "model": round(mr["model_one_cycle"], 6)},
            "two_cycle_residual": round(mr["measured_two_cycles"], 7),
            "cycle_factor": round(mr["cycle_factor"], 3) if mr["cycle_factor"] else None,
            "ms_per_extra_cycle": round(per_cycle, 3), "network_forward_ms": round(fwd, 3),
            "windows": {str(w): {"residual_after_ideal_correction": round(x["residual_after_ideal_correction"], 6),
                                 "v_cycles_replaced_at_most": (round(x["v_cycles_replaced_at_most"], 3)
                                                               if x["v_cycles_replaced_at_most"] is not None else None),
                                 "ms_saved_at_most": (round(x["v_cycles_replaced_at_most"] * per_cycle, 3)
                                                      if x["v_cycles_replaced_at_most"] is not None else None)}
                        for w, x in mr["windows"].items()}}
        mg, mg4 = r["multigrid"], r["multigrid_to_1e-4"]
        out[f"n{R - 2}"]["multigrid_seed"] = {
            "sweeps": mg["sweeps"], "ms": round(mg["ms"], 3), "reached": mg["reached"],
            "rel_residual": round(mg["rel_residual"], 7), "speedup": round(plain["ms"] / max(mg["ms"], 1e-9), 2),
            "v_cycles_toward_1e-4": {"rel_residual": round(mg4["rel_residual"], 7), "ms": round(mg4["ms"], 3),
                                     "reached": mg4["reached"]},
            "local_refinement_bound": refine}
    ms = net.time_forward(1, 1024, 1024, reps=10)
    fl = net.forward_flops(1, 1024, 1024)
    # useful work: the first layer reads 2 real input channels of its 16 and the
    # output layer writes 1 real channel of its 16 (the rest is MFMA padding)
    levels = [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0]
    real = [(2 if i == 0 else ci, 1 if i == len(PN.LAYERS) - 1 else co) for i, (ci, co) in enumerate(PN.LAYERS)]
    useful = sum(2.0 * 9 * ci * co * 1024 * 1024 / 4 ** lv for (ci, co), lv in zip(real, levels))
    out["network_forward_1024"] = {"ms": round(ms, 4), "tflops": round(fl / (ms * 1e-3) / 1e12, 2),
                                   "mfma_frac_bf16_dense": round(fl / (ms * 1e-3) / 2.5e15, 5), "flops": fl,
                                   "useful_flops": useful,
                                   "useful_tflops": round(useful / (ms * 1e-3) / 1e12, 2),
                                   "useful_frac_bf16_dense": round(useful / (ms * 1e-3) / 2.5e15, 5),
                                   "note": "flops count the padded channels the MFMA units execute; useful_flops "
                                           "the real ones (2 input channels, 1 output channel)"}
    net.close()
    return out


def cpu_baseline(n: int, kd: int, kp: int, params, sample_n: int):
    """The oracle (oracle/fs2d_oracle.c, a bit-exact restatement of the
    reference CPU solver, single-threaded like the reference) timed on one full
    step of the same workload at sample_n^2 interior cells, scaled to n^2 by
    the cell count (every stage is linear in the cells; 4096^2 / 2048^2 costs
    3.98x on the build host)."""
    import numpy as np

    import oracle
    from fs2d_amd import workloads as W

    O = oracle.Oracle()
    Rs = sample_n + 2
    diff, visc, dt = params
    rho, u, v = W.vortex_state(Rs, seed=3)
    z = np.zeros((Rs, Rs), np.float32)
    # pinned to one host core for the sample (the reference solver is single-threaded)
    cores = os.sched_getaffinity(0)
    pin = min(cores)
    os.sched_setaffinity(0, {pin})
    try:
        t0 = time.perf_counter()
        O.solve(rho, z, diff, u, v, z, z, visc, dt, kd, kp, oracle.GS_LEX)
        sec = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, cores)
    scale = (n * n) / float(sample_n * sample_n)
    cpu_model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": round(1.0 / (sec * scale), 6), "unit": "steps/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model, "host_cpus": os.cpu_count(), "pinned_core": pin,
            "scaled": scale != 1.0,
            "sample": f"one full step at {sample_n}x{sample_n}+ring (K_d=K_p={kd}, same dt/viscosity/diffusion, "
                      f"seeded 64-vortex start) took {sec:.2f} s on host core {pin} (affinity pinned); "
                      + (f"SCALED x{scale:g} by cell count to {n}x{n} (every stage is linear in the cells), "
                         if scale != 1.0 else "")
                      + "oracle/fs2d_oracle.c, bit-identical to the reference CPU solver"}


def run_render(R: int, state, reps: int = 20) -> dict:
    """Visual output of the resident state (include/fs2d_render.h), R x R cells:
    density -> RGBA8 (4 B read + 4 B written per cell) and velocity -> line
    vertices (40 B written per cell, u and v read on every 8th row and column),
    each into device memory, timed with HIP events on the handle's stream."""
    import fs2d_amd

    with fs2d_amd.Solver(R, launch_mode=LAUNCH_MODE) as s:
        s.upload(*state)
        ms_d, ms_v = s.time_render(reps)
    cells = R * R
    bd = 8 * cells
    bv = 40 * cells + 8 * ((R + 7) // 8) ** 2
    gd, gv = bd / (ms_d * 1e-3) / 1e9, bv / (ms_v * 1e-3) / 1e9
    return {"density_rgba8": {"ms": round(ms_d, 4), "gbs_algorithmic": round(gd, 1),
                              "roofline_frac": round(gd / HBM_PEAK_GBS, 4), "kernel": "k_render_density"},
            "velocity_lines": {"ms": round(ms_v, 4), "gbs_algorithmic": round(gv, 1),
                               "roofline_frac": round(gv / HBM_PEAK_GBS, 4), "kernel": "k_render_velocity"},
            "cells": cells}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    # a step's buffer rotation repeats within 12 steps: 15 warm-up steps
    # capture every step graph before the timed region
    ap.add_argument("--warmup", type=int, default=15)
    ap.add_argument("--n", type=int, default=4096, help="interior cells per side")
    ap.add_argument("--iters", type=int, default=80, help="K_d = K_p")
    ap.add_argument("--order", choices=("gs", "jacobi"), default="gs")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-jacobi", action="store_true")
    ap.add_argument("--no-8192", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--launch", choices=("graph", "eager"), default="graph",
                    help="replay captured step graphs (default) or issue every launch")
    ap.add_argument("--gs-fixed", default="",
                    help="COLS,SWEEPS,LAG: fix the Gauss-Seidel launch policy instead of tuning it (for counter "
                         "runs, whose serialised dispatches distort the tuner's timings)")
    ap.add_argument("--cpu-sample-n", type=int, default=2048, help="interior size of the timed CPU sample")
    args = ap.parse_args()
    global LAUNCH_MODE, GS_FIXED
    LAUNCH_MODE = 1 if args.launch == "eager" else 0
    if args.gs_fixed:
        tw, m, lag = (int(x) for x in args.gs_fixed.split(","))
        GS_FIXED = {"gs_tile_cols": tw, "gs_block": m, "gs_lag": lag}
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only; --gpus {args.gpus} is not supported (a single grid would "
              "shard as row strips with a halo exchange per stage, which is not built)", file=sys.stderr)
        return 2

    from fs2d_amd import workloads as W

    n, K = args.n, args.iters
    R = n + 2
    params = (1e-5, 1e-6, 0.05)  # diffusion, viscosity, dt
    state = W.vortex_state(R, seed=3)
    order = 0 if args.order == "gs" else 1

    g = run_gpu(order, n, K, K, args.steps, args.warmup, state, params)
    sec_step, launches = g["sec_step"], g["launches"]
    cells = n * n
    bpc = bytes_per_cell_step(K, K, params[0] != 0, params[1] != 0)
    step_gbs = bpc * cells / sec_step / 1e9
    pol = g.get("policy") or {"diffuse": None, "pressure": None}
    d_fields = (1 if params[0] != 0 else 0) + (2 if params[1] != 0 else 0)
    rl_p = solve_roofline(order, n, K, g["ms_pressure"], "pressure", 1, pol["pressure"])
    if "ms_joint" in g:
        # a Gauss-Seidel step: the joint launches (three diffuses with the first
        # pressure solve pipelined behind them), then the second pressure solve
        dom = solve_roofline(order, n, K, g["ms_joint"], "joint", d_fields + 1, pol["joint"])
        dom["share_of_step"] = round(g["ms_joint"] / (sec_step * 1e3), 4)
        dom["other_solve"] = dict(rl_p, share_of_step=round(g["ms_pressure"] / (sec_step * 1e3), 4))
        dom["separately"] = {"diffuse_ms": round(g["ms_diffuse"], 3), "pressure_ms": round(g["ms_pressure"], 3),
                             "note": "the same solves run one after the other: diffuse + pressure"}
    else:
        rl_d = solve_roofline(order, n, K, g["ms_diffuse"], "diffuse", d_fields, pol["diffuse"])
        # the dominant kernel of a step: one diffuse solve (three fields) against two pressure solves
        per_step = {"diffuse": g["ms_diffuse"], "pressure": 2 * g["ms_pressure"]}
        dom, other = (rl_d, rl_p) if per_step["diffuse"] >= per_step["pressure"] else (rl_p, rl_d)
        dom = dict(dom)
        dom["share_of_step"] = round(max(per_step.values()) / (sec_step * 1e3), 4)
        dom["other_solve"] = dict(other, share_of_step=round(min(per_step.values()) / (sec_step * 1e3), 4))
    out = {
        "metric": f"steps/s (plain stable-fluids step, {n}^2 interior, K={K})",
        "value": round(1.0 / sec_step, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(sec_step * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 64-vortex velocity, smoothed-noise density)",
        "config": {"workload": f"stable-fluids step {n}x{n}+ring, dt=0.05, visc=1e-6, diff=1e-5, "
                               f"K_d=K_p={K}, {'Gauss-Seidel (reference order)' if order == 0 else 'Jacobi'}",
                   "order": args.order, "n": n, "iterations": K, "launch": args.launch,
                   **({"gs_fixed": args.gs_fixed} if args.gs_fixed else {})},
        "cell_updates_per_s": round(cells / sec_step, 1),
        "step_hbm_gbs_algorithmic": round(step_gbs, 1),
        "step_roofline_frac": round(step_gbs / HBM_PEAK_GBS, 4),
        "bytes_per_cell_step": bpc,
        "launches_per_step": launches,
        "stream_copy_gbs": round(fs2d_amd_copy_bandwidth(), 1),
        "roofline": dom,
    }
    if order == 0:
        out["gs_launch_policy"] = dict(pol, chosen="timed on the first solve of each kind (tile columns, sweeps "
                                                   "per unit, lag); identical results for every choice")
    if order == 0 and not args.no_jacobi:
        j = run_gpu(1, n, K, K, args.steps, args.warmup, state, params)
        js = j["sec_step"]
        jg = bpc * cells / js / 1e9
        jp = solve_roofline(1, n, K, j["ms_pressure"], "pressure", 1, None)
        valu = profile_value(1, n, K, "valu_instructions_per_pressure_solve")
        if valu:
            jp["valu_issue_frac"] = round(valu / (j["ms_pressure"] * 1e-3) / VALU_ISSUE_PEAK, 4)
        out["jacobi"] = {"value": round(1.0 / js, 3), "unit": "steps/s", "ms_per_step": round(js * 1e3, 3),
                         "cell_updates_per_s": round(cells / js, 1),
                         "step_gbs_algorithmic": round(jg, 1),
                         "step_algorithmic_over_peak": round(jg / HBM_PEAK_GBS, 3),
                         "pressure_solve": jp, "launches_per_step": j["launches"],
                         "note": "temporally blocked sweeps: algorithmic bytes over the peak is not a roofline "
                                 "fraction; pressure_solve.hbm_frac_counters is"}
    if not args.no_render:
        out["render"] = run_render(R, state)
    del state
    if not args.no_8192 and n == 4096:
        n8, R8 = 8192, 8194
        st8 = W.vortex_state(R8, seed=3)
        big = {}
        for name, o in (("gs", 0), ("jacobi", 1)):
            if o == 1 and args.no_jacobi:
                continue
            sec8 = run_gpu(o, n8, K, K, 5, max(args.warmup, 13), st8, params, time_solve=False)["sec_step"]
            g8 = bpc * n8 * n8 / sec8 / 1e9
            big[name] = {"value": round(1.0 / sec8, 3), "unit": "steps/s", "ms_per_step": round(sec8 * 1e3, 3),
                         "step_gbs_algorithmic": round(g8, 1)}
            if g8 <= HBM_PEAK_GBS:
                big[name]["step_roofline_frac"] = round(g8 / HBM_PEAK_GBS, 4)
            else:
                big[name]["step_algorithmic_over_peak"] = round(g8 / HBM_PEAK_GBS, 3)
        del st8
        out["n8192"] = big
    if not args.no_batch:
        out["batched"] = run_batch(256, max(20, 5 * args.steps), max(args.warmup, 13))
    if not args.no_learned:
        out["learned_guess"] = run_learned_guess()
    if not args.no_cpu:
        cb = cpu_baseline(n, K, K, params, min(args.cpu_sample_n, n))
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = round(out["value"] / cb["value"], 1)
    print(json.dumps(finite(out), allow_nan=False))
    return 0


if __name__ == "__main__":
    sys.exit(main())
