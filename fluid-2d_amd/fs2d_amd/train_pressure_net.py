"""Train the pressure network and store its parameters.

    PYTHONPATH=fluid-2d_amd python -m fs2d_amd.train_pressure_net [--out PATH] [--seeds 10] [--epochs 4]

Training data: seeded 64-vortex runs at 1024 x 1024 (the GPU solver, Jacobi
order, K = 40); at steps 10, 20, ..., 500 both projections of steps k-1 and k
are solved for 20000 sweeps and each projection of step k gives a pair
(divergence, the handle's previous pressure) labelled by the exact DCT solve
of the pressure equation; 256 x 256 crops, batch 16, Adam with lr 1e-3.  Run
on a machine with the GPU; the parameters go to --out (.npz).
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

from . import pressure_net as PN

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "pressure_net_params.npz")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=DEFAULT_OUT)
    ap.add_argument("--seeds", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--crops-per-pair", type=int, default=20)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--max-seconds", type=float, default=600.0, help="stop training after this long")
    args = ap.parse_args()

    def log(msg):
        print(f"[{time.strftime('%H:%M:%S')}] {msg}", flush=True)

    t0 = time.time()
    steps = list(range(10, 501, 10))
    data = PN.make_pairs(args.n + 2, seeds=range(100, 100 + args.seeds), steps=steps, crop=args.crop,
                         crops_per_pair=args.crops_per_pair, log=log)
    log(f"{len(data[0])} crops of {args.crop}^2 in {time.time() - t0:.0f} s")
    net = PN.PressureNet(seed=0)
    t1 = time.time()
    rng = np.random.default_rng(0)
    D, PV, T = data
    batch, step = 16, 0
    losses = []
    for ep in range(args.epochs):
        order = rng.permutation(len(D))
        tot, cnt = 0.0, 0
        for b0 in range(0, len(order) - batch + 1, batch):
            idx = np.sort(order[b0:b0 + batch])
            tot += net.train_step(D[idx], PV[idx], T[idx], args.lr)
            cnt += 1
            step += 1
            if step % 200 == 0:
                log(f"epoch {ep} step {step}: mean loss {tot / cnt:.4e}")
            if time.time() - t1 > args.max_seconds:
                break
        losses.append(tot / max(cnt, 1))
        log(f"epoch {ep}: mean loss {losses[-1]:.4e} ({time.time() - t1:.0f} s)")
        if time.time() - t1 > args.max_seconds:
            break
    net.save(args.out)
    log(f"saved {args.out} after {step} steps")
    return 0


if __name__ == "__main__":
    sys.exit(main())
