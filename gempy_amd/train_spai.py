"""Train the learned preconditioner on generated co-kriging systems.

    python -m gempy_amd.train_spai [--steps 2000] [--out gempy_amd/spai_weights.npz]

One system per step from ``synthetic.training_system(seed, N)`` with seeds
1000-2999 and N drawn in [2,000, 20,000]; Adam on the local SPAI residual,
computed and differentiated by libgkhip.so on the GPU.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    from gempy_amd import learned
    net = learned.SpaiNetwork(seed=0)
    t0 = time.perf_counter()
    hist = learned.train(net, steps=args.steps, seeds=range(1000, 3000), batch=args.batch, lr=args.lr,
                         log=lambda m: print(m, file=sys.stderr, flush=True))
    net.save(args.out or learned.WEIGHTS)
    print(json.dumps({"steps": args.steps, "seconds": time.perf_counter() - t0,
                      "loss_first_100": float(np.mean(hist[:100])), "loss_last_100": float(np.mean(hist[-100:]))}))


if __name__ == "__main__":
    main()
