"""Train the learned warm start's U-Net on the GPU and write its weights.

    python -m citcom_amd.train_warm --out citcom_amd/weights/warm_unet.npz

Samples come from the solver itself: time loops of the 64^3-element
workload (tdep64: 65^3 nodes, 5 levels) at Rayleigh numbers and viscosity
contrasts drawn log-uniformly from [1e5, 1e7] and [1e2, 1e5], each solve
to the workload's own criterion with the extrapolated warm start; every
solve with three earlier ones gives one (input, target) pair (unet.py,
chip_warm_record).  Training runs on the project's BF16 MFMA kernels with
Adam; held-out cases measure the loss against the extrapolation's
(the mean square of the targets).  The weights, the loss history and the
sample statistics are written to --out.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(msg):
    print(f"[train_warm] {msg}", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--base", default="tdep64")
    ap.add_argument("--cases", type=int, default=25)
    ap.add_argument("--steps", type=int, default=80, help="recorded samples per case")
    ap.add_argument("--holdout-cases", type=int, default=2)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--time-budget", type=float, default=0.0, help="stop training after this many seconds (0: none)")
    ap.add_argument("--out", required=True)
    args = ap.parse_args(argv)

    import torch

    from citcom_amd import unet
    t_start = time.perf_counter()
    cases = unet.draw_cases(args.cases + args.holdout_cases, seed=args.seed)
    train_cases, hold_cases = cases[:args.cases], cases[args.cases:]
    t0 = time.perf_counter()
    feats, targs, info = unet.make_samples(args.base, train_cases, args.steps, log=log)
    t_data = time.perf_counter() - t0
    hold = None
    hold_info = []
    if hold_cases:
        hf, ht, hold_info = unet.make_samples(args.base, hold_cases, args.steps // 2, log=log)
        hold = (hf, ht)
    p = unet.problem_for(args.base, 1e6, 1e4)
    dims = p.dims()
    log(f"{feats.shape[0]} samples at {dims} nodes in {t_data:.0f} s; device memory "
        f"{torch.cuda.memory_allocated() / 2**30:.1f} GiB")
    wch = (1.0, 1.0, 1.0, 2.0)
    base_hold = None
    if hold is not None:
        w = torch.tensor(wch, dtype=torch.float32, device="cuda")
        base_hold = float((hold[1] ** 2 * w).mean())
        log(f"held-out: the extrapolation alone scores {base_hold:.4f}")
    net = unet.UNet(seed=args.seed)
    t0 = time.perf_counter()
    hist = []
    ep = 0
    while ep < args.epochs:
        h = unet.train(net, feats, targs, dims, epochs=1, batch=args.batch,
                       lr=args.lr * (0.05 + 0.95 * 0.5 * (1 + np.cos(np.pi * ep / max(1, args.epochs)))),
                       wch=wch, seed=args.seed + ep, log=lambda m: None, holdout=hold, cosine=False)
        rec = dict(h[0], epoch=ep)
        hist.append(rec)
        log(f"epoch {ep}: loss {rec['loss']:.4f}" + (f", held-out {rec['holdout']:.4f}" if hold else "") +
            f" ({rec['s']:.1f} s)")
        ep += 1
        if args.time_budget and time.perf_counter() - t_start > args.time_budget:
            log(f"time budget reached after {ep} epochs")
            break
    t_train = time.perf_counter() - t0
    nw, nb = unet.param_counts()
    meta = {"base": args.base, "dims": list(dims), "train_cases": train_cases, "holdout_cases": hold_cases,
            "samples": int(feats.shape[0]), "holdout_samples": int(hold[0].shape[0]) if hold else 0,
            "data_s": t_data, "train_s": t_train, "epochs": ep, "batch": args.batch, "lr": args.lr,
            "loss_weights": wch, "history": hist, "holdout_extrapolation_score": base_hold,
            "params": nw + nb, "sample_iterations": [i["iterations"] for i in info],
            "holdout_iterations": [i["iterations"] for i in hold_info]}
    unet.save(net, args.out, meta)
    log(f"wrote {args.out}: {nw + nb} parameters, data {t_data:.0f} s, training {t_train:.0f} s")
    print(json.dumps({k: meta[k] for k in ("samples", "data_s", "train_s", "epochs", "holdout_extrapolation_score")}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
