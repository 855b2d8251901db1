"""Train the learned lower bound of the batched path on the GPU.

    python -m zolo.train_l0 [--train 8192] [--val 2048] [--steps 3000] [--out zolo/l0_weights.json]

Matrices come from the batched workload generator (``zolo.workloads.batch``),
labels are log10(sigma_min / sigma_max) from an FP64 SVD of the FP32 data,
features and the network's forward/backward/Adam steps run in the library's
own kernels (``l0_features_device``, ``l0_train_device``).  The pinball loss
at quantile tau biases the prediction below the label; the margin is then
set on the validation matrices so that at most ``--over`` of them get a bound
above their true sigma_min.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np


def init_params(rng, mu, sd, n_param):
    P = np.zeros(n_param, dtype=np.float64)
    P[:16] = mu
    P[16:32] = sd
    o = 32
    P[o:o + 64 * 16] = rng.standard_normal(64 * 16) / math.sqrt(16)
    o += 64 * 16 + 64
    P[o:o + 64 * 64] = rng.standard_normal(64 * 64) / math.sqrt(64)
    o += 64 * 64 + 64
    P[o:o + 64] = rng.standard_normal(64) / math.sqrt(64)
    o += 64
    P[o] = -3.0  # b3: start near the middle of the label range
    return P


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=8192)
    ap.add_argument("--val", type=int, default=2048)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--tau", type=float, default=0.05)
    ap.add_argument("--over", type=float, default=0.005, help="allowed fraction of overestimated bounds")
    ap.add_argument("--seed", type=int, default=12345)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)

    import torch
    import zolo as zl
    from zolo import workloads

    def log(msg):
        print(f"[train_l0] {msg}", flush=True)

    t0 = time.time()
    feats, labels = [], []
    for part, (count, seed) in enumerate(((args.train, args.seed), (args.val, args.seed + 1))):
        done = 0
        F, T = [], []
        while done < count:  # chunks bound the memory of the generator
            k = min(2048, count - done)
            mats, _, _ = workloads.batch(k, seed * 1000 + done)
            f, _ = zl.l0_features_device(mats)
            F.append(f.cpu().numpy())
            T.append(workloads.log_cond(mats))
            done += k
            del mats
            torch.cuda.empty_cache()
            log(f"{'train' if part == 0 else 'val'} {done}/{count} ({time.time() - t0:.0f} s)")
        feats.append(np.concatenate(F))
        labels.append(np.concatenate(T))
    Ftr, Fva = feats
    Ttr, Tva = labels
    mu = Ftr.mean(0)
    sd = Ftr.std(0) + 1e-6
    rng = np.random.default_rng(args.seed)
    P = init_params(rng, mu, sd, zl.l0_param_count())
    dP = torch.tensor(P, dtype=torch.float32, device="cuda")
    dF = torch.tensor(Ftr, dtype=torch.float32, device="cuda")
    dT = torch.tensor(Ttr, dtype=torch.float32, device="cuda")
    hist = []
    stages = ((args.steps // 2, args.lr), (args.steps - args.steps // 2, args.lr / 5))
    for steps, lr in stages:
        h = zl.l0_train_device(dP, dF, dT, steps, lr=lr, tau=args.tau)
        hist.extend(h.tolist())
        log(f"{steps} steps at lr {lr:g}: loss {h[0]:.4f} -> {h[-1]:.4f}")
    yv = zl.l0_forward_device(dP, torch.tensor(Fva, dtype=torch.float32, device="cuda")).cpu().numpy()
    resid = yv - Tva
    margin = max(0.0, float(np.quantile(resid, 1.0 - args.over)))
    over = float(np.mean(resid - margin > 0))
    P = dP.cpu().numpy().astype(np.float64)
    P[-1] = margin
    meta = {
        "params": [float(x) for x in P.astype(np.float32)],
        "tau": args.tau, "margin": margin, "train_count": args.train, "val_count": args.val,
        "val_overestimate_rate": over, "val_median_underestimate_decades": float(np.median(margin - resid)),
        "val_mae_decades": float(np.mean(np.abs(resid))), "loss_first": hist[0], "loss_last": hist[-1],
        "steps": args.steps, "lr": args.lr, "seed": args.seed,
        "generator": "zolo.workloads.batch: n ~ U{32..512}, log10 kappa ~ U[1, 12], Haar U, V, s geometric",
    }
    out = args.out or os.path.join(os.path.dirname(os.path.abspath(__file__)), "l0_weights.json")
    with open(out, "w") as f:
        json.dump(meta, f)
    log(f"margin {margin:.3f} decades, validation overestimates {over:.4f}, MAE {meta['val_mae_decades']:.3f}; "
        f"wrote {out} ({time.time() - t0:.0f} s)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
