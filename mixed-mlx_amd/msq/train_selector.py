"""Trains the learned tier selector on the GPU and saves its weights.

    python mixed-mlx_amd/msq/train_selector.py [--out DIR] [--count 4096]

Labels come from the solver itself: every tier of msq.selector.TIERS is run
on seeded synthetic batches (batched uneven workloads with seeds 10 and 11
and a mixture batch with orders 32..256, spectra log-uniform / power law /
clustered).  A held-out batch (seed 12) measures what the selector gains over
the fixed BF16 tier.  The network is trained with the library's forward,
backward and Adam kernels; the weights (standardisation folded in) and the
per-tier costs go to msq/data/selector.npz (and --out, if given).
"""
from __future__ import annotations

import argparse
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from msq import capi, generators as G, selector as SEL  # noqa: E402


def log(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--epochs", type=int, default=300)
    args = ap.parse_args()

    import torch

    rt = capi.Runtime(0)
    t0 = time.time()
    data = []
    sources = [("uneven seed 10", lambda: G.batched_uneven(count=args.count, config=10)),
               ("uneven seed 11", lambda: G.batched_uneven(count=args.count, config=11)),
               ("mixture 32..256", lambda: list(G.mixture(count=args.count, config=4,
                                                           order_range=(32, 256)))),
               ("uneven seed 12 (held out)", lambda: G.batched_uneven(count=args.count,
                                                                      config=12))]
    for name, make in sources:
        mats = make()
        log(f"{name}: {len(mats)} matrices generated ({time.time() - t0:.0f} s)")
        data.append(SEL.label_batch(rt, mats, log=log))
        del mats
        log(f"{name}: labelled ({time.time() - t0:.0f} s)")

    train_parts, held = data[:-1], data[-1]
    x = np.concatenate([d[0] for d in train_parts])
    y = np.concatenate([d[1] for d in train_parts])
    sw = np.concatenate([d[2] for d in train_parts])
    co = np.concatenate([d[3] for d in train_parts])
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    std[std < 1e-6] = 1.0
    xs = (x - mean) / std
    params, hist = SEL.train(rt, xs, y, epochs=args.epochs, log=log)
    log(f"trained: loss {hist[0]:.4f} -> {hist[-1]:.4f} ({time.time() - t0:.0f} s)")
    # Mean cost of each tier per matrix, attempts that failed included.
    tier_cost = SEL.cost_units(sw, co).mean(axis=0)
    sel = SEL.Selector(SEL.fold_standardisation(params, mean, std), tier_cost)

    # Held-out batch: selector vs the fixed BF16 tier vs the best choice.
    hx, hy, hsw, hco, hord = held
    dev = torch.device("cuda")
    P = torch.from_numpy(sel.params).to(dev)
    X = torch.from_numpy(np.ascontiguousarray(hx, dtype=np.float32)).to(dev)
    logits = torch.zeros(hx.shape[0] * capi.MLP_OUTPUTS, device=dev)
    rt.mlp_forward(P, X, logits)
    prob = 1.0 / (1.0 + np.exp(-logits.cpu().numpy().reshape(-1, capi.MLP_OUTPUTS)))
    choice = sel.choose(prob)
    c_sel, f_sel = SEL.policy_cost(choice, hy, hsw, hco, hord)
    c_fix, f_fix = SEL.policy_cost(np.full(len(choice), 1), hy, hsw, hco, hord)
    cost = SEL.cost_units(hsw, hco) * (hord[:, None] / 256.0) ** 3
    plain = cost[:, 0]
    best = np.where(hy > 0, cost, np.inf).min(axis=1)
    acc = float(np.mean((prob > 0.5) == (hy > 0.5)))
    log(f"held out: prediction accuracy {acc:.3f}; cost units selector {c_sel:.0f} "
        f"({f_sel} fallbacks), fixed bf16_r0.5 {c_fix:.0f} ({f_fix} fallbacks), plain "
        f"{plain.sum():.0f}, best possible {best.sum():.0f} (sweep units x (n/256)^3)")
    log("tier choice histogram:", np.bincount(choice, minlength=len(SEL.TIERS)).tolist())
    extra = dict(feature_mean=mean, feature_std=std, loss_history=np.array(hist),
                 heldout_cost=np.array([c_sel, c_fix, plain.sum(), best.sum()]))
    sel.save(**extra)
    if args.out:
        sel.save(pathlib.Path(args.out) / "selector.npz", **extra)
    log(f"saved {SEL.WEIGHTS}")
    rt.close()


if __name__ == "__main__":
    main()
