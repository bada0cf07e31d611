"""Trains the learned tier selector on the GPU and saves its weights.

    python mixed-mlx_amd/msq/train_selector.py [--out DIR] [--count 4096]

Labels come from the solver itself: every tier of msq.selector.TIERS is run
on seeded synthetic batches (batched uneven workloads with seeds 10 and 11,
and mixture batches with orders 32..256 and 32..512, spectra log-uniform /
power law / clustered), and each tier's batched solve is timed.  Two
held-out batches (uneven seed 12, mixture 32..512 seed 6) measure what the
selector gains over the fixed BF16 tier.  The network is trained with the
library's forward, backward and Adam kernels; the weights (standardisation
folded in) and the per-tier unit costs go to msq/data/selector.npz (and
--out, if given).
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
    train_sources = [
        ("uneven seed 10", lambda: G.batched_uneven(count=args.count, config=10)),
        ("uneven seed 11", lambda: G.batched_uneven(count=args.count, config=11)),
        ("mixture 32..256", lambda: list(G.mixture(count=args.count, config=4,
                                                    order_range=(32, 256)))),
        ("mixture 32..512", lambda: list(G.mixture(count=args.count // 2, config=5,
                                                    order_range=(32, 512)))),
    ]
    held_sources = [
        ("uneven seed 12 (held out)", lambda: G.batched_uneven(count=args.count, config=12)),
        ("mixture 32..512 seed 6 (held out)",
         lambda: list(G.mixture(count=args.count // 4, config=6, order_range=(32, 512)))),
    ]

    def label(sources):
        out = []
        for name, make in sources:
            mats = make()
            log(f"{name}: {len(mats)} matrices generated ({time.time() - t0:.0f} s)")
            out.append(SEL.label_batch(rt, mats, log=log))
            del mats
            log(f"{name}: labelled ({time.time() - t0:.0f} s)")
        return out

    train_parts = label(train_sources)
    held_parts = label(held_sources)
    x = np.concatenate([d[0] for d in train_parts])
    y = np.concatenate([d[1] for d in train_parts])
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    std[std < 1e-6] = 1.0
    xs = (x - mean) / std
    params, hist = SEL.train(rt, xs, y, epochs=args.epochs, log=log)
    log(f"trained: loss {hist[0]:.4f} -> {hist[-1]:.4f} ({time.time() - t0:.0f} s)")
    # Unit cost of each tier, pooled over the training batches: total wall
    # time over total work.
    tier_ms = np.sum([d[5] for d in train_parts], axis=0)
    work = float(np.sum([np.sum(SEL.work_units(d[4])) for d in train_parts]))
    tier_cost = tier_ms / work
    log("tier unit costs (ms per (n/256)^3):",
        " ".join(f"{t[0]}={c:.4f}" for t, c in zip(SEL.TIERS, tier_cost)))
    sel = SEL.Selector(SEL.fold_standardisation(params, mean, std), tier_cost)

    # Held-out batches: selector vs the fixed BF16 tier vs plain vs the best
    # choice, in the modelled time of each batch's own measured tier costs.
    dev = torch.device("cuda")
    P = torch.from_numpy(sel.params).to(dev)
    held_cost = []
    for (name, _), (hx, hy, hsw, hco, hord, hms) in zip(held_sources, held_parts):
        X = torch.from_numpy(np.ascontiguousarray(hx, dtype=np.float32)).to(dev)
        logits = torch.zeros(hx.shape[0] * capi.MLP_OUTPUTS, device=dev)
        rt.mlp_forward(P, X, logits)
        prob = 1.0 / (1.0 + np.exp(-logits.cpu().numpy().reshape(-1, capi.MLP_OUTPUTS)))
        choice = sel.choose(prob)
        uc = SEL.unit_costs(hms, hord)
        c_sel, f_sel = SEL.policy_cost(choice, hy, uc, hord)
        c_fix, f_fix = SEL.policy_cost(np.full(len(choice), 1), hy, uc, hord)
        c_plain, _ = SEL.policy_cost(np.zeros(len(choice), int), hy, uc, hord)
        w = SEL.work_units(hord)
        # Best choice in hindsight: the cheapest tier that met the target
        # (the plain path where none did).
        best = float(np.sum(w * np.where(hy > 0, uc[None, :], np.inf).min(axis=1, initial=uc[0])))
        acc = float(np.mean((prob > 0.5) == (hy > 0.5)))
        log(f"{name}: prediction accuracy {acc:.3f}; modelled ms: selector {c_sel:.1f} "
            f"({f_sel} fallbacks), fixed bf16_r0.5 {c_fix:.1f} ({f_fix} fallbacks), plain "
            f"{c_plain:.1f}, best possible {best:.1f}")
        log("  tier choice histogram:", np.bincount(choice, minlength=len(SEL.TIERS)).tolist())
        held_cost.append([c_sel, c_fix, c_plain, best])
    extra = dict(feature_mean=mean, feature_std=std, loss_history=np.array(hist),
                 heldout_cost=np.array(held_cost))
    sel.save(**extra)
    if args.out:
        sel.save(pathlib.Path(args.out) / "selector.npz", **extra)
    log(f"saved {SEL.WEIGHTS}")
    rt.close()


if __name__ == "__main__":
    main()
