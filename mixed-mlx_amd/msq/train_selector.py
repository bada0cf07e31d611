"""Trains the learned tier selector on the GPU and saves its weights.

    python mixed-mlx_amd/msq/train_selector.py [--samples 100000] [--out DIR]

Labels come from the solver itself: every choice of msq.selector.CHOICES
(eight tiers and the plain path) is run on seeded synthetic batches
generated on the GPU, and each choice's batched solve is timed.  Two kinds
of batch make up the samples, in equal numbers of matrices' work:

* uneven: n = 16 (1 + floor(16 u)), cond 10^U[1, 5], spectra alternately
  log-uniform and four tight clusters (the batched uneven workload's
  distribution);
* mixture: n uniform in [32, 512], spectra one third each log-uniform (cond
  up to 1e6), power law k^-alpha (alpha in [0.5, 3]) and four clusters (the
  learned-tier workload's distribution).

Bases are Haar-random (QR of a Gaussian matrix in binary64 on the GPU, sign
fixed), A = Q diag(lambda) Q^T and the known root Q diag(sqrt(lambda)) Q^T.
Held-out batches of both kinds (other seeds) measure the selector against
the fixed BF16 tier and the plain path, in modelled time (each batch's own
measured choice costs) and in measured time of the batched solves.  The
network is trained with the library's forward, backward and Adam kernels;
the weights (standardisation folded in) and the per-choice unit costs go to
msq/data/selector.npz (and --out, if given) with the measured latency of a
plain-path fallback solve per order (fallback_latency.py), the labelled
samples to --data, if given.
"""
from __future__ import annotations

import argparse
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from msq import capi, generators as G, selector as SEL, solver as MS  # noqa: E402


def log(*a):
    print(*a, flush=True)


def batch_params(kind, count, seed):
    """Sorted orders and per-matrix spectra of one synthetic batch."""
    rng = np.random.default_rng(seed)
    if kind == "uneven":
        orders = 16 * (1 + rng.integers(0, 16, count))
        specs = [("log_uniform" if i % 2 == 0 else "clustered", 10.0 ** rng.uniform(1, 5), 0.0)
                 for i in range(count)]
    else:
        orders = rng.integers(32, 513, count)
        specs = []
        for _ in range(count):
            k = int(rng.integers(0, 3))
            specs.append((("log_uniform", "power_law", "clustered")[k],
                          10.0 ** (6.0 * rng.uniform()), rng.uniform(0.5, 3.0)))
    order = np.argsort(orders, kind="stable")
    orders = orders[order]
    specs = [specs[i] for i in order]
    lams = []
    for n, (s, cond, alpha) in zip(orders, specs):
        u = rng.uniform(size=2 * int(n) + 4)
        if s == "log_uniform":
            lams.append(G.log_uniform_spectrum(u, int(n), cond))
        elif s == "power_law":
            lams.append(G.power_law_spectrum(int(n), alpha))
        else:
            lams.append(G.clustered_spectrum(u, int(n), cond))
    return [int(n) for n in orders], lams


def gpu_batch(kind, count, seed, dev):
    """A synthetic batch built on the GPU, orders ascending: (orders, A, X)
    with A and the known root X in the batch layout (binary64)."""
    import torch

    orders, lams = batch_params(kind, count, seed)
    lds, offs, _, total, _ = capi.batch_layout(orders)
    A = torch.zeros(total, dtype=torch.float64, device=dev)
    X = torch.zeros(total, dtype=torch.float64, device=dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    i = 0
    while i < count:
        n = orders[i]
        j = i
        while j < count and orders[j] == n and j - i < 256:
            j += 1
        k = j - i
        Gm = torch.randn(k, n, n, generator=gen, dtype=torch.float64, device=dev)
        Q, R = torch.linalg.qr(Gm)
        d = torch.sign(torch.diagonal(R, dim1=1, dim2=2))
        d[d == 0] = 1.0
        Q = Q * d[:, None, :]
        lam = torch.from_numpy(np.stack(lams[i:j])).to(dev)
        Qt = Q.transpose(1, 2).contiguous()
        Am = torch.bmm(Q * lam[:, None, :], Qt)
        Xm = torch.bmm(Q * torch.sqrt(lam)[:, None, :], Qt)
        Am = 0.5 * (Am + Am.transpose(1, 2))
        Xm = 0.5 * (Xm + Xm.transpose(1, 2))
        for t in range(k):
            ld, off = int(lds[i + t]), int(offs[i + t])
            A[off:off + ld * ld].view(ld, ld)[:n, :n] = Am[t]
            X[off:off + ld * ld].view(ld, ld)[:n, :n] = Xm[t]
        i = j
    return orders, A, X


def batch_errors(orders, roots, known):
    """Relative Frobenius error of every root against the known root (orders
    ascending, so each leading dimension's matrices are contiguous)."""
    import torch

    lds, offs, _, _, _ = capi.batch_layout(orders)
    out = torch.empty(len(orders), dtype=torch.float64, device=roots.device)
    i = 0
    while i < len(orders):
        ld = int(lds[i])
        j = i
        while j < len(orders) and int(lds[j]) == ld:
            j += 1
        s, e = int(offs[i]), int(offs[j - 1]) + ld * ld
        d = (roots[s:e] - known[s:e]).view(j - i, ld * ld)
        r = known[s:e].view(j - i, ld * ld)
        out[i:j] = torch.linalg.vector_norm(d, dim=1) / torch.linalg.vector_norm(r, dim=1)
        i = j
    return out.cpu().numpy()


def label(rt, orders, A, Xk, runs, log=None):
    """Every choice on the batch: (features, success, sweeps, corrections,
    choice_ms), success = converged and within 1e-6 of the known root."""
    import torch

    count = len(orders)
    feats = torch.zeros(count * capi.MLP_INPUTS, dtype=torch.float32, device=A.device)
    rt.matrix_features(orders, A, feats)
    X = torch.zeros_like(A)
    nc = len(SEL.CHOICES)
    success = np.zeros((count, nc), dtype=np.float32)
    sweeps = np.zeros((count, nc), dtype=np.float32)
    corrections = np.zeros((count, nc), dtype=np.float32)
    ms = np.zeros(nc)
    for k in range(nc):
        opts = SEL.tier_options(k)
        rt.solve_batched(orders, A, X, opts)
        times = []
        for _ in range(runs):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = rt.solve_batched(orders, A, X, opts)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        ms[k] = float(np.median(times))
        err = batch_errors(orders, X, Xk)
        success[:, k] = reps.converged & (err <= MS.TARGET)
        sweeps[:, k] = reps.sweeps
        corrections[:, k] = reps.corrections
        if log:
            log(f"    {SEL.CHOICES[k]}: success {success[:, k].mean():.3f}, sweeps "
                f"{sweeps[:, k].mean():.2f}, corrections {corrections[:, k].mean():.2f}, "
                f"{ms[k]:.1f} ms")
    return (feats.cpu().numpy().reshape(count, capi.MLP_INPUTS), success, sweeps, corrections,
            ms)


def timed(fn, reps=3):
    import torch

    fn()
    out = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        out.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--samples", type=int, default=100000)
    ap.add_argument("--out", default=None)
    ap.add_argument("--data", default=None, help="npz file for the labelled samples")
    ap.add_argument("--epochs", type=int, default=200)
    args = ap.parse_args()

    import torch

    dev = torch.device("cuda")
    torch.cuda.set_device(0)
    rt = capi.Runtime(0)
    t0 = time.time()
    # Chunks: uneven 4096 matrices, mixture 5000, alternating, until the
    # sample count is reached.
    chunks = []
    total = 0
    seed = 1000
    while total < args.samples:
        kind = "uneven" if len(chunks) % 2 == 0 else "mixture"
        count = min(4096 if kind == "uneven" else 5000, args.samples - total)
        chunks.append((kind, count, seed))
        total += count
        seed += 1
    parts = []
    for c, (kind, count, sd) in enumerate(chunks):
        orders, A, Xk = gpu_batch(kind, count, sd, dev)
        log(f"chunk {c + 1}/{len(chunks)} {kind} seed {sd}: {count} matrices "
            f"({time.time() - t0:.0f} s)")
        f, y, sw, co, ms = label(rt, orders, A, Xk, runs=3 if kind == "uneven" else 1, log=log)
        parts.append((kind, np.array(orders), f, y, sw, co, ms))
        del A, Xk
        torch.cuda.empty_cache()
    x = np.concatenate([p[2] for p in parts])
    y = np.concatenate([p[3] for p in parts])
    orders_all = np.concatenate([p[1] for p in parts])
    log(f"{len(x)} labelled samples ({time.time() - t0:.0f} s); success rate per choice: "
        + " ".join(f"{n}={v:.3f}" for n, v in zip(SEL.CHOICES, y.mean(axis=0))))
    if args.data:
        np.savez(args.data, x=x, y=y, orders=orders_all,
                 kinds=np.array([p[0] for p in parts]),
                 chunk_sizes=np.array([len(p[1]) for p in parts]),
                 choice_ms=np.array([p[6] for p in parts]))
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    std[std < 1e-6] = 1.0
    params, hist = SEL.train(rt, (x - mean) / std, y[:, :SEL.PLAIN], epochs=args.epochs, log=log)
    log(f"trained: loss {hist[0]:.4f} -> {hist[-1]:.4f} ({time.time() - t0:.0f} s)")
    # Unit cost of each choice, pooled over the batches: wall time over work.
    choice_ms = np.sum([p[6] for p in parts], axis=0)
    work = float(np.sum([np.sum(SEL.work_units(p[1])) for p in parts]))
    cost = choice_ms / work
    log("unit costs (ms per (n/256)^3): "
        + " ".join(f"{n}={c:.4f}" for n, c in zip(SEL.CHOICES, cost)))
    from msq import fallback_latency as FL
    lat_orders, lat_ms = FL.measure(rt)
    sel = SEL.Selector(SEL.fold_standardisation(params, mean, std), cost, lat_orders, lat_ms)

    # Held-out batches: modelled and measured time of the selector's choices
    # against the fixed BF16 tier and the plain path.
    held = []
    for kind, count, sd in (("uneven", 4096, 2000), ("mixture", 3000, 2001)):
        orders, A, Xk = gpu_batch(kind, count, sd, dev)
        log(f"held-out {kind} seed {sd}: {count} matrices")
        f, yk, _, _, ms = label(rt, orders, A, Xk, runs=1, log=log)
        prob = sel.probabilities(rt, orders, A)
        per_matrix = sel.choose(prob, orders)
        batch_choice = sel.choose_batch(prob, orders)
        uc = SEL.unit_costs(ms, orders)
        c_sel, f_sel = SEL.policy_cost(batch_choice, yk, uc, orders, sel.latency)
        c_pm, f_pm = SEL.policy_cost(per_matrix, yk, uc, orders, sel.latency)
        c_fix, f_fix = SEL.policy_cost(np.full(count, SEL.FIXED), yk, uc, orders, sel.latency)
        c_plain, _ = SEL.policy_cost(np.full(count, SEL.PLAIN), yk, uc, orders, sel.latency)
        w = SEL.work_units(orders)
        best = float(np.sum(w * np.where(yk > 0, uc[None, :], np.inf).min(axis=1)))
        acc = float(np.mean((prob > 0.5) == (yk[:, :SEL.PLAIN] > 0.5)))
        X = torch.zeros_like(A)
        t_sel = timed(lambda: SEL.solve_selected(rt, sel, orders, A, X))
        _, fb_sel, _ = SEL.solve_selected(rt, sel, orders, A, X)
        err_sel = batch_errors(orders, X, Xk)
        t_fix = timed(lambda: MS.solve_with_fallback(rt, orders, A, X, SEL.tier_options(SEL.FIXED)))
        t_plain = timed(lambda: MS.solve_with_fallback(rt, orders, A, X, SEL.tier_options(SEL.PLAIN)))
        log(f"  prediction accuracy {acc:.3f}; modelled ms: batch choice {c_sel:.1f} "
            f"({f_sel} fallbacks), per-matrix {c_pm:.1f} ({f_pm}), fixed {c_fix:.1f} ({f_fix}), "
            f"plain {c_plain:.1f}, best possible {best:.1f}")
        log(f"  measured ms: selector {t_sel:.1f} ({len(fb_sel)} fallbacks, worst error "
            f"{err_sel.max():.2e}), fixed {t_fix:.1f}, plain {t_plain:.1f}")
        log("  choices: " + " ".join(f"{n}={c}" for n, c in zip(
            SEL.CHOICES, np.bincount(batch_choice, minlength=len(SEL.CHOICES)))))
        held.append([c_sel, c_pm, c_fix, c_plain, best, t_sel, t_fix, t_plain, acc])
        del A, Xk, X
        torch.cuda.empty_cache()
    extra = dict(feature_mean=mean, feature_std=std, loss_history=np.array(hist),
                 heldout=np.array(held), samples=np.array(len(x)))
    sel.save(**extra)
    if args.out:
        sel.save(pathlib.Path(args.out) / "selector.npz", **extra)
    log(f"saved {SEL.WEIGHTS}")
    rt.close()


if __name__ == "__main__":
    main()
