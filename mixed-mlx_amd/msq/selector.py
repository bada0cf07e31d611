"""Learned tier selector of the batched SPD square-root solver.

No counterpart in the reference.  For each matrix of a batch the selector
chooses one of eight solver configurations ("tiers"): the plain FP32 path,
which always meets the target and is the fallback, and seven low-bit tiers
that stop Jacobi early (``correction_rate``) and leave the rest of the
accuracy to the Sylvester corrections.  A low-bit tier that misses the
binary64 residual check is re-solved on the plain path, so a wrong choice
costs time, never accuracy.

* features: ``Runtime.matrix_features`` (HIP kernel; 16 per matrix);
* network: 16 -> 64 -> 64 -> 8 ReLU MLP, output k = logit of "tier k meets
  the target within its correction budget"; forward, backward and Adam are the
  library's HIP kernels (``Runtime.mlp_*``);
* labels: the solver's own runs of every tier on seeded synthetic batches
  (``label_batch``), success = converged and relative Frobenius error against
  the construction-known root <= 1e-6;
* decision: the tier of least expected cost c_t + (1 - p_t) F c_plain, with
  p_t the predicted success and c_t the tier's measured cost: the wall time of
  the tier's batched solve of each labelled batch divided by the batch's work
  sum (n / 256)^3, pooled over the batches (milliseconds per unit of work; a
  matrix of order n costs c_t (n / 256)^3, so the factor cancels in the
  choice).  A failed choice pays the plain path again, in a second solve call
  whose latency is that of its slowest matrix rather than the matrix's share
  of a batch: F = FALLBACK_WEIGHT weights it accordingly, so a low-bit tier is
  chosen only when its predicted success is high.

Feature standardisation is folded into the first layer when the weights are
saved, so inference runs the raw features through the network unchanged.
"""
from __future__ import annotations

import pathlib

import numpy as np

from . import capi
from . import solver as MS

WEIGHTS = pathlib.Path(__file__).resolve().parent / "data" / "selector.npz"
FALLBACK_WEIGHT = 10.0

# (name, basis/correction tier, correction_rate); tier 0 is the plain path.
TIERS = [
    ("plain_fp32", None, 0.0),
    ("bf16_r0.5", "bf16", 0.5),
    ("bf16_r1", "bf16", 1.0),
    ("bf16_r2", "bf16", 2.0),
    ("bf16_r4", "bf16", 4.0),
    ("bf16_r8", "bf16", 8.0),
    ("fp8_r1", "fp8", 1.0),
    ("mxfp4_r1", "mxfp4", 1.0),
]
assert len(TIERS) == capi.MLP_OUTPUTS


def tier_options(k, max_corrections=12):
    name, tier, rate = TIERS[k]
    if tier is None:
        return MS.fallback_options()
    return MS.tier_options(tier, rate=rate, max_corrections=max_corrections)


def work_units(orders):
    """(n / 256)^3 per matrix: the batched solve's time scales with n^3."""
    return (np.asarray(orders, dtype=np.float64) / 256.0) ** 3


def unit_costs(tier_ms, orders):
    """Milliseconds per unit of work of each tier on one labelled batch."""
    return np.asarray(tier_ms, dtype=np.float64) / float(np.sum(work_units(orders)))


def policy_cost(choice, success, unit_cost, orders):
    """Modelled time (ms) of a per-matrix tier choice on a labelled batch: a
    matrix costs its tier's unit cost times (n / 256)^3, and every choice that
    failed pays the plain path again; returns (cost, failures)."""
    rows = np.arange(len(choice))
    w = work_units(orders)
    c = np.asarray(unit_cost, dtype=np.float64)
    ok = success[rows, choice] > 0
    return float(np.sum(w * (c[choice] + np.where(ok, 0.0, c[0])))), int(np.sum(~ok))


# -------------------------------------------------------------- labelling
def true_errors(roots, mats, offs, lds):
    """Relative Frobenius error of each solved root against the known root."""
    Xh = roots.cpu().numpy()
    out = np.empty(len(mats))
    for i, (m, off, ld) in enumerate(zip(mats, offs, lds)):
        n = m["order"]
        Xi = Xh[off:off + ld * ld].reshape(ld, ld, order="F")[:n, :n]
        out[i] = np.linalg.norm(Xi - m["X"]) / np.linalg.norm(m["X"])
    return out


def label_batch(rt, mats, target=MS.TARGET, log=None):
    """Runs every tier on the batch.  Returns (features, success, sweeps,
    corrections, orders, tier_ms): arrays of shape (count, 16), (count, 8) x 3,
    (count,), (8,); tier_ms is the wall time of each tier's batched solve (the
    median of three runs after a first, so graph capture and workspace growth
    are excluded)."""
    import time

    import torch

    orders = [m["order"] for m in mats]
    _, aflat = capi.pack_batch([m["A"] for m in mats])
    lds, offs, _, total, _ = capi.batch_layout(orders)
    A = torch.from_numpy(aflat).cuda()
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    feats = torch.zeros(len(mats) * capi.MLP_INPUTS, dtype=torch.float32, device="cuda")
    rt.matrix_features(orders, A, feats)
    count, tiers = len(mats), len(TIERS)
    success = np.zeros((count, tiers), dtype=np.float32)
    sweeps = np.zeros((count, tiers))
    corrections = np.zeros((count, tiers))
    tier_ms = np.zeros(tiers)
    for k in range(tiers):
        opts = tier_options(k)
        rt.solve_batched(orders, A, X, opts)
        runs = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = rt.solve_batched(orders, A, X, opts)
            torch.cuda.synchronize()
            runs.append((time.perf_counter() - t0) * 1e3)
        tier_ms[k] = float(np.median(runs))
        err = true_errors(X, mats, offs, lds)
        ok = reps.converged & (err <= target)
        success[:, k] = ok
        sweeps[:, k] = reps.sweeps
        corrections[:, k] = reps.corrections
        if log:
            log(f"  tier {TIERS[k][0]}: success {ok.mean():.3f}, mean sweeps "
                f"{sweeps[:, k].mean():.2f}, mean corrections {corrections[:, k].mean():.2f}, "
                f"{tier_ms[k]:.1f} ms")
    return (feats.cpu().numpy().reshape(count, capi.MLP_INPUTS), success, sweeps, corrections,
            np.array(orders), tier_ms)


# ------------------------------------------------------------------ network
def init_params(seed=0):
    """He-initialised weights, zero biases (float32, the library's layout)."""
    rng = np.random.default_rng(seed)
    h, i, o = capi.MLP_HIDDEN, capi.MLP_INPUTS, capi.MLP_OUTPUTS
    parts = [rng.standard_normal((h, i)) * np.sqrt(2.0 / i), np.zeros(h),
             rng.standard_normal((h, h)) * np.sqrt(2.0 / h), np.zeros(h),
             rng.standard_normal((o, h)) * np.sqrt(1.0 / h), np.zeros(o)]
    return np.concatenate([p.reshape(-1) for p in parts]).astype(np.float32)


def fold_standardisation(params, mean, std):
    """Weights that take raw features: W1' = W1 / std, b1' = b1 - W1' mean."""
    p = np.array(params, dtype=np.float64)
    h, i = capi.MLP_HIDDEN, capi.MLP_INPUTS
    W1 = p[:h * i].reshape(h, i) / std[None, :]
    b1 = p[h * i:h * i + h] - W1 @ mean
    p[:h * i] = W1.reshape(-1)
    p[h * i:h * i + h] = b1
    return p.astype(np.float32)


def train(rt, x, y, mask=None, epochs=200, batch=512, lr=3e-3, decay=1e-4, seed=0,
          params=None, log=None):
    """Trains on standardised inputs x (N x 16) and targets y (N x 8) with the
    library's kernels; returns (params, per-epoch mean loss)."""
    import torch

    n = x.shape[0]
    mask = np.ones_like(y, dtype=np.float32) if mask is None else mask.astype(np.float32)
    dev = torch.device("cuda")
    X = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(dev)
    Y = torch.from_numpy(np.ascontiguousarray(y, dtype=np.float32)).to(dev)
    M = torch.from_numpy(np.ascontiguousarray(mask)).to(dev)
    P = torch.from_numpy(init_params(seed) if params is None else np.array(params, np.float32)).to(dev)
    G = torch.zeros_like(P)
    m1 = torch.zeros_like(P)
    m2 = torch.zeros_like(P)
    bmax = min(batch, n)
    logits = torch.zeros(bmax * capi.MLP_OUTPUTS, device=dev)
    h1 = torch.zeros(bmax * capi.MLP_HIDDEN, device=dev)
    h2 = torch.zeros(bmax * capi.MLP_HIDDEN, device=dev)
    losses = torch.zeros(bmax, device=dev)
    rng = np.random.default_rng(seed + 1)
    history, step = [], 0
    for ep in range(epochs):
        order = torch.from_numpy(rng.permutation(n)).to(dev)
        total, weight = 0.0, 0.0
        for s in range(0, n, bmax):
            idx = order[s:s + bmax]
            b = idx.numel()
            xb = X[idx].contiguous()
            yb = Y[idx].contiguous()
            mb = M[idx].contiguous()
            mt = float(mb.sum())
            if mt == 0:
                continue
            rt.mlp_forward(P, xb, logits[:b * capi.MLP_OUTPUTS], h1[:b * capi.MLP_HIDDEN],
                           h2[:b * capi.MLP_HIDDEN])
            rt.mlp_backward(P, xb, h1[:b * capi.MLP_HIDDEN], h2[:b * capi.MLP_HIDDEN],
                            logits[:b * capi.MLP_OUTPUTS], yb, mb, mt, G, losses[:b])
            step += 1
            rt.mlp_adam(P, G, m1, m2, lr, step, decay=decay)
            total += float(losses[:b].sum())
            weight += mt
        history.append(total / max(weight, 1.0))
        if log and (ep % 25 == 0 or ep == epochs - 1):
            log(f"  epoch {ep}: loss {history[-1]:.4f}")
    return P.cpu().numpy(), history


# ------------------------------------------------------------------ selector
class Selector:
    """Trained network (raw-feature weights) and per-tier unit costs (ms per
    unit of work, see unit_costs)."""

    def __init__(self, params, tier_cost):
        self.params = np.asarray(params, dtype=np.float32)
        self.tier_cost = np.asarray(tier_cost, dtype=np.float64)

    @classmethod
    def load(cls, path=WEIGHTS):
        with np.load(path, allow_pickle=False) as z:
            return cls(z["params"], z["tier_cost"])

    def save(self, path=WEIGHTS, **extra):
        path = pathlib.Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, params=self.params, tier_cost=self.tier_cost, **extra)

    def probabilities(self, rt, orders, matrices):
        """P(tier k succeeds) per matrix, (count, 8), from device matrices."""
        import torch

        count = len(orders)
        dev = matrices.device
        feats = torch.zeros(count * capi.MLP_INPUTS, dtype=torch.float32, device=dev)
        rt.matrix_features(orders, matrices, feats)
        P = torch.from_numpy(self.params).to(dev)
        logits = torch.zeros(count * capi.MLP_OUTPUTS, dtype=torch.float32, device=dev)
        rt.mlp_forward(P, feats, logits)
        z = logits.cpu().numpy().reshape(count, capi.MLP_OUTPUTS).astype(np.float64)
        return 1.0 / (1.0 + np.exp(-z))

    def choose(self, prob, orders=None, fallback_weight=FALLBACK_WEIGHT, candidates=None):
        """Tier per matrix minimising c_t + (1 - p_t) F c_plain (tier 0: p = 1),
        over `candidates` (tier indices; all tiers when None; the plain tier
        is always a candidate)."""
        p = np.array(prob, dtype=np.float64)
        p[:, 0] = 1.0
        expected = self.tier_cost[None, :] + (1.0 - p) * fallback_weight * self.tier_cost[0]
        if candidates is not None:
            allowed = np.zeros(len(self.tier_cost), dtype=bool)
            allowed[list(candidates)] = True
            allowed[0] = True
            expected = np.where(allowed[None, :], expected, np.inf)
        return np.argmin(expected, axis=1)

    def options(self, rt, orders, matrices, candidates=None):
        tiers = self.choose(self.probabilities(rt, orders, matrices), orders,
                            candidates=candidates)
        opts = [tier_options(int(k)) for k in range(len(TIERS))]
        return [opts[int(k)] for k in tiers], tiers


# Tiers the batched solve runs as one precision group: BF16 basis and
# corrections at stop rates 0.5 to 4.  A second low-bit precision group (FP8,
# MXFP4) costs more in one batched solve than the per-matrix cost model
# predicts, because each group adds its own launches to every correction step.
# Measured on MI355X, in the same process on the same batches:
#   * 1000-matrix mixture, orders 32..512: 364.7 ms against 378.4 ms for the
#     fixed BF16 tier, with 17 fallbacks against 45; all eight tiers give
#     376.0 ms;
#   * 4096-matrix uneven batch: 89.1 ms against 75.8 ms.  The fixed tier
#     never fails there, so the features and the matrices sent to the plain
#     path are pure cost.
BATCH_CANDIDATES = (1, 2, 3, 4)


# Batch gate: when the network expects less than this fraction of the batch
# to fail on the gentlest BF16 tier (mean of 1 - p), the whole batch runs
# that tier.  The few fallback re-solves it may save cost less than routing
# matrices to other tiers adds to the batched solve.  Measured: the uneven
# 4096-matrix batch expects 0.3% (12.5 matrices; none fail) and gates to the
# fixed tier; the 32..512 mixture expects 4.6% (45.9; 45 fail) and is routed
# per matrix.
GATE_FAILURE_FRACTION = 0.005


def solve_selected(rt, selector, orders, matrices, roots, candidates=BATCH_CANDIDATES,
                   gate=GATE_FAILURE_FRACTION):
    """Selector-chosen tiers with the plain-path fallback.  Returns
    (reports, fallback_indices, tiers)."""
    prob = selector.probabilities(rt, orders, matrices)
    if gate is not None and float(np.mean(1.0 - prob[:, 1])) < gate:
        tiers = np.ones(len(orders), dtype=np.int64)
    else:
        tiers = selector.choose(prob, orders, candidates=candidates)
    opts_all = [tier_options(int(k)) for k in range(len(TIERS))]
    if np.all(tiers == tiers[0]):  # one record for the whole batch
        opts = opts_all[int(tiers[0])]
    else:
        opts = [opts_all[int(k)] for k in tiers]
    reports, fallback = MS.solve_with_fallback(rt, orders, matrices, roots, opts)
    return reports, fallback, tiers
