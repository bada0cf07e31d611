"""Learned tier selector of the batched SPD square-root solver.

No counterpart in the reference.  For each matrix of a batch the selector
chooses one of eight solver configurations ("tiers"): the product format of
the reconstruction and corrections -- MXFP4, scaled FP8 or BF16 (both with
BF16 corrections), or FP32 -- crossed with two Jacobi stop rules, stopping
once the frozen-factor corrections are predicted to contract by 0.5 per step
(fewer corrections) or by 2 (the Jacobi stops earlier, more corrections).
The plain FP32 path is a ninth choice the network does not score: it always
meets the target (with its second Jacobi pass for unresolved spectra) and is
the fallback of any tier that misses the binary64 residual check, so a wrong
choice costs time, never accuracy.

The tier space follows the format list of the learned-tier workload
({MXFP4, FP8, BF16, FP32}); its second axis is the Jacobi stop rule instead
of a budget of one or two corrections.  Measured on MI355X (4096-matrix
uneven batch and a 1000-matrix 32..512 mixture): capping the corrections at
two makes the Jacobi run 2-4 more sweeps (each costing about ten correction
steps) and sends 78% of the uneven batch (BF16 products) or 70% of the
mixture to the fallback; the two stop rules cover the same trade between
Jacobi sweeps and corrections without that cliff.

* features: ``Runtime.matrix_features`` (HIP kernel; 16 per matrix);
* network: 16 -> 64 -> 64 -> 8 ReLU MLP, output k = logit of "tier k meets
  the target within its correction budget"; trained in float32 by the
  library's forward, backward and Adam kernels; inference in BF16
  (``Runtime.mlp_forward_bf16``: BF16 weights and activations, float
  accumulation);
* labels: the solver's own runs of every tier and of the plain path on
  seeded synthetic batches (``label_batch``), success = converged and
  relative Frobenius error against the construction-known root <= 1e-6;
* decision: per matrix of order n, the choice of least expected time
  w c_t + (1 - p_t) (L(n) + w c_plain), w = (n / 256)^3, with p_t the
  predicted success, c_t the measured cost of choice t (the wall time of its
  batched solve of each labelled batch divided by the batch's work sum,
  pooled over the batches) and L(n) the measured latency of the plain-path
  solve of one matrix of order n (fallback_latency.py): a failed choice is
  solved again in a second batched call, whose time is that of its slowest
  matrix rather than the matrix's share of a batch.  The plain path costs
  w c_plain.  Each product format used in a batch adds its own launches to
  every correction step of the batched solve, which the per-matrix costs do
  not model: formats chosen for less than GROUP_SHARE_MIN of the batch's work
  are dropped (their matrices take their best remaining choice), and a batch
  whose best single choice is expected within SINGLE_MARGIN of the
  per-matrix choices runs that single choice.  A batch expected to pass on
  the fixed BF16 tier (GATE_FAILURE_FRACTION) runs it, and a batch that runs
  the fixed tier as a whole stops its orders above FIXED_BATCH_LARGE_ORDER at
  stop rate FIXED_BATCH_LARGE_RATE (``batch_options``);
* batch gate from a sample: batches of GATE_SAMPLE_MIN matrices or more
  first score one matrix in GATE_SAMPLE_STRIDE (stratified by order,
  ``gate_sample``); when the sample's mean predicted failure of the fixed tier
  is below the gate (times GATE_SAMPLE_MARGIN), the batch runs the fixed tier
  without the other matrices' features (the features are most of the
  selector's cost: 1.5 ms for the whole 4096-matrix uneven batch, 0.3 ms for
  its sample); otherwise the rest of the batch is scored and the full
  decision above runs.

Feature standardisation is folded into the first layer when the weights are
saved, so inference runs the raw features through the network unchanged.
"""
from __future__ import annotations

import pathlib

import numpy as np

from . import capi
from . import solver as MS

WEIGHTS = pathlib.Path(__file__).resolve().parent / "data" / "selector.npz"
# Weight of the expected fallback latency (1: as measured).
FALLBACK_WEIGHT = 1.0

# (name, product format, correction_rate); the network's output k scores
# TIERS[k].  PLAIN (= 8) is the plain FP32 path.
TIERS = [
    ("mxfp4_r0.5", "mxfp4", 0.5),
    ("mxfp4_r2", "mxfp4", 2.0),
    ("fp8_r0.5", "fp8", 0.5),
    ("fp8_r2", "fp8", 2.0),
    ("bf16_r0.5", "bf16", 0.5),
    ("bf16_r2", "bf16", 2.0),
    ("fp32_r0.5", "fp32", 0.5),
    ("fp32_r2", "fp32", 2.0),
]
assert len(TIERS) == capi.MLP_OUTPUTS
PLAIN = len(TIERS)
CHOICES = [t[0] for t in TIERS] + ["plain_fp32"]
FIXED = 4  # bf16_r0.5: the fixed tier the selector is compared with
GROUP_SHARE_MIN = 0.1
# Measured on MI355X (held-out batches, same process): choices that mix two
# product formats ran 4.9 ms (8%) slower on the 4096-matrix uneven batch than
# their best single choice, so a single choice within 10% is preferred.
SINGLE_MARGIN = 0.1
# A batch the network expects to pass on the fixed tier (mean predicted
# failure below 0.5%) runs the fixed tier: on the uneven batch the tiers'
# costs lie within a few percent of each other, below what the pooled cost
# model resolves (measured: the uneven batch expects 0.3% and none fail; the
# 32..512 mixture expects about 5% and 5.6% fail).
GATE_FAILURE_FRACTION = 0.005
# The gate estimated from a stratified sample of one matrix in 16 (batches of
# 2048 or more, so at least 128 matrices), held to the same threshold: near
# the gate the fixed tier and the per-matrix choices cost about the same, so
# an estimate that lands on the wrong side of it costs little; far from it a
# stratified sample does not mistake the batch.  Measured on MI355X from the
# network's per-matrix predictions over 4096-matrix batches, resampled 4000
# times: the uneven batch's sample mean at one in 16 is 0.0025 +- 0.00044
# (5.7 standard deviations below the gate; two seeds), the 32..512
# mixture's 0.065 +- 0.013 (4.5 above); neither was misjudged.  One in 16
# instead of one in 8 takes the sample's features from two waves of
# workgroups to one: the gate's features and inference 0.41 -> 0.30 ms.  A
# misjudged batch costs time (fallback solves), never accuracy.
GATE_SAMPLE_STRIDE = 16
GATE_SAMPLE_MIN = 2048
GATE_SAMPLE_MARGIN = 1.0
# A batch that runs the fixed tier as a whole stops the Jacobi of its orders
# above FIXED_BATCH_LARGE_ORDER at stop rate FIXED_BATCH_LARGE_RATE instead of
# the tier's 0.5: those orders carry most of the Jacobi's work, and the extra
# correction they need on average runs inside the correction steps that the
# batch's smaller matrices take anyway.  Measured on MI355X (blocked
# eigensolver, graphs), two independently seeded 4096-matrix uneven batches:
# 36.45 -> 35.43 ms and 36.26 -> 35.31 ms, no fallbacks, the largest
# correction count unchanged (6 and 7); thresholds 96..160 came within 0.1 ms
# of each other, rate 1.5 raised the largest count to 8, and rate 1 on the
# orders up to 128 instead lengthened the correction tail (38.4 ms).
FIXED_BATCH_LARGE_ORDER = 128
FIXED_BATCH_LARGE_RATE = 1.0


def gate_sample(orders):
    """Batch positions of the gate's sample: positions sorted by order (stable),
    every GATE_SAMPLE_STRIDE-th from the middle of the first stride, in
    position order; None for batches below GATE_SAMPLE_MIN."""
    count = len(orders)
    if count < GATE_SAMPLE_MIN:
        return None
    # Orders are at most 512: as uint16 the stable sort is numpy's radix sort
    # (4x faster than on 32-bit keys; the gate's host time precedes its
    # kernels).
    by_order = np.argsort(np.asarray(orders).astype(np.uint16), kind="stable")
    return np.sort(by_order[GATE_SAMPLE_STRIDE // 2::GATE_SAMPLE_STRIDE])


def choice_format(k):
    """Product format of choice k ("fp32" for the plain path)."""
    return "fp32" if k == PLAIN else TIERS[k][1]


def tier_options(k, max_corrections=12):
    if k == PLAIN:
        return MS.fallback_options()
    name, tier, rate = TIERS[k]
    return MS.tier_options(tier, rate=rate, max_corrections=max_corrections)


def work_units(orders):
    """(n / 256)^3 per matrix: the batched solve's time scales with n^3."""
    return (np.asarray(orders, dtype=np.float64) / 256.0) ** 3


def unit_costs(tier_ms, orders):
    """Milliseconds per unit of work of each tier on one labelled batch."""
    return np.asarray(tier_ms, dtype=np.float64) / float(np.sum(work_units(orders)))


def policy_cost(choice, success, unit_cost, orders, latency=None):
    """Modelled time (ms) of a per-matrix choice on a labelled batch: a
    matrix costs its choice's unit cost times (n / 256)^3; the choices that
    failed pay the plain path again (unit_cost[PLAIN] times their work) in
    one second call, which takes at least `latency(n)` of its largest failed
    matrix when `latency` is given.  success and unit_cost hold the eight
    tiers, then the plain path.  Returns (cost, failures)."""
    rows = np.arange(len(choice))
    w = work_units(orders)
    c = np.asarray(unit_cost, dtype=np.float64)
    ok = success[rows, choice] > 0
    total = float(np.sum(w * (c[choice] + np.where(ok, 0.0, c[PLAIN]))))
    if latency is not None and not ok.all():
        total += float(np.max(latency(np.asarray(orders)[~ok])))
    return total, int(np.sum(~ok))


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


def label_batch(rt, mats, target=MS.TARGET, log=None, runs=3):
    """Runs every tier and the plain path on the batch.  Returns (features,
    success, sweeps, corrections, orders, tier_ms): arrays of shape (count,
    16), (count, 9) x 3, (count,), (9,); tier_ms is the wall time of each
    choice's batched solve (the median of `runs` runs after a first, so graph
    capture and workspace growth are excluded)."""
    import time

    import torch

    orders = [m["order"] for m in mats]
    _, aflat = capi.pack_batch([m["A"] for m in mats])
    lds, offs, _, total, _ = capi.batch_layout(orders)
    A = torch.from_numpy(aflat).cuda()
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    feats = torch.zeros(len(mats) * capi.MLP_INPUTS, dtype=torch.float32, device="cuda")
    rt.matrix_features(orders, A, feats)
    count, tiers = len(mats), len(CHOICES)
    success = np.zeros((count, tiers), dtype=np.float32)
    sweeps = np.zeros((count, tiers))
    corrections = np.zeros((count, tiers))
    tier_ms = np.zeros(tiers)
    for k in range(tiers):
        opts = tier_options(k)
        rt.solve_batched(orders, A, X, opts)
        times = []
        for _ in range(runs):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            reps = rt.solve_batched(orders, A, X, opts)
            torch.cuda.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        tier_ms[k] = float(np.median(times))
        err = true_errors(X, mats, offs, lds)
        ok = reps.converged & (err <= target)
        success[:, k] = ok
        sweeps[:, k] = reps.sweeps
        corrections[:, k] = reps.corrections
        if log:
            log(f"  {CHOICES[k]}: success {ok.mean():.3f}, mean sweeps "
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
    """Trained network (raw-feature weights), per-choice unit costs (ms per
    unit of work, see unit_costs: the eight tiers, then the plain path) and
    the measured latency of a plain-path fallback solve per order."""

    def __init__(self, params, choice_cost, fallback_orders=None, fallback_ms=None):
        self.params = np.asarray(params, dtype=np.float32)
        self.choice_cost = np.asarray(choice_cost, dtype=np.float64)
        assert self.choice_cost.shape == (len(CHOICES),)
        self.fallback_orders = (np.array([1.0]) if fallback_orders is None
                                else np.asarray(fallback_orders, dtype=np.float64))
        self.fallback_ms = (np.array([0.0]) if fallback_ms is None
                            else np.asarray(fallback_ms, dtype=np.float64))
        self._p16 = None

    @classmethod
    def load(cls, path=WEIGHTS):
        with np.load(path, allow_pickle=False) as z:
            lat = (z["fallback_orders"], z["fallback_ms"]) if "fallback_ms" in z.files else (None, None)
            return cls(z["params"], z["choice_cost"], *lat)

    def save(self, path=WEIGHTS, **extra):
        path = pathlib.Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez(path, params=self.params, choice_cost=self.choice_cost,
                 fallback_orders=self.fallback_orders, fallback_ms=self.fallback_ms, **extra)

    def latency(self, orders):
        """Measured latency (ms) of the plain-path solve of one matrix, per
        order (linear interpolation over the measured grid)."""
        return np.interp(np.asarray(orders, dtype=np.float64), self.fallback_orders,
                         self.fallback_ms)

    def bf16_params(self, rt, dev):
        """The weights as BF16 codes on the device (encoded once, by the
        library's BF16 codec)."""
        import torch

        if self._p16 is None or self._p16.device != dev:
            P = torch.from_numpy(self.params).to(dev)
            p16 = torch.zeros(capi.MLP_PARAMS, dtype=torch.int16, device=dev)
            rt.quantize(capi.BF16, P, p16, None)
            self._p16 = p16
        return self._p16

    def infer(self, rt, feats, rows=None):
        """P(tier k succeeds), (count, 8), for device features (count x 16):
        BF16 inference; with `rows` (host indices), the probabilities of those
        rows only (the network runs on every row: cheaper than a device
        gather and its index upload)."""
        import torch

        count = feats.numel() // capi.MLP_INPUTS
        logits = torch.empty(count * capi.MLP_OUTPUTS, dtype=torch.float32, device=feats.device)
        rt.mlp_forward_bf16(self.bf16_params(rt, feats.device), feats, logits)
        z = logits.cpu().numpy().reshape(count, capi.MLP_OUTPUTS)
        if rows is not None:
            z = z[rows]
        return 1.0 / (1.0 + np.exp(-z.astype(np.float64)))

    def probabilities(self, rt, orders, matrices):
        """P(tier k succeeds) per matrix, (count, 8), from device matrices:
        features, then BF16 inference."""
        import torch

        count = len(orders)
        feats = torch.zeros(count * capi.MLP_INPUTS, dtype=torch.float32, device=matrices.device)
        rt.matrix_features(orders, matrices, feats)
        return self.infer(rt, feats)

    def choose_for_batch(self, rt, orders, matrices, fallback_weight=FALLBACK_WEIGHT):
        """Choices for one batched solve of device matrices: the gate from
        the sample first (large batches), else every matrix scored and
        ``choose_batch``.  Returns (choices, number of matrices scored)."""
        import torch

        count = len(orders)
        dev = matrices.device
        feats = torch.zeros(count * capi.MLP_INPUTS, dtype=torch.float32, device=dev)
        sample = gate_sample(orders)
        if sample is not None:
            rt.matrix_features(orders, matrices, feats, members=sample)
            ps = self.infer(rt, feats, rows=sample)
            if float(np.mean(1.0 - ps[:, FIXED])) < GATE_SAMPLE_MARGIN * GATE_FAILURE_FRACTION:
                return np.full(count, FIXED), len(sample)
            rest = np.setdiff1d(np.arange(count), sample)
            rt.matrix_features(orders, matrices, feats, members=rest)
        else:
            rt.matrix_features(orders, matrices, feats)
        prob = self.infer(rt, feats)
        return self.choose_batch(prob, orders, fallback_weight), count

    def expected(self, prob, orders, fallback_weight=FALLBACK_WEIGHT):
        """Expected time (ms) of every choice for every matrix, (count, 9):
        w c_t + (1 - p_t) (F L(n) + w c_plain) for the tiers, w c_plain for
        the plain path."""
        w = work_units(orders)[:, None]
        p = np.ones((prob.shape[0], len(CHOICES)))
        p[:, :PLAIN] = prob
        c = self.choice_cost[None, :]
        lat = fallback_weight * self.latency(orders)[:, None]
        return w * c + (1.0 - p) * (lat + w * c[:, PLAIN:PLAIN + 1])

    def choose(self, prob, orders, fallback_weight=FALLBACK_WEIGHT, candidates=None):
        """Choice per matrix (0..7 a tier, PLAIN the plain path) minimising
        the expected time, over `candidates` (choice indices; all when None;
        the plain path is always a candidate)."""
        expected = self.expected(prob, orders, fallback_weight)
        if candidates is not None:
            allowed = np.zeros(len(CHOICES), dtype=bool)
            allowed[list(candidates)] = True
            allowed[PLAIN] = True
            expected = np.where(allowed[None, :], expected, np.inf)
        return np.argmin(expected, axis=1)

    def choose_batch(self, prob, orders, fallback_weight=FALLBACK_WEIGHT):
        """Per-matrix choices for one batched solve: the fixed tier for a
        batch expected to pass on it; otherwise formats chosen for less than
        GROUP_SHARE_MIN of the work are dropped, and a single choice is taken
        when it is expected within SINGLE_MARGIN of the per-matrix ones."""
        if float(np.mean(1.0 - prob[:, FIXED])) < GATE_FAILURE_FRACTION:
            return np.full(prob.shape[0], FIXED)
        w = work_units(orders)
        expected = self.expected(prob, orders, fallback_weight)
        allowed = list(range(len(CHOICES)))
        while True:
            mask = np.isin(np.arange(len(CHOICES)), allowed)[None, :]
            choice = np.argmin(np.where(mask, expected, np.inf), axis=1)
            share = {}
            for k, wk in zip(choice, w):
                f = choice_format(int(k))
                share[f] = share.get(f, 0.0) + wk
            total = float(np.sum(w))
            small = [f for f, v in share.items()
                     if v < GROUP_SHARE_MIN * total and len(share) > 1]
            if not small:
                break
            drop = min(small, key=lambda f: share[f])
            allowed = [k for k in allowed if choice_format(k) != drop or k == PLAIN]
        per_matrix = float(np.sum(expected[np.arange(len(choice)), choice]))
        single = np.sum(expected, axis=0)
        best = int(np.argmin(single))
        if single[best] <= (1.0 + SINGLE_MARGIN) * per_matrix:
            return np.full(len(choice), best)
        return choice


_CHOICE_OPTIONS = []


def choice_options():
    """Solve options of every choice, then the fixed tier's format at
    FIXED_BATCH_LARGE_RATE (built once: ctypes records)."""
    if not _CHOICE_OPTIONS:
        _CHOICE_OPTIONS.extend(tier_options(k) for k in range(len(CHOICES)))
        _CHOICE_OPTIONS.append(MS.tier_options(TIERS[FIXED][1], rate=FIXED_BATCH_LARGE_RATE,
                                               max_corrections=12))
    return _CHOICE_OPTIONS


def batch_options(choice, orders):
    """Options for one batched solve: one record when the batch makes one
    choice (the fixed tier's with its large orders at FIXED_BATCH_LARGE_RATE),
    else per-matrix records."""
    opts = choice_options()
    if np.all(choice == choice[0]):
        if int(choice[0]) != FIXED:
            return opts[int(choice[0])]
        large = np.asarray(orders) > FIXED_BATCH_LARGE_ORDER
        if not large.any():
            return opts[FIXED]
        return capi.options_array([opts[FIXED], opts[-1]], large)
    return capi.options_array(opts, choice)


def solve_selected(rt, selector, orders, matrices, roots):
    """Selector-chosen tiers with the plain-path fallback.  Returns
    (reports, fallback_indices, choices)."""
    # One array of orders for every call below (a list would be converted
    # by each).
    orders = np.ascontiguousarray(orders, dtype=np.uint32)
    choice, _ = selector.choose_for_batch(rt, orders, matrices)
    reports, fallback = MS.solve_with_fallback(rt, orders, matrices, roots,
                                               batch_options(choice, orders))
    return reports, fallback, choice
