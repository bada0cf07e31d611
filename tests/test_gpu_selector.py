"""GPU checks of the learned tier selector's kernels against the CPU oracle
(oracle/mlp.py), its training, and what the trained network gains."""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, say
from msq import capi
from msq import generators as G
from msq import selector as SEL
from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _spd(rng, n, cond):
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = cond ** (rng.random(n) - 1.0)
    A = (Q * lam) @ Q.T
    return 0.5 * (A + A.T)


def test_features_match_oracle(runtime):
    rng = np.random.default_rng(11)
    orders = [5, 16, 17, 64, 200, 300, 512]
    mats = [_spd(rng, n, 10.0 ** (1 + i % 4)) for i, n in enumerate(orders)]
    mats.append(np.diag(np.array([10.0] + [5.0] * 30 + [0.1])))
    orders.append(32)
    say("features", orders)
    o, aflat = capi.pack_batch(mats)
    F = torch.zeros(len(o) * capi.MLP_INPUTS, dtype=torch.float32, device="cuda")
    runtime.matrix_features(o, dev(aflat), F)
    got = host(F).reshape(len(o), capi.MLP_INPUTS)
    for A, g in zip(mats, got):
        want = M.features(A)
        # Same arithmetic (binary64; power steps on A rounded to float32),
        # different summation orders; float32 output.
        plain = [k for k in range(M.INPUTS) if k not in (13, 14)]
        np.testing.assert_allclose(g[plain], want[plain], rtol=2e-6, atol=2e-6)
        # 13, 14: relative change of the last power step, a difference of two
        # nearly equal Rayleigh quotients: compared on the linear scale, to
        # the summation-order rounding of a quotient (~ n eps).
        n = A.shape[0]
        np.testing.assert_allclose(10.0 ** g[[13, 14]].astype(np.float64), 10.0 ** want[[13, 14]],
                                   rtol=1e-5, atol=8 * n * 2.0**-52)


def test_feature_subset_rows_equal_the_full_call(runtime):
    """msq_matrix_features_subset: the listed rows bit for bit as the full
    call computes them, the other rows untouched."""
    rng = np.random.default_rng(12)
    orders = [int(n) for n in rng.integers(1, 513, 40)]
    mats = [_spd(rng, n, 10.0 ** rng.uniform(1, 5)) for n in orders]
    o, aflat = capi.pack_batch(mats)
    A = dev(aflat)
    full = torch.zeros(len(o) * capi.MLP_INPUTS, dtype=torch.float32, device="cuda")
    runtime.matrix_features(o, A, full)
    members = sorted(rng.choice(len(o), 17, replace=False).tolist())
    part = torch.full((len(o) * capi.MLP_INPUTS,), 7.0, dtype=torch.float32, device="cuda")
    runtime.matrix_features(o, A, part, members=members)
    f, p = host(full).reshape(len(o), -1), host(part).reshape(len(o), -1)
    say("feature subset", len(members), "of", len(o))
    assert np.array_equal(p[members].view(np.uint32), f[members].view(np.uint32))
    rest = np.setdiff1d(np.arange(len(o)), members)
    assert np.all(p[rest] == 7.0)
    with pytest.raises(capi.MsqError):
        runtime.matrix_features(o, A, part, members=[len(o)])


def test_sampled_gate_agrees_with_the_full_batch(runtime):
    """The batched uneven workload (first 2048 matrices): the gate's sample
    sends the batch to the fixed tier, as the full batch's mean predicted
    failure does; the sample scores one matrix in GATE_SAMPLE_STRIDE."""
    if not SEL.WEIGHTS.exists():
        pytest.skip("no trained selector weights")
    sel = SEL.Selector.load()
    mats = G.batched_uneven(count=2048)
    o, flat = capi.pack_batch([m["A"] for m in mats])
    A = dev(flat)
    choice, scored = sel.choose_for_batch(runtime, o, A)
    prob = sel.probabilities(runtime, o, A)
    full_fail = float(np.mean(1.0 - prob[:, SEL.FIXED]))
    sample = SEL.gate_sample(o)
    say("gate: full-batch predicted failure", full_fail, "sample",
        float(np.mean(1.0 - prob[sample, SEL.FIXED])), "scored", scored)
    assert full_fail < SEL.GATE_FAILURE_FRACTION
    assert scored == len(sample) == len(o) // SEL.GATE_SAMPLE_STRIDE
    assert (choice == SEL.FIXED).all()
    assert (sel.choose_batch(prob, o) == SEL.FIXED).all()


def test_features_reject_orders_above_512(runtime):
    A = np.eye(520)
    o, aflat = capi.pack_batch([A])
    F = torch.zeros(capi.MLP_INPUTS, dtype=torch.float32, device="cuda")
    with pytest.raises(capi.MsqError):
        runtime.matrix_features(o, dev(aflat), F)


def _batch(seed, b):
    rng = np.random.default_rng(seed)
    p = (rng.standard_normal(M.PARAMS) * 0.3).astype(np.float32)
    x = rng.standard_normal((b, M.INPUTS)).astype(np.float32)
    y = (rng.random((b, M.OUTPUTS)) < 0.4).astype(np.float32)
    m = (rng.random((b, M.OUTPUTS)) < 0.9).astype(np.float32)
    return p, x, y, m


@pytest.mark.parametrize("b", [1, 37, 700])
def test_forward_and_backward_match_oracle(runtime, b):
    p, x, y, m = _batch(20 + b, b)
    P, X, Y, Mk = dev(p), dev(x), dev(y), dev(m)
    z = torch.zeros(b * M.OUTPUTS, device="cuda")
    h1 = torch.zeros(b * M.HIDDEN, device="cuda")
    h2 = torch.zeros(b * M.HIDDEN, device="cuda")
    runtime.mlp_forward(P, X, z, h1, h2)
    zo, h1o, h2o = M.forward(p.astype(np.float64), x.astype(np.float64))
    np.testing.assert_allclose(host(z).reshape(b, -1), zo, rtol=1e-4, atol=1e-4)
    np.testing.assert_allclose(host(h2).reshape(b, -1), h2o, rtol=1e-4, atol=1e-4)
    g = torch.zeros(M.PARAMS, device="cuda")
    losses = torch.zeros(b, device="cuda")
    runtime.mlp_backward(P, X, h1, h2, z, Y, Mk, float(m.sum()), g, losses)
    go = M.backward(p.astype(np.float64), x.astype(np.float64), y, m)
    gg = host(g)
    scale = np.abs(go).max()
    say("backward", b, "max |g - oracle| / max|g|", np.abs(gg - go).max() / scale)
    assert np.abs(gg - go).max() <= 1e-4 * scale
    lo = M.loss(p.astype(np.float64), x.astype(np.float64), y, m)
    np.testing.assert_allclose(host(losses).sum() / m.sum(), lo, rtol=1e-4)


@pytest.mark.parametrize("b", [1, 300, 5000])
def test_forward_bf16_matches_oracle(runtime, b):
    """BF16 inference (weights, inputs and activations in BF16, float
    accumulation in index order) bit for bit against the oracle, and within
    BF16 rounding of the float32 network."""
    p, x, _, _ = _batch(50 + b, b)
    x[: min(b, 7)] *= 1e3  # large inputs: rounding ties and saturating sums
    p16 = torch.zeros(M.PARAMS, dtype=torch.int16, device="cuda")
    runtime.quantize(capi.BF16, dev(p), p16, None)
    z = torch.zeros(b * M.OUTPUTS, device="cuda")
    runtime.mlp_forward_bf16(p16, dev(x), z)
    got = host(z).reshape(b, M.OUTPUTS)
    want = M.forward_bf16(p, x)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # Unscaled rows: within BF16 rounding of the float32 network.
    k = min(b, 7)
    if b > k:
        z32, _, _ = M.forward(p.astype(np.float64), x[k:].astype(np.float64))
        diff = np.abs(got[k:] - z32).max() / np.abs(z32).max()
        say("bf16 vs float32 network: max difference / max |z|", float(diff))
        assert diff < 0.05


def test_backward_is_deterministic(runtime):
    p, x, y, m = _batch(31, 300)
    P, X, Y, Mk = dev(p), dev(x), dev(y), dev(m)
    z = torch.zeros(300 * M.OUTPUTS, device="cuda")
    h1 = torch.zeros(300 * M.HIDDEN, device="cuda")
    h2 = torch.zeros(300 * M.HIDDEN, device="cuda")
    runtime.mlp_forward(P, X, z, h1, h2)
    g1 = torch.zeros(M.PARAMS, device="cuda")
    g2 = torch.zeros(M.PARAMS, device="cuda")
    runtime.mlp_backward(P, X, h1, h2, z, Y, Mk, float(m.sum()), g1)
    runtime.mlp_backward(P, X, h1, h2, z, Y, Mk, float(m.sum()), g2)
    assert torch.equal(g1, g2)


def test_adam_matches_oracle(runtime):
    rng = np.random.default_rng(40)
    p = rng.standard_normal(M.PARAMS).astype(np.float32)
    P = dev(p.copy())
    m1 = torch.zeros(M.PARAMS, device="cuda")
    m2 = torch.zeros(M.PARAMS, device="cuda")
    po, a, v = p.astype(np.float64), np.zeros(M.PARAMS), np.zeros(M.PARAMS)
    for step in range(1, 4):
        g = rng.standard_normal(M.PARAMS).astype(np.float32)
        runtime.mlp_adam(P, dev(g), m1, m2, 1e-2, step, decay=1e-3)
        po, a, v = M.adam(po, g.astype(np.float64), a, v, 1e-2, step, decay=1e-3)
    np.testing.assert_allclose(host(P), po, rtol=1e-5, atol=1e-6)


def test_training_fits_a_learnable_rule(runtime):
    """Targets that a feature threshold decides: the GPU-trained network
    learns them (loss falls, accuracy > 95 %)."""
    rng = np.random.default_rng(50)
    x = rng.standard_normal((4000, M.INPUTS)).astype(np.float32)
    y = np.stack([(x[:, k % M.INPUTS] + 0.5 * x[:, (k + 3) % M.INPUTS] > 0.2 * k - 0.7)
                  for k in range(M.OUTPUTS)], axis=1).astype(np.float32)
    params, hist = SEL.train(runtime, x, y, epochs=40, batch=256, lr=3e-3)
    z, _, _ = M.forward(params.astype(np.float64), x.astype(np.float64))
    acc = np.mean((z > 0) == (y > 0.5))
    say("training", hist[0], hist[-1], acc)
    assert hist[-1] < 0.5 * hist[0] and acc > 0.95


def test_trained_selector_beats_the_fixed_tier(runtime):
    """Held-out batch (mixture seed 13, never used in training): the committed
    network's batch choices cost no more than the fixed BF16 tier or the
    plain path, each with the plain-path fallback, in the modelled time of the
    batch's own measured choice costs (every choice is run on the batch and
    timed; a failure costs the measured latency of a fallback solve), and the
    BF16 inference agrees with the float32 network on the choices."""
    if not SEL.WEIGHTS.exists():
        pytest.skip("no trained selector weights")
    sel = SEL.Selector.load()
    mats = list(G.mixture(count=768, config=13, order_range=(32, 256)))
    say("selector held-out batch", len(mats))
    x, success, sweeps, corr, orders, choice_ms = SEL.label_batch(runtime, mats, log=say)
    o, flat = capi.pack_batch([m["A"] for m in mats])
    prob = sel.probabilities(runtime, o, dev(flat))
    logits = torch.zeros(len(mats) * M.OUTPUTS, device="cuda")
    runtime.mlp_forward(dev(sel.params), dev(x), logits)
    p32 = 1.0 / (1.0 + np.exp(-host(logits).reshape(len(mats), -1).astype(np.float64)))
    agree = np.mean(sel.choose(prob, o) == sel.choose(p32, o))
    choice = sel.choose_batch(prob, o)
    uc = SEL.unit_costs(choice_ms, orders)
    c_sel, f_sel = SEL.policy_cost(choice, success, uc, orders, sel.latency)
    c_fix, f_fix = SEL.policy_cost(np.full(len(mats), SEL.FIXED), success, uc, orders, sel.latency)
    c_plain, _ = SEL.policy_cost(np.full(len(mats), SEL.PLAIN), success, uc, orders, sel.latency)
    say("modelled ms: selector", c_sel, f_sel, "fixed", c_fix, f_fix, "plain", c_plain,
        "choices", np.bincount(choice, minlength=len(SEL.CHOICES)).tolist(),
        "bf16/fp32 inference agreement", agree)
    assert success[:, SEL.PLAIN].mean() >= 0.99  # the fallback meets the target
    assert agree >= 0.95
    assert c_sel <= min(c_fix, c_plain)
