"""CPU checks of the selector's oracle and host logic (no GPU).

The oracle's backward is pinned by central finite differences of its loss;
its features by matrices whose features are known in closed form and by
eigenvalue bounds that hold for any SPD matrix.
"""
import numpy as np
import pytest

from oracle import mlp as M


def _params(seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(M.PARAMS) * 0.3


def test_parameter_count_matches_library_layout():
    assert M.PARAMS == 5768
    from msq import capi  # binding constants (no library call)
    assert capi.MLP_PARAMS == M.PARAMS


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(2)
    x = rng.standard_normal((7, M.INPUTS))
    y = (rng.random((7, M.OUTPUTS)) < 0.5).astype(float)
    m = (rng.random((7, M.OUTPUTS)) < 0.8).astype(float)
    g = M.backward(p, x, y, m)
    # Random coordinates of every layer; ReLU kinks are measure-zero.
    idx = np.concatenate([rng.choice(M.PARAMS, 60, replace=False),
                          np.arange(M.PARAMS - 8, M.PARAMS)])
    h = 1e-6
    for q in idx:
        e = np.zeros(M.PARAMS)
        e[q] = h
        fd = (M.loss(p + e, x, y, m) - M.loss(p - e, x, y, m)) / (2 * h)
        assert abs(fd - g[q]) <= 1e-6 + 1e-5 * abs(fd), (q, fd, g[q])


def test_backward_plants_a_sign_error_and_fails():
    """A gradient with one flipped layer sign is caught by the same check."""
    rng = np.random.default_rng(3)
    p = _params(4)
    x = rng.standard_normal((5, M.INPUTS))
    y = (rng.random((5, M.OUTPUTS)) < 0.5).astype(float)
    m = np.ones_like(y)
    g = M.backward(p, x, y, m)
    bad = g.copy()
    bad[-8:] *= -1.0  # b3 flipped
    q = M.PARAMS - 3
    e = np.zeros(M.PARAMS)
    e[q] = 1e-6
    fd = (M.loss(p + e, x, y, m) - M.loss(p - e, x, y, m)) / 2e-6
    assert abs(fd - g[q]) < 1e-6 and abs(fd - bad[q]) > 1e-4


def test_adam_first_step_is_signed_learning_rate():
    g = np.array([3.0, -0.5, 2.0, 1e-3])
    p, m1, m2 = M.adam(np.zeros(4), g, np.zeros(4), np.zeros(4), lr=0.01, step=1, eps=0.0)
    want = -0.01 * np.sign(g)
    np.testing.assert_allclose(p, want, rtol=1e-12)


def test_adam_decay_enters_the_gradient():
    """Weight decay d adds d p to the gradient before the moments: the first
    bias-corrected step is -lr sign(g + d p) (chosen so that every sign
    differs from sign(g)); the second is -lr m1_hat / sqrt(m2_hat) with the
    moments of the two decayed gradients, in exact arithmetic."""
    from fractions import Fraction as Fr
    p0 = np.array([1.0, -2.0, 0.5])
    g = np.array([-0.5, 0.5, -0.375])
    decay, lr = 1.0, 0.125
    p1, m1, m2 = M.adam(p0, g, np.zeros(3), np.zeros(3), lr=lr, step=1, eps=0.0, decay=decay)
    assert np.all(np.sign(g + decay * p0) != np.sign(g))
    np.testing.assert_allclose(p1, p0 - lr * np.sign(g + decay * p0), rtol=1e-15)
    g2 = np.array([0.25, -0.25, 0.5])
    p2, _, _ = M.adam(p1, g2, m1, m2, lr=lr, step=2, eps=0.0, decay=decay)
    b1, b2 = Fr(9, 10), Fr(999, 1000)
    for i in range(3):
        ga = Fr(g[i]) + Fr(decay) * Fr(p0[i])
        gb = Fr(g2[i]) + Fr(decay) * Fr(p1[i])
        a = b1 * (1 - b1) * ga + (1 - b1) * gb
        v = b2 * (1 - b2) * ga * ga + (1 - b2) * gb * gb
        want = float(Fr(p1[i])) - lr * float(a / (1 - b1 * b1)) / np.sqrt(float(v / (1 - b2 * b2)))
        assert p2[i] == pytest.approx(want, rel=1e-13)


def test_feature_8_is_the_spread_of_the_log_diagonal():
    """f8 = standard deviation of log10 diag(A): half the diagonal 1, half 100
    gives exactly 1; a constant diagonal gives 0."""
    A = np.diag([1.0, 100.0, 1.0, 100.0]) + 0.01 * (np.ones((4, 4)) - np.eye(4))
    assert M.features(A)[8] == pytest.approx(1.0, abs=1e-14)
    assert M.features(5.0 * np.eye(6) + 0.1 * np.ones((6, 6)))[8] == pytest.approx(0.0, abs=1e-7)


def test_features_13_14_power_step_changes_in_closed_form():
    """f13 / f14 = log10 of the relative change of the last power step.  For
    a diagonal matrix the Rayleigh quotients of both iterations are exact
    rationals: q_k = sum_i w_i e_i^(2k+1) / sum_i w_i e_i^(2k) with e = d,
    w = 1 (lambda_max, 16 steps from the ones vector) and e = c - d,
    w = s_i^2 (lambda_min, 24 steps on c I - A from the hash start s,
    c = 1.0001 lambda_max)."""
    from fractions import Fraction as Fr
    d = [2.0, 1.0, 0.875]  # exact in float32
    f = M.features(np.diag(d))

    def quotients(e, w, steps):
        q = []
        for k in range(steps):
            num = sum(wi * ei ** (2 * k + 1) for ei, wi in zip(e, w))
            den = sum(wi * ei ** (2 * k) for ei, wi in zip(e, w))
            q.append(num / den)
        return q

    e_max = [Fr(x) for x in d]
    q = quotients(e_max, [Fr(1)] * 3, M.POWER_MAX)
    lmax = q[-1]
    want13 = np.log10(float(abs(q[-1] - q[-2]) / lmax))
    assert f[13] == pytest.approx(want13, abs=1e-6)
    c = Fr(1.0001) * lmax
    s = [Fr(x) for x in M.start_vector(3)]
    mu = quotients([c - x for x in e_max], [si * si for si in s], M.POWER_MIN)
    want14 = np.log10(float(abs(mu[-1] - mu[-2]) / abs(mu[-1])))
    assert -6.0 < want14 < -1.0  # a change the features resolve, not the 1e-30 floor
    assert f[14] == pytest.approx(want14, abs=1e-6)


def test_features_of_a_diagonal_matrix():
    # Separated extremes.  The lambda_max iteration contracts by 5 / 10 per
    # step: after 16 steps its Rayleigh quotient is within ~30 * 2^-32 of 10
    # (relative 4e-9); the lambda_min one by 5.001 / 9.901, exact to rounding.
    d = np.array([10.0] + [5.0] * 30 + [0.1])
    f = M.features(np.diag(d))
    n, t = len(d), d.mean()
    assert f[0] == n / 256 and abs(f[1] - np.log2(n) / 8) < 1e-15
    np.testing.assert_allclose(f[2], 2.0, atol=1e-8)           # log10(10 / 0.1)
    np.testing.assert_allclose(f[3], np.log10(10 / t), atol=1e-8)
    np.testing.assert_allclose(f[4], np.log10(0.1 / t), atol=1e-8)
    np.testing.assert_allclose(f[5], 0.01, atol=1e-9)          # Gershgorin low / lmax
    np.testing.assert_allclose(f[6], 0.0, atol=1e-8)
    np.testing.assert_allclose(f[7], -2.0, atol=1e-12)
    assert f[9] == pytest.approx(6.0) and f[10] == pytest.approx(6.0)  # r_i = 0: clip 1e6
    assert f[11] == 0.0 and f[15] == 0.0
    np.testing.assert_allclose(f[12], np.log10(np.sqrt(np.sum(d * d) * n) / d.sum()), atol=1e-12)


def test_features_bound_the_spectrum_of_random_spd_matrices():
    rng = np.random.default_rng(5)
    for n in (16, 33, 100):
        Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
        lam = 10.0 ** rng.uniform(-3, 0, n)
        A = (Q * lam) @ Q.T
        A = 0.5 * (A + A.T)
        f = M.features(A)
        t = np.trace(A) / n
        lmax_est = t * 10 ** f[3]
        lmin_est = t * 10 ** f[4]
        # Rayleigh quotients of the float32-rounded matrix the power steps
        # use: lambda_max estimate from below, lambda_min from above.
        lam32 = np.linalg.eigvalsh(A.astype(np.float32).astype(np.float64))
        assert lmax_est <= lam32.max() * (1 + 1e-12) and lmax_est >= 0.5 * lam32.max()
        assert lmin_est >= lam32.min() * (1 - 1e-9)
        # ... which are the spectrum of A to float32 rounding.
        assert abs(lam32.max() - lam.max()) <= 2.0**-23 * np.linalg.norm(A)
        # Gershgorin discs contain the spectrum.
        assert f[5] * lmax_est <= lam.min() + 1e-12
        assert lmax_est * 10 ** f[6] >= lam.max() - 1e-12


def test_fold_standardisation_equals_standardised_inputs():
    from msq import selector as SEL
    rng = np.random.default_rng(6)
    p = _params(7).astype(np.float32)
    x = rng.standard_normal((9, M.INPUTS)) * 3 + 1
    mean, std = x.mean(0), x.std(0)
    z_std, _, _ = M.forward(p, (x - mean) / std)
    z_raw, _, _ = M.forward(SEL.fold_standardisation(p, mean, std), x)
    np.testing.assert_allclose(z_raw, z_std, rtol=1e-5, atol=1e-5)


def test_choice_minimises_expected_cost_and_policy_cost():
    from msq import selector as SEL
    # eight tiers, then the plain path (index 8, always succeeds); unit costs
    # in ms per (n / 256)^3; a failed choice pays L(n) = 20 ms plus the plain
    # path's share again.
    cost = np.array([3.0, 4.0, 5.0, 6.0, 6.0, 7.0, 7.5, 7.5, 10.0])
    sel = SEL.Selector(np.zeros(M.PARAMS, np.float32), cost, [1, 1000], [20.0, 20.0])
    orders = [256, 256]
    prob = np.array([[0.2, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0],    # tier 1 sure, 0 unlikely
                     [0.0, 0.3, 0.1, 0.0, 0.0, 0.0, 0.0, 0.0]])   # nothing likely: plain
    assert sel.choose(prob, orders).tolist() == [1, SEL.PLAIN]
    e = sel.expected(prob, orders)
    assert e[0, 0] == 3.0 + 0.8 * (20.0 + 10.0) and e[0, 1] == 4.0 and e[1, SEL.PLAIN] == 10.0
    prob[0, 1] = 0.97          # 4 + 0.03 * 30 = 4.9 < 5: still tier 1
    assert sel.choose(prob, orders).tolist() == [1, SEL.PLAIN]
    prob[0, 1] = 0.95          # 4 + 1.5 > 5: tier 2
    assert sel.choose(prob, orders).tolist() == [2, SEL.PLAIN]
    # candidates restrict the tiers; the plain path stays a candidate
    assert sel.choose(prob, orders, candidates=[6]).tolist() == [6, SEL.PLAIN]
    # a small matrix (n = 128, work 1/8): the latency term dominates
    assert sel.choose(np.array([[0.9] * 8]), [128]).tolist() == [SEL.PLAIN]
    success = np.array([[0, 1, 1, 1, 1, 1, 1, 1, 1], [0, 0, 0, 0, 0, 0, 0, 0, 1]], dtype=float)
    # Orders 256 and 128: work units 1 and 1/8; failures pay the plain path
    # in one second call of latency max L.
    total, fails = SEL.policy_cost(np.array([0, 1]), success, cost, [256, 128])
    assert fails == 2 and total == (3 + 10) + (4 + 10) / 8
    total, fails = SEL.policy_cost(np.array([0, 1]), success, cost, [256, 128],
                                   lambda o: np.asarray(o) / 10.0)
    assert total == (3 + 10) + (4 + 10) / 8 + 25.6
    total, fails = SEL.policy_cost(np.array([2, SEL.PLAIN]), success, cost, [256, 128])
    assert fails == 0 and total == 5 + 10 / 8
    np.testing.assert_allclose(SEL.unit_costs([9.0, 4.5], [256, 128]), [8.0, 4.0])


def test_batch_choice_drops_small_format_groups_and_prefers_a_single_choice():
    from msq import selector as SEL
    cost = np.array([1.0, 1.0, 1.0, 1.0, 2.0, 2.0, 3.0, 3.0, 6.0])
    sel = SEL.Selector(np.zeros(M.PARAMS, np.float32), cost)
    n = 100
    orders = [256] * n
    prob = np.zeros((n, 8))
    prob[:, 4] = 1.0           # the fixed (BF16) tier always works: the gate
    prob[:40, 0] = 1.0
    assert (sel.choose_batch(prob, orders) == SEL.FIXED).all()
    prob[:, 4] = 0.99          # 1% expected to fail on the fixed tier
    prob[:40, 0] = 0.0
    prob[:3, 0] = 1.0          # MXFP4 works for 3% of the work only
    ch = sel.choose(prob, orders)
    assert (ch[:3] == 0).all() and (ch[3:] == 4).all()
    # MXFP4's group (3% of the work) is dropped: one choice for the batch
    assert (sel.choose_batch(prob, orders) == 4).all()
    # 40% MXFP4: a saving beyond the single-choice margin, kept per matrix
    prob[:40, 0] = 1.0
    bc = sel.choose_batch(prob, orders)
    assert (bc[:40] == 0).all() and (bc[40:] == 4).all()


def test_gate_sample_is_stratified_by_order():
    """One matrix in GATE_SAMPLE_STRIDE of every order class, deterministic,
    in position order; small batches are scored whole."""
    from msq import selector as SEL
    assert SEL.gate_sample([64] * (SEL.GATE_SAMPLE_MIN - 1)) is None
    rng = np.random.default_rng(4)
    orders = (16 * (1 + rng.integers(0, 16, 4096))).tolist()
    sample = SEL.gate_sample(orders)
    assert np.array_equal(sample, SEL.gate_sample(orders))
    assert np.all(np.diff(sample) > 0) and len(sample) == 4096 // SEL.GATE_SAMPLE_STRIDE
    o = np.asarray(orders)
    for n in np.unique(o):
        members = int(np.sum(o == n))
        picked = int(np.sum(o[sample] == n))
        assert abs(picked - members / SEL.GATE_SAMPLE_STRIDE) <= 1
    # a batch sorted by order is sampled across its whole range too
    sorted_orders = sorted(orders)
    s2 = SEL.gate_sample(sorted_orders)
    assert s2[0] < SEL.GATE_SAMPLE_STRIDE and s2[-1] >= 4096 - SEL.GATE_SAMPLE_STRIDE


def test_forward_bf16_rounding_and_closed_forms():
    """Pins of the BF16 forward: the BF16 rounding of float32 values (ties to
    even, NaN quieted, exhaustive against the codec oracle on a sample); an
    exact network on small integers; and the rounding of a hidden activation
    at 1 + 2^-9 (a tie between BF16 1 and 1 + 2^-8: rounds to 1)."""
    from oracle import codecs as C
    rng = np.random.default_rng(8)
    v = rng.standard_normal(20000).astype(np.float32) * np.float32(1e3)
    v[:4] = [np.nan, np.inf, -0.0, 1.0 + 2.0 ** -9]
    codes, _ = C.encode("bf16", v)
    want = C.decode("bf16", codes, None, v.size)
    got = M._bf16(v)
    assert np.array_equal(got.view(np.uint32), np.asarray(want, np.float32).view(np.uint32))
    # W1 = identity on the first 16 hidden units, W2 = identity, W3 sums the
    # first 16 hidden units into output 0; integer inputs: exact.
    parts = {k: np.zeros(s) for k, s in M.SIZES}
    parts["W1"][:16, :16] = np.eye(16)
    parts["W2"][:16, :16] = np.eye(16)
    parts["W3"][0, :16] = 1.0
    parts["b3"][1] = 3.0
    params = M.pack(parts).astype(np.float32)
    x = np.arange(32, dtype=np.float32).reshape(2, 16) - 8.0
    z = M.forward_bf16(params, x)
    relu = np.maximum(x, 0)
    assert z[0, 0] == relu[0].sum() and z[1, 0] == relu[1].sum() and np.all(z[:, 1] == 3.0)
    parts = {k: np.zeros(s) for k, s in M.SIZES}
    parts["b1"][0] = 1.0 + 2.0 ** -9   # bias rounds to 1 in BF16
    parts["W2"][0, 0] = 1.0
    parts["W3"][0, 0] = 1.0
    z = M.forward_bf16(M.pack(parts).astype(np.float32), np.zeros((1, 16), np.float32))
    assert z[0, 0] == 1.0
    # float32 network with the same (representable) weights differs only by
    # the activations' BF16 rounding
    p = (rng.standard_normal(M.PARAMS) * 0.3).astype(np.float32)
    p = M._bf16(p)
    x = M._bf16(rng.standard_normal((64, 16)).astype(np.float32))
    z16 = M.forward_bf16(p, x)
    z64, _, _ = M.forward(p.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(z16 - z64)) <= 0.05 * (np.abs(z64).max() + 1)


def test_batch_options_records():
    """Per-matrix options reach the library as raw msq_solve_options records
    equal byte for byte to the ctypes structures; a batch that runs the fixed
    tier as a whole stops its orders above FIXED_BATCH_LARGE_ORDER at
    FIXED_BATCH_LARGE_RATE and the rest at the tier's rate; any other single
    choice is one record."""
    import ctypes
    from msq import capi, selector as SEL

    opts = SEL.choice_options()
    assert len(opts) == len(SEL.CHOICES) + 1

    def rec(raw):
        return capi.SolveOptions.from_buffer_copy(raw.tobytes())

    orders = np.array([16, 128, 129, 144, 256, 64], dtype=np.uint32)
    fixed = np.full(len(orders), SEL.FIXED)
    arr = SEL.batch_options(fixed, orders)
    assert arr.dtype == capi.OPTIONS_RECORD and arr.itemsize == ctypes.sizeof(capi.SolveOptions)
    for n, r in zip(orders, arr):
        o = rec(r)
        want = SEL.FIXED_BATCH_LARGE_RATE if n > SEL.FIXED_BATCH_LARGE_ORDER else SEL.TIERS[SEL.FIXED][2]
        assert o.jacobi.correction_rate == np.float32(want)
        assert o.correction_precision == opts[SEL.FIXED].correction_precision
        assert o.max_corrections == opts[SEL.FIXED].max_corrections
    # small orders only: the tier's own record
    small = SEL.batch_options(fixed[:2], orders[:2])
    assert isinstance(small, capi.SolveOptions)
    assert bytes(small) == bytes(opts[SEL.FIXED])
    # another single choice: its record
    one = SEL.batch_options(np.full(len(orders), SEL.PLAIN), orders)
    assert bytes(one) == bytes(opts[SEL.PLAIN])
    # mixed choices: record k for matrix i, byte for byte
    mixed = np.array([0, 4, 8, 5, 6, 1])
    arr = SEL.batch_options(mixed, orders)
    for k, r in zip(mixed, arr):
        assert r.tobytes() == bytes(opts[k])
