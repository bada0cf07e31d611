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
    cost = np.array([10.0, 7.0, 6.0, 5.0, 4.0, 3.0, 7.5, 7.5])
    sel = SEL.Selector(np.zeros(M.PARAMS, np.float32), cost)
    prob = np.array([[0.0, 1.0, 1.0, 1.0, 1.0, 0.2, 1.0, 1.0],    # r4 sure, r8 unlikely
                     [0.0, 0.3, 0.1, 0.0, 0.0, 0.0, 0.0, 0.0]])   # nothing likely: plain
    ch = sel.choose(prob, fallback_weight=1.0)
    assert ch.tolist() == [4, 0]
    # A heavier fallback weight turns away from the uncertain tiers only.
    assert sel.choose(prob, fallback_weight=10.0).tolist() == [4, 0]
    prob[0, 4] = 0.95
    assert sel.choose(prob, fallback_weight=1.0).tolist() == [4, 0]
    assert sel.choose(prob, fallback_weight=10.0).tolist() == [3, 0]
    success = np.array([[1, 1, 1, 1, 1, 0, 1, 1], [1, 0, 0, 0, 0, 0, 0, 0]], dtype=float)
    # Orders 256 and 128: work units 1 and 1/8.
    total, fails = SEL.policy_cost(np.array([5, 1]), success, cost, [256, 128])
    assert fails == 2 and total == (3 + 10) + (7 + 10) / 8
    total, fails = SEL.policy_cost(ch, success, cost, [256, 128])
    assert fails == 0 and total == 4 + 10 / 8
    # Unit costs: batch time over total work.
    np.testing.assert_allclose(SEL.unit_costs([9.0, 4.5], [256, 128]), [8.0, 4.0])