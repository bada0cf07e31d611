import numpy as np

from msq import generators as G


def test_haar_orthogonal_and_deterministic():
    Q1 = G.haar_orthogonal([G.matrix_seed(3, 5), G.matrix_seed(3, 6)], 24)
    Q2 = G.haar_orthogonal([G.matrix_seed(3, 5)], 24)
    np.testing.assert_array_equal(Q1[0], Q2[0])
    for Q in Q1:
        np.testing.assert_allclose(Q.T @ Q, np.eye(24), atol=1e-13)
    assert not np.allclose(Q1[0], Q1[1])


def test_spectra_pin_condition():
    u = np.random.default_rng(0).random(200)
    lam = G.log_uniform_spectrum(u, 64, 1e4)
    assert lam.min() == 1e-4 and lam.max() == 1.0
    lam = G.clustered_spectrum(u, 64, 1e3)
    assert lam.min() == 1e-3 and lam.max() == 1.0
    assert np.all(lam > 0)
    lam = G.power_law_spectrum(10, 2.0)
    np.testing.assert_allclose(lam, 1.0 / np.arange(1, 11) ** 2)


def test_batched_uneven_shape():
    mats = G.batched_uneven(count=24)
    orders = [m["order"] for m in mats]
    assert all(o % 16 == 0 and 16 <= o <= 256 for o in orders)
    for m in mats[:6]:
        A, X = m["A"], m["X"]
        np.testing.assert_allclose(X @ X, A, atol=1e-12)
        w = np.linalg.eigvalsh(A)
        np.testing.assert_allclose(w.max() / w.min(), m["condition"], rtol=1e-6)
    assert [m["kind"] for m in mats[:2]] == ["log_uniform", "clustered"]


def test_mixture_kinds():
    mats = list(G.mixture(count=9, order_range=(32, 40)))
    assert {m["kind"] for m in mats} <= {"log_uniform", "power_law", "clustered"}
    assert all(32 <= m["order"] <= 40 for m in mats)
