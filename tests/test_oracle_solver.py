"""CPU checks pinning the solver oracle (oracle/solver.py).

* schedule: closed-form round-robin properties of makeJacobiSchedule
  (src/metal_solver.cpp:50-79) and the reference's table shapes;
* Jacobi: eigenvalues against LAPACK, orthogonality, off-norm decrease, the
  reference's fixed 8 sweeps;
* products: the reference's known answer for the tiled kernels (identity basis,
  rhs 2, H = 0.5 gives a correction of exactly 1; src/metal_checks.cpp:537-601);
* end to end: the frozen-factor iteration reaches the FP64 LAPACK root and the
  construction-known root, and the CPU plain path (1_baseline/spd.cpp:348-430)
  reproduces the reference's behaviour on its controlled problem (residual of
  X0 ~1e-6, ~1e-12 after one correction; recorded in the reference's own
  probe run on n = 64, cond 100).
"""
import numpy as np
import pytest

from oracle import solver as S
from msq import generators as G


@pytest.mark.parametrize("n", [2, 3, 7, 16, 33])
def test_schedule_is_round_robin(n):
    padded, partners, pair_indices, pair_members = S.jacobi_schedule(n)
    assert padded == n + n % 2
    assert partners.shape == (padded - 1, padded)
    seen = set()
    for r in range(padded - 1):
        # every round is a perfect matching of 0..padded-1
        assert sorted(pair_members[r]) == list(range(padded))
        for k in range(padded // 2):
            p, q = int(pair_members[r, 2 * k]), int(pair_members[r, 2 * k + 1])
            assert p < q and partners[r, p] == q and partners[r, q] == p
            assert pair_indices[r, p] == k and pair_indices[r, q] == k
            seen.add((p, q))
    # every pair exactly once per sweep
    assert len(seen) == padded * (padded - 1) // 2


def test_schedule_closed_form():
    """Slot x of round t holds P_t(x) = 1 + ((x - 1 - t) mod (N - 1)), P_t(0) = 0;
    slot k pairs with slot N - 1 - k (the form the HIP kernels use)."""
    for n in (6, 9, 32):
        padded, _, _, pair_members = S.jacobi_schedule(n)
        for t in range(padded - 1):
            P = [0] + [1 + ((x - 1 - t) % (padded - 1)) for x in range(1, padded)]
            for k in range(padded // 2):
                a, b = P[k], P[padded - 1 - k]
                assert (min(a, b), max(a, b)) == (int(pair_members[t, 2 * k]),
                                                  int(pair_members[t, 2 * k + 1]))


def _spd(n, cond, seed=1):
    Q = G.haar_orthogonal([G.matrix_seed(7, seed)], n)[0]
    lam = np.geomspace(1.0 / cond, 1.0, n)
    A, X = G.assemble(Q, lam)
    return A, X, lam


@pytest.mark.parametrize("n", [5, 16, 24])
def test_jacobi_eigenvalues_and_orthogonality(n):
    A, X, lam = _spd(n, 100.0)
    V, roots, info = S.jacobi_f32(A, sweeps=8)
    d = np.sort(np.asarray(info["diagonal"], np.float64))
    np.testing.assert_allclose(d, np.linalg.eigvalsh(A), rtol=1e-4, atol=1e-6)
    assert S.orthogonality_error(V) < 1e-5
    VAV = V.astype(np.float64).T @ A @ V.astype(np.float64)
    off = np.sqrt(np.sum(VAV**2) - np.sum(np.diag(VAV) ** 2))
    assert off < 1e-5 * np.linalg.norm(A)


def test_jacobi_tolerance_mode_stops():
    A, _, _ = _spd(20, 1e3)
    _, _, info = S.jacobi_f32(A, off_tolerance=1e-5, max_sweeps=30, return_history=True)
    h = info["history"]
    assert info["sweeps"] < 30 and h[-1] <= 1e-5 * info["matrix_norm"]
    assert all(b < a for a, b in zip(h, h[1:]))


def test_fp32_angles_match_reference_formula():
    """angle_dtype=float32 restates jacobi_rotations_f32 exactly; float64 angles
    give the same rotation to float32 precision."""
    A, _, _ = _spd(12, 10.0)
    V32, r32, _ = S.jacobi_f32(A, sweeps=2, angle_dtype=np.float32)
    V64, r64, _ = S.jacobi_f32(A, sweeps=2, angle_dtype=np.float64)
    np.testing.assert_allclose(r32, r64, rtol=1e-5)


def test_tiled_kernel_known_answer():
    """src/metal_checks.cpp:537-601: identity basis and roots 1 give X0 = I, and
    rhs 2 with H = 0.5 gives a correction of exactly 1."""
    n = 16
    V = np.eye(n, dtype=np.float32)
    roots = np.ones(n, np.float32)
    np.testing.assert_array_equal(S.reconstruct_f32(V, roots), np.eye(n, dtype=np.float32))
    _, H = S.frozen_factor(V, roots)
    assert np.all(H == 0.5)
    E = S.correction_f32(V, H, np.full((n, n), 2.0))
    np.testing.assert_array_equal(E, np.ones((n, n), np.float32))


@pytest.mark.parametrize("cond", [10.0, 1e3])
def test_solve_reaches_fp64_root(cond):
    A, X, lam = _spd(32, cond, seed=3)
    Xref = S.reference_root(A)
    assert S.relative_error(Xref, X) < 1e-13
    Xs, info = S.solve_gpu_resident(A, sweeps=8, max_iterations=6, tolerance=1e-14)
    assert S.relative_error(Xs, Xref) < 1e-9
    r = info["residuals"]
    assert r[-1] < r[0]
    assert S.relative_symmetry_error(Xs) < 1e-12


def test_residual_closed_form():
    A, X, _ = _spd(10, 10.0)
    R, rel = S.residual_fp64(A, X)
    assert rel < 1e-14
    R2, rel2 = S.residual_fp64(A, 1.01 * X)
    np.testing.assert_allclose(R2, A - 1.0201 * (X @ X), atol=1e-13)


def test_accelerate_mixed_controlled_problem():
    """The CPU plain path on the reference's controlled problem (n = 64,
    cond 100): X0 residual ~1e-6, then ~1e-12, then FP64 level."""
    A = S.controlled_matrix(64, 100.0)
    Xref = S.reference_root(A)
    X, steps = S.accelerate_mixed(A, Xref, max_corrections=2, tolerance=0.0)
    res = [s[0] for s in steps]
    assert 1e-7 < res[0] < 1e-5
    assert res[1] < 1e-10
    assert steps[-1][1] < 1e-12


@pytest.mark.parametrize("sweeps", [1, 2, 3])
def test_tile_jacobi_at_order_16_is_one_scalar_sweep(sweeps):
    """One block pair: a blocked sweep is one sweep of the circle schedule over
    all 16 indices, so the blocked oracle must follow jacobi_f32 to float
    rounding (a wrong sign, index or accumulation order would not)."""
    A = G.single_matrix(9, 16, 1e-2, index=0)[0]
    V1, d1, _ = S.jacobi_tile(A, sweeps=sweeps)
    V2, _, info = S.jacobi_f32(A, sweeps=sweeps, angle_dtype=np.float64)
    scale = np.linalg.norm(A)
    assert np.max(np.abs(d1 - info["diagonal"])) <= 1e-5 * scale
    assert np.max(np.abs(V1 - V2)) <= 1e-4


@pytest.mark.parametrize("n", [40, 150])
def test_tile_jacobi_converges_to_lapack(n):
    """Larger orders: eigenvalues against LAPACK, V orthogonal, V^T A V equal to
    the rotated matrix, off(A) decreasing sweep by sweep."""
    A = G.single_matrix(9, n, 1e-3, index=1)[0]
    V, d, info = S.jacobi_tile(A, off_tolerance=1e-12, return_history=True)
    lam = np.linalg.eigvalsh(A)
    assert np.max(np.abs(np.sort(d) - lam)) <= 1e-11 * np.linalg.norm(A)
    assert np.max(np.abs(V.T @ V - np.eye(n))) <= 1e-12
    assert np.max(np.abs(V.T @ A @ V - info["rotated"])) <= 1e-12
    h = info["history"]
    assert all(b < a for a, b in zip(h, h[1:]))
    assert info["sweeps"] <= 12
