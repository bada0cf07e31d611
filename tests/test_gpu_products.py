"""MFMA products, the binary64 residual and the Jacobi kernels against the
oracle and closed forms."""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, say, jacobi_method
from msq import capi
from msq import generators as G
from oracle import codecs as C
from oracle import solver as S

pytestmark = pytest.mark.gpu

TIERS = [capi.FP32, capi.BF16, capi.FP8_E4M3FN_SCALED, capi.MXFP4_E2M1]


def random_basis(n, seed):
    return G.haar_orthogonal([G.matrix_seed(11, seed)], n)[0].astype(np.float32)


def quantized(tier, M):
    """Tier rounding of a float matrix stored K-contiguous (rows of M)."""
    M = np.ascontiguousarray(M, np.float32)
    if tier == capi.FP32:
        return M.astype(np.float64)
    if tier == capi.BF16:
        c, _ = C.encode("bf16", M.reshape(-1))
        return C.decode("bf16", c, None, M.size).reshape(M.shape).astype(np.float64)
    if tier == capi.FP8_E4M3FN_SCALED:
        c, s = C.fp8_scaled(M.reshape(-1))
        return (C.decode("fp8_e4m3fn", c, None, M.size).reshape(M.shape) / np.float64(s))
    # MXFP4 blocks of 32 along the contiguous (K) direction of each row
    ld = M.shape[1]
    c, s = C.encode("mxfp4_e2m1", M.reshape(-1))
    return C.decode("mxfp4_e2m1", c, s, M.size).reshape(M.shape).astype(np.float64)


@pytest.mark.parametrize("tier", TIERS)
def test_reconstruction_products(runtime, tier):
    orders = [16, 32, 48, 64, 96, 200]
    Vs = [random_basis(n, i) for i, n in enumerate(orders)]
    roots = [np.geomspace(0.01, 1.0, n).astype(np.float32) for n in orders]
    o, vflat = capi.pack_batch(Vs, np.float32)
    _, _, voffs, total, vtotal = capi.batch_layout(orders)
    rflat = np.zeros(vtotal, np.float32)
    for r, v0 in zip(roots, voffs):
        rflat[v0:v0 + r.size] = r
    out = torch.zeros(total, dtype=torch.float32, device="cuda")
    say(f"reconstruct tier {tier}")
    runtime.reconstruct(tier, o, dev(vflat), dev(rflat), out)
    got = capi.unpack_batch(o, host(out))
    for V, r, X, n in zip(Vs, roots, got, orders):
        ld = (n + 31) // 32 * 32
        Gt = np.zeros((ld, ld), np.float32)       # rows of G = V diag(sqrt s)
        Gt[:n, :n] = V * np.sqrt(r)[None, :]
        Gq = quantized(tier, Gt)[:n, :n]
        want = Gq @ Gq.T
        # FP32 accumulation bound (3_spectral/product_benchmark.cpp:422-444 form);
        # the FP8 MFMA (f8f6f4, 128 K per instruction) accumulates inside an
        # instruction more coarsely than FP32, so FP8 gets u = 2^-16 per term.
        u = 2.0**-16 if tier == capi.FP8_E4M3FN_SCALED else 2.0**-24
        bound = (2 * n + 8) * u * np.linalg.norm(Gq) ** 2
        assert np.linalg.norm(X - want) <= bound, (tier, n)


@pytest.mark.parametrize("tier", TIERS)
def test_correction_known_answer(runtime, tier):
    """src/metal_checks.cpp:537-601: identity basis, roots 1, rhs 2 -> 1."""
    orders = [16, 16, 64]
    o, vflat = capi.pack_batch([np.eye(n, dtype=np.float32) for n in orders], np.float32)
    _, rflat = capi.pack_batch([np.full((n, n), 2.0, np.float32) for n in orders], np.float32)
    _, _, voffs, total, vtotal = capi.batch_layout(orders)
    roots = np.zeros(vtotal, np.float32)
    for n, v0 in zip(orders, voffs):
        roots[v0:v0 + n] = 1.0
    out = torch.zeros(total, dtype=torch.float32, device="cuda")
    runtime.correction(tier, o, dev(vflat), dev(roots), dev(rflat), out)
    for E, n in zip(capi.unpack_batch(o, host(out)), orders):
        tol = 0 if tier != capi.FP8_E4M3FN_SCALED else 1e-6
        np.testing.assert_allclose(E, np.ones((n, n)), rtol=0, atol=tol)
    # reconstruction of the identity basis with unit roots is the identity
    X = torch.zeros(total, dtype=torch.float32, device="cuda")
    runtime.reconstruct(tier, o, dev(vflat), dev(roots), X)
    for Xi, n in zip(capi.unpack_batch(o, host(X)), orders):
        np.testing.assert_allclose(Xi, np.eye(n), rtol=0, atol=tol)


def test_correction_fp32_matches_oracle(runtime):
    orders = [32, 80, 128]
    As = []
    for i, n in enumerate(orders):
        A, X, _ = G.single_matrix(5, n, 1e-2, index=i)
        As.append(A)
    Vs, roots, Rs = [], [], []
    for A in As:
        V, r, _ = S.jacobi_f32(A, sweeps=6)
        X = S.symmetrize(S.reconstruct_f32(V, r))
        R, _ = S.residual_fp64(A, X)
        Vs.append(V)
        roots.append(r)
        Rs.append(R.astype(np.float32))
    o, vflat = capi.pack_batch(Vs, np.float32)
    _, rflat = capi.pack_batch(Rs, np.float32)
    _, _, voffs, total, vtotal = capi.batch_layout(orders)
    sflat = np.zeros(vtotal, np.float32)
    for r, v0 in zip(roots, voffs):
        sflat[v0:v0 + r.size] = r
    out = torch.zeros(total, dtype=torch.float32, device="cuda")
    runtime.correction(capi.FP32, o, dev(vflat), dev(sflat), dev(rflat), out)
    for E, V, r, R in zip(capi.unpack_batch(o, host(out)), Vs, roots, Rs):
        _, H = S.frozen_factor(V, r)
        want = S.correction_f32(V, H, R)
        assert np.linalg.norm(E - want) <= 1e-5 * np.linalg.norm(want)


def test_residual_fp64(runtime):
    orders = [16, 50, 130]
    As, Xs = [], []
    for i, n in enumerate(orders):
        A, X, _ = G.single_matrix(6, n, 1e-3, index=i)
        As.append(A)
        Xs.append(X * (1 + 1e-4 * np.random.default_rng(i).standard_normal((n, n))))
        Xs[-1] = 0.5 * (Xs[-1] + Xs[-1].T)
    o, aflat = capi.pack_batch(As)
    _, xflat = capi.pack_batch(Xs)
    _, _, _, total, _ = capi.batch_layout(orders)
    R = torch.zeros(total, dtype=torch.float64, device="cuda")
    R32 = torch.zeros(total, dtype=torch.float32, device="cuda")
    rel = runtime.residual_fp64(o, dev(aflat), dev(xflat), R, R32)
    for Ri, A, X, ri in zip(capi.unpack_batch(o, host(R)), As, Xs, rel):
        want, wrel = S.residual_fp64(A, X)
        assert np.max(np.abs(Ri - want)) <= 1e-14 * np.abs(X).max() ** 2 * A.shape[0]
        assert abs(ri - wrel) <= 1e-10 * wrel


def _jacobi_batch(runtime, mats, options):
    orders = [m.shape[0] for m in mats]
    o, aflat = capi.pack_batch(mats)
    _, _, _, total, vtotal = capi.batch_layout(orders)
    V = torch.zeros(total, dtype=torch.float32, device="cuda")
    D = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    Rt = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    reps = runtime.jacobi_batched(o, dev(aflat), V, D, Rt, options)
    return (capi.unpack_batch(o, host(V)), capi.unpack_vectors(o, host(D)),
            capi.unpack_vectors(o, host(Rt)), reps)


def test_jacobi_fixed_sweeps_parity(runtime):
    """The reference's fixed 8 sweeps (src/metal_solver.cpp:210-213) against
    the oracle restatement on the same inputs."""
    orders = [5, 16, 31, 64, 100, 128, 150, 256]
    mats = [G.single_matrix(7, n, 1e-2, index=i)[0] for i, n in enumerate(orders)]
    say("jacobi fixed 8 sweeps", orders)
    # The scalar eigensolver at every order: the oracle's schedule.
    with jacobi_method(runtime, capi.JACOBI_SCALAR):
        Vs, Ds, Rts, reps = _jacobi_batch(runtime, mats, capi.jacobi_options(fixed_sweeps=8))
    for A, V, d, rt, rep in zip(mats, Vs, Ds, Rts, reps):
        n = A.shape[0]
        assert rep.sweeps == 8 and not rep.invalid_spectrum
        # The library pads orders to a multiple of 16 (zero rows beyond n);
        # the oracle runs the same schedule (for n % 16 == 0 it is the
        # reference's own).
        Vo, ro, info = S.jacobi_f32(A, sweeps=8, pad_to=(n + 15) // 16 * 16)
        # Hoffman-Wielandt: the sorted diagonal of an orthogonal similarity of A
        # is within its off-diagonal norm of the eigenvalues; the FP32 iterates
        # also carry the rounding of S (n - 1) rounds, ~ eps ||A||_F sqrt(S n).
        lam = np.linalg.eigvalsh(A)
        drift = 2.0 * 2.0**-23 * np.linalg.norm(A) * np.sqrt(8.0 * (n - 1))
        gpu_gap = np.linalg.norm(np.sort(d) - lam)
        ora_gap = np.linalg.norm(np.sort(info["diagonal"]) - lam)
        assert gpu_gap <= rep.off_norm + drift, (n, gpu_gap, rep.off_norm, drift)
        assert ora_gap <= info["off_norm"] + drift, (n, ora_gap)
        if n <= 128:
            # Converged in 8 sweeps: the diagonals agree to the FP32 rounding
            # that each diagonal entry collects over its 8 (n - 1) updates, a
            # random walk of ~2^-24 max|lambda| per update (4 sigma here); the
            # two sides round differently (FMA contraction), so no tighter.
            walk = 4.0 * 2.0**-24 * np.abs(lam).max() * np.sqrt(8.0 * (n - 1))
            np.testing.assert_allclose(np.sort(d), np.sort(info["diagonal"]), rtol=0,
                                       atol=walk)
            np.testing.assert_allclose(np.sort(d), lam, rtol=0, atol=walk + rep.off_norm)
        np.testing.assert_allclose(rt, np.sqrt(d), rtol=1e-7)
        assert S.orthogonality_error(V) < 2e-6 * np.sqrt(n)
        # V diag(d) V^T reproduces A up to the off-diagonal part left
        err = np.linalg.norm((V * d) @ V.T - A)
        assert err <= rep.off_norm + 2e-6 * np.sqrt(n) * np.linalg.norm(A)


def test_jacobi_tolerance_mode(runtime):
    orders = [16, 64, 128, 200, 256]
    mats = [G.single_matrix(8, n, 1e-4, index=i)[0] for i, n in enumerate(orders)]
    say("jacobi tolerance", orders)
    Vs, Ds, _, reps = _jacobi_batch(runtime, mats, capi.jacobi_options(off_tol=1e-7))
    for A, V, d, rep in zip(mats, Vs, Ds, reps):
        assert 3 <= rep.sweeps < 30
        hist = list(rep.off_history[: rep.sweeps])
        assert hist[-1] <= 1e-5
        VAV = V.astype(np.float64).T @ A @ V.astype(np.float64)
        off = np.sqrt(max(np.sum(VAV**2) - np.sum(np.diag(VAV) ** 2), 0))
        assert off <= 2e-5 * np.linalg.norm(A)


@pytest.mark.parametrize("fixed", [3, 0])
def test_jacobi_orders_with_v_in_device_memory(runtime, fixed):
    """The scalar eigensolver (MSQ_JACOBI_SCALAR).  Padded orders 144..256 keep
    V in device memory and update it once per
    group of rounds, as many as the LDS beside A holds history for (24 to 144
    rounds), groups ending at sweep boundaries.  One matrix per padded order
    (odd orders, so the padding is live), with fixed sweeps (the last group
    of each sweep is one round short of whole chunks) and with the tolerance
    stop.  V must stay in step with A's rounds: the off-norm of V^T A V
    (binary64, from the returned V) equals the kernel's own off-norm of its
    rotated A to the FP32 drift, V is orthogonal and V diag(d) V^T
    reproduces A up to the off-norm left.  After 3 of the oracle's sweeps of
    the same padded schedule the off-norm is of the same size; it is not
    compared closer, because the early iterates are chaotic (float32 against
    binary64 angles, or a 1e-7 relative perturbation of A, move the oracle's
    own off-norm after 3 sweeps by 2 to 9% on these matrices)."""
    orders = [137, 150, 170, 185, 200, 215, 230, 250]  # padded 144, 160, ..., 256
    mats = [G.single_matrix(12, n, 1e-3, index=i)[0] for i, n in enumerate(orders)]
    say("jacobi V in device memory", orders, "fixed", fixed)
    opts = capi.jacobi_options(fixed_sweeps=fixed) if fixed else capi.jacobi_options(off_tol=1e-7)
    with jacobi_method(runtime, capi.JACOBI_SCALAR):
        Vs, Ds, _, reps = _jacobi_batch(runtime, mats, opts)
    for A, V, d, rep in zip(mats, Vs, Ds, reps):
        n = A.shape[0]
        N = (n + 15) // 16 * 16
        normA = np.linalg.norm(A)
        drift = 2.0 * 2.0**-23 * normA * np.sqrt(max(rep.sweeps, 1) * N)
        V64 = V.astype(np.float64)
        VAV = V64.T @ A @ V64
        off = np.sqrt(max(np.sum(VAV**2) - np.sum(np.diag(VAV) ** 2), 0.0))
        assert abs(off - rep.off_norm) <= 2.0 * drift, (n, off, rep.off_norm, drift)
        if fixed:
            assert rep.sweeps == fixed
            _, _, info = S.jacobi_f32(A, sweeps=fixed, pad_to=N)
            assert 0.75 * info["off_norm"] <= rep.off_norm <= 1.33 * info["off_norm"], (
                n, rep.off_norm, info["off_norm"])
        else:
            assert 3 <= rep.sweeps < 30
            assert rep.off_history[rep.sweeps - 1] <= 1e-5
        lam = np.linalg.eigvalsh(A)
        assert np.linalg.norm(np.sort(d) - lam) <= rep.off_norm + drift
        assert S.orthogonality_error(V) < 2e-6 * np.sqrt(n)
        err = np.linalg.norm((V * d) @ V.T - A)
        assert err <= rep.off_norm + 4e-6 * np.sqrt(n) * normA, (n, err, rep.off_norm)


@pytest.mark.parametrize("n", [257, 300, 520])
def test_jacobi_large_order_blocked_path(runtime, n):
    """Orders above 256: the blocked Jacobi (64 x 64 subproblems, rotations as
    MFMA products; padded to a multiple of 64).  Eigenvalues against LAPACK
    within the off-norm left plus the FP32 drift of the sweeps, V
    orthogonal, V diag V^T reproduces A."""
    A = G.single_matrix(9, n, 1e-2)[0]
    say("jacobi blocked", n)
    Vs, Ds, _, reps = _jacobi_batch(runtime, [A], capi.jacobi_options(off_tol=1e-7))
    rep = reps[0]
    lam = np.linalg.eigvalsh(A)
    drift = 2.0 * 2.0**-23 * np.linalg.norm(A) * np.sqrt(rep.sweeps * n)
    assert np.linalg.norm(np.sort(Ds[0]) - lam) <= rep.off_norm + drift
    assert rep.off_norm <= 1e-7 * rep.matrix_norm or rep.sweeps >= 3
    assert S.orthogonality_error(Vs[0]) < 2e-6 * np.sqrt(n)
    err = np.linalg.norm((Vs[0] * Ds[0]) @ Vs[0].T - A)
    assert err <= rep.off_norm + 4e-6 * np.sqrt(n) * np.linalg.norm(A)


def test_jacobi_large_orders_batched_classes(runtime):
    """Many orders above 256 in one batch: the matrices of each padded order
    advance together (one launch per kernel per block round for the class)
    and leave the active list as they stop.  Every matrix's eigenvectors,
    diagonal and report equal those of the same matrix solved alone, bit for
    bit, and pass the accuracy checks of the single-matrix test."""
    orders = [300, 260, 520, 400, 64, 300, 257, 384]
    conds = [1e2, 1e4, 1e1, 1e3, 1e2, 1e5, 1e3, 1e2]
    mats = [G.single_matrix(11, n, 1.0 / c, index=i)[0]
            for i, (n, c) in enumerate(zip(orders, conds))]
    say("jacobi blocked batch", orders)
    opts = capi.jacobi_options(off_tol=1e-7)
    Vs, Ds, _, reps = _jacobi_batch(runtime, mats, opts)
    for k in (0, 2, 5):
        Va, Da, _, ra = _jacobi_batch(runtime, [mats[k]], opts)
        assert np.array_equal(Va[0], Vs[k]) and np.array_equal(Da[0], Ds[k])
        assert ra[0].sweeps == reps[k].sweeps and ra[0].off_norm == reps[k].off_norm
    for A, V, d, rep in zip(mats, Vs, Ds, reps):
        n = A.shape[0]
        lam = np.linalg.eigvalsh(A)
        drift = 2.0 * 2.0**-23 * np.linalg.norm(A) * np.sqrt(rep.sweeps * n)
        assert np.linalg.norm(np.sort(d) - lam) <= rep.off_norm + drift
        assert S.orthogonality_error(V) < 2e-6 * np.sqrt(n)
        err = np.linalg.norm((V * d) @ V.T - A)
        assert err <= rep.off_norm + 4e-6 * np.sqrt(n) * np.linalg.norm(A)


def test_orthogonality_measure(runtime):
    orders = [16, 64]
    Vs = [random_basis(n, 3 + i) for i, n in enumerate(orders)]
    Vs[1] = Vs[1] * np.float32(1.001)
    o, vflat = capi.pack_batch(Vs, np.float32)
    got = runtime.orthogonality(o, dev(vflat))
    for g, V in zip(got, Vs):
        assert abs(g - S.orthogonality_error(V)) <= 1e-6 + 1e-3 * S.orthogonality_error(V)


def test_mxfp4_two_term_reconstruction(runtime):
    """The solve's MXFP4 reconstruction: G = V diag(sqrt s) carried as hi + lo
    (hi the MXFP4 rounding of G, lo that of G - hi in float), X0 =
    sym(hi hi^T + 2 hi lo^T) = hi hi^T + hi lo^T + lo hi^T, against the same
    expansion built on the host from the solve's own Jacobi basis; and the
    expansion is much closer to the root than hi hi^T alone."""
    orders = [32, 64, 100, 256]
    mats = [G.single_matrix(15, n, 1e-3, index=i)[0] for i, n in enumerate(orders)]
    jac = capi.jacobi_options(off_tol=1e-7)
    opts = capi.solve_options(capi.MXFP4_E2M1, capi.BF16, jac, min_corrections=0,
                              max_corrections=0, residual_tolerance=0.0)
    o, aflat = capi.pack_batch(mats)
    _, offs, voffs, total, vtotal = capi.batch_layout(o)
    A = dev(aflat)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    say("mxfp4 two-term reconstruction", orders)
    runtime.solve_batched(o, A, X, opts)
    V = torch.zeros(total, dtype=torch.float32, device="cuda")
    D = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    Rt = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    runtime.jacobi_batched(o, A, V, D, Rt, jac)
    Vh, Rh = host(V), host(Rt)
    got = capi.unpack_batch(o, host(X))
    for M, Xg, n, off, v0 in zip(mats, got, orders, offs, voffs):
        ld = (n + 31) // 32 * 32
        Vm = Vh[off:off + ld * ld].reshape(ld, ld, order="F")
        s = Rh[v0:v0 + ld]
        Gs = (Vm * np.sqrt(s)[None, :]).astype(np.float32)   # rows of G, K-contiguous
        hi = quantized(capi.MXFP4_E2M1, Gs)
        lo = quantized(capi.MXFP4_E2M1, (Gs - hi.astype(np.float32)).astype(np.float32))
        C = hi @ hi.T + 2.0 * (hi @ lo.T)
        want = (0.5 * (C + C.T))[:n, :n]
        u = 2.0**-24
        bound = 2 * ((2 * n + 8) * u * (np.linalg.norm(hi) ** 2 + 2 * np.linalg.norm(hi) *
                                         np.linalg.norm(lo)) + 2 * u * np.linalg.norm(C))
        assert np.linalg.norm(Xg - want) <= bound, (n, np.linalg.norm(Xg - want), bound)
        root = S.reference_root(M)
        one = (hi @ hi.T)[:n, :n]
        e1, e2 = S.relative_error(one, root), S.relative_error(Xg, root)
        say(n, "X0 error one term", e1, "two terms", e2)
        assert e2 < 0.5 * e1
