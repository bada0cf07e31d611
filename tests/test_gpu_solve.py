"""End-to-end GPU-resident solves against the FP64 LAPACK oracle."""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, say
from msq import capi
from msq import generators as G
from msq import solver as MS
from oracle import solver as S

pytestmark = pytest.mark.gpu


def solve(runtime, mats, options):
    o, aflat = capi.pack_batch(mats)
    _, _, _, total, _ = capi.batch_layout(o)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    reps = runtime.solve_batched(o, dev(aflat), X, options)
    return capi.unpack_batch(o, host(X)), reps


def test_config0_n64_fp32_one_correction(runtime):
    """n = 64, cond 1e2, FP32: Jacobi to off < 1e-7 ||A||_F (or the FP32
    floor), X0, one correction, compared with the FP64 LAPACK root."""
    A, Xc, lam = G.single_matrix(0, 64, 1e-2)
    Xref = S.reference_root(A)
    opts = capi.solve_options(capi.FP32, capi.FP32, capi.jacobi_options(off_tol=1e-7),
                              min_corrections=1, max_corrections=1, residual_tolerance=0.0)
    (X,), (rep,) = solve(runtime, [A], opts)
    say("config0", rep)
    assert rep["corrections"] == 1
    err = S.relative_error(X, Xref)
    assert err <= 1e-6
    assert rep["residuals"][1] < rep["residuals"][0]
    assert S.relative_symmetry_error(X) <= 1e-12


@pytest.mark.parametrize("tier", ["plain_fp32", "bf16"])
def test_batched_uneven_every_matrix(runtime, tier):
    mats = G.batched_uneven(count=48)
    As = [m["A"] for m in mats]
    opts = MS.plain_fp32_options() if tier == "plain_fp32" else MS.tier_options("bf16")
    Xs, reps = solve(runtime, As, opts)
    worst = 0.0
    for m, X, rep in zip(mats, Xs, reps):
        Xref = S.reference_root(m["A"])
        err = S.relative_error(X, Xref)
        worst = max(worst, err)
        assert rep["status"] == "converged", (m["order"], m["condition"], rep)
        assert err <= 1e-6, (m["order"], m["condition"], err, rep)
    say(tier, "worst", worst)


@pytest.mark.parametrize("tier, tolerance", [("bf16", 5e-3), ("fp8", 0.15), ("mxfp4", 0.5)])
def test_basis_acceptance_and_escalation(runtime, tier, tolerance):
    """The reference's basis acceptance (src/metal_solver.cpp:217-236): V is
    rounded to the basis tier and kept when ||V_t^T V_t - I||_F / sqrt(n) is
    within the tier's tolerance (src/configuration.cpp:109-118), else X0 comes
    from the FP32 basis.  The reported orthogonality equals msq_orthogonality
    of the same Jacobi V quantised and decoded by the codec API, the
    escalation flag follows the tolerance, and every root meets 1e-6."""
    orders = [32, 48, 64, 100, 128, 160, 200, 256]
    mats = [G.single_matrix(14, n, 1e-3, index=i)[0] for i, n in enumerate(orders)]
    opts = MS.tier_options(tier)
    o, aflat = capi.pack_batch(mats)
    _, _, _, total, vtotal = capi.batch_layout(o)
    A = dev(aflat)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    say("basis acceptance", tier, orders)
    reps = runtime.solve_batched(o, A, X, opts)
    # The same Jacobi (same options, deterministic), quantised by the codec API.
    V = torch.zeros(total, dtype=torch.float32, device="cuda")
    D = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    Rt = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    runtime.jacobi_batched(o, A, V, D, Rt, opts.jacobi)
    basis = MS.TIER_PRECISIONS[tier][0]
    decoded = torch.zeros_like(V)
    if basis == capi.FP8_E4M3FN_SCALED:
        codes = torch.zeros(total, dtype=torch.uint8, device="cuda")
        # Per-matrix scale sigma = 448 / amax, applied per segment.
        lds, offs, _, _, _ = capi.batch_layout(o)
        for off, ld in zip(offs, lds):
            seg = V[off:off + ld * ld]
            sig = torch.zeros(1, dtype=torch.float32, device="cuda")
            runtime.quantize_fp8_scaled(seg, codes[off:off + ld * ld], sig)
            runtime.decode_fp8_scaled(codes[off:off + ld * ld], sig, decoded[off:off + ld * ld])
    else:
        codes = torch.zeros(capi.storage_bytes(basis, total), dtype=torch.uint8, device="cuda")
        scales = torch.zeros(max(1, capi.scale_bytes(basis, total)), dtype=torch.uint8,
                             device="cuda")
        runtime.quantize_and_decode(basis, V, codes, scales, decoded)
    orth = runtime.orthogonality(o, decoded)
    got = np.array([r["basis_orthogonality"] for r in reps])
    esc = np.array([r["precision_escalations"] for r in reps])
    say("orthogonality", np.round(orth, 5).tolist(), "escalated", esc.tolist())
    np.testing.assert_allclose(got, orth, rtol=1e-3, atol=1e-6)
    assert (esc == (orth > tolerance)).all()
    # Roots to 1e-6: directly, or (an MXFP4 basis, whose X0 can leave the
    # BF16 corrections diverging) after the plain-path re-solve.
    reps, fallback = MS.solve_with_fallback(runtime, o, A, X, opts)
    say("fallbacks", fallback)
    if tier != "mxfp4":
        assert not fallback
    Xs = capi.unpack_batch(o, host(X))
    for Ai, Xi, rep in zip(mats, Xs, reps):
        assert rep["status"] == "converged", rep
        assert S.relative_error(Xi, S.reference_root(Ai)) <= 1e-6


def test_config1_n1024_bf16(runtime):
    """n = 1024, cond 1e4: BF16 reconstruction, FP64 residual, 2 corrections."""
    A, Xc, lam = G.single_matrix(1, 1024, 1e-4)
    say("config1 n=1024")
    opts = capi.solve_options(capi.BF16, capi.FP32, capi.jacobi_options(off_tol=1e-7),
                              min_corrections=2, max_corrections=2, residual_tolerance=0.0)
    (X,), (rep,) = solve(runtime, [A], opts)
    err = S.relative_error(X, Xc)
    say("config1", rep["sweeps"], rep["residuals"], err)
    assert err <= 1e-6


def test_invalid_spectrum_reported(runtime):
    A = G.single_matrix(2, 32, 1e-2)[0]
    A = A - 2.0 * np.eye(32)   # indefinite
    (X,), (rep,) = solve(runtime, [A], MS.plain_fp32_options())
    assert rep["status"] == "invalid_spectrum"


@pytest.mark.parametrize("tier", ["fp8", "mxfp4"])
def test_config2_n4096_low_bit_products(runtime, tier):
    """n = 4096, lambda log-uniform in [1e-3, 1]: blocked MFMA Jacobi, then the
    reconstruction and 3 corrections with every product in the tier
    (per-tensor-scaled FP8 E4M3 / MXFP4 with E8M0 block scales).  Native
    low-bit corrections meet the reference's verification tolerance of the
    storage width (src/configuration.cpp:98-107: 1e-2 for 8 bits, 0.25 for 4
    bits); FP8's shrink the FP64 residual at every step, 4-bit ones stay at
    their noise floor.  With BF16 corrections on the same low-bit basis the root
    meets 1e-6 against the construction root."""
    A, Xc, lam = G.single_matrix(2, 4096, 1e-3)
    p = capi.FP8_E4M3FN_SCALED if tier == "fp8" else capi.MXFP4_E2M1
    say("config2", tier)
    opts = capi.solve_options(p, p, capi.jacobi_options(off_tol=1e-7), min_corrections=3,
                              max_corrections=3, residual_tolerance=0.0)
    (X,), (rep,) = solve(runtime, [A], opts)
    res = rep["residuals"][:rep["corrections"] + 1]
    err = S.relative_error(X, Xc)
    say("config2", tier, "residuals", res, "error", err)
    assert rep["corrections"] == 3
    assert err <= (1e-2 if tier == "fp8" else 0.25)
    if tier == "fp8":
        assert all(b < a for a, b in zip(res, res[1:]))
    else:
        # 4-bit corrections work at their own noise floor, which the two-term
        # reconstruction already starts below: they stay within it.
        assert max(res) <= 0.5
    # BF16 corrections on the low-bit basis, with the plain-path fallback
    # available: both converge on their own (the MXFP4 reconstruction is the
    # two-term expansion hi + lo of G, about 0.03 off the root, where one
    # E2M1 term alone leaves 0.17 and the corrections stall).
    opts = capi.solve_options(p, capi.BF16, capi.jacobi_options(off_tol=1e-7),
                              max_corrections=8, residual_tolerance=0.0,
                              forward_tolerance=MS.ESTIMATE_TARGET)
    o, aflat = capi.pack_batch([A])
    _, _, _, total, _ = capi.batch_layout(o)
    Xd = torch.zeros(total, dtype=torch.float64, device="cuda")
    (rep,), fallback = MS.solve_with_fallback(runtime, o, dev(aflat), Xd, opts)
    err = S.relative_error(capi.unpack_batch(o, host(Xd))[0], Xc)
    say("config2", tier, "+ bf16 corrections", rep["corrections"], "fallback", fallback,
        "error", err)
    assert rep["status"] == "converged" and err <= 1e-6
    assert not fallback


def _edge_batch():
    """Matrices whose roots are known exactly or nearly so, at orders that
    hit the schedule's and the layout's edges: 1 x 1, 2 x 2, odd orders,
    orders just below and at a multiple of 32, the largest order of the LDS
    kernel and the first of the blocked one."""
    rng = np.random.default_rng(5)
    mats = [np.array([[9.0]]),                                   # root 3
            np.array([[2.0, 1.0], [1.0, 2.0]]),                  # eigenvalues 1, 3
            4.0 * np.eye(3),                                     # root 2 I
            np.eye(17),                                          # root I
            np.diag(np.geomspace(1e-3, 1.0, 31)),                # already diagonal
            G.single_matrix(16, 33, 1e-2)[0]]
    # Exactly repeated eigenvalues: Q diag(1 x 100, 0.01 x 155) Q^T.
    Q = G.haar_orthogonal([G.matrix_seed(16, 1)], 255)[0]
    lam = np.r_[np.ones(100), np.full(155, 1e-2)]
    mats.append((Q * lam) @ Q.T)
    mats.append(G.single_matrix(16, 256, 1e-3, index=2)[0])
    mats.append(G.single_matrix(16, 257, 1e-3, index=3)[0])
    return [0.5 * (M + M.T) for M in mats]


@pytest.mark.parametrize("tier", ["plain", "bf16"])
def test_edge_orders_and_structures(runtime, tier):
    """Orders 1, 2, 3, 17, 31, 33, 255, 256, 257 in one batch (the LDS kernel
    and the blocked one together), identity, scaled identity, diagonal and
    exactly repeated spectra: every root meets 1e-6 against the FP64 LAPACK
    root and is exactly symmetric, and the identity's root is the identity."""
    mats = _edge_batch()
    opts = MS.plain_fp32_options() if tier == "plain" else MS.tier_options("bf16")
    say("edge batch", tier, [M.shape[0] for M in mats])
    o, aflat = capi.pack_batch(mats)
    _, _, _, total, _ = capi.batch_layout(o)
    Ad = dev(aflat)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    reps, fallback = MS.solve_with_fallback(runtime, o, Ad, X, opts)
    got = capi.unpack_batch(o, host(X))
    for M, Xg, rep in zip(mats, got, reps):
        err = S.relative_error(Xg, S.reference_root(M))
        say(M.shape[0], rep["status"], rep["sweeps"], rep["corrections"], err)
        assert rep["status"] == "converged" and err <= 1e-6, (M.shape[0], rep, err)
        assert np.abs(Xg - Xg.T).max() == 0.0
    # The identity: V = I and every root 1, so every product is exact.
    assert np.array_equal(got[3], np.eye(17))
    assert not fallback


def test_invalid_batches_are_errors_and_leave_the_runtime_usable(runtime):
    """Empty batches, zero orders and null buffers are rejected before any
    device work; the runtime then solves the next batch normally."""
    A = torch.zeros(1024, dtype=torch.float64, device="cuda")
    X = torch.zeros(1024, dtype=torch.float64, device="cuda")
    opts = MS.plain_fp32_options()
    with pytest.raises(capi.MsqError, match="empty batch"):
        runtime.solve_batched([], A, X, opts)
    with pytest.raises(capi.MsqError, match="outside"):
        runtime.solve_batched([0], A, X, opts)
    with pytest.raises(capi.MsqError, match="null"):
        runtime.solve_batched([16], 0, X, opts)
    (Xg,), (rep,) = solve(runtime, [np.diag([4.0, 9.0])], opts)
    assert rep["status"] == "converged" and np.allclose(Xg, np.diag([2.0, 3.0]), atol=1e-12)


def test_graph_replays_match_eager(runtime):
    """Graph mode: the first call of a signature runs eagerly, the second is
    captured, later ones replay.  Every replay -- with eager work, host
    synchronisation and new input contents (same buffers) in between -- must
    give the eager result bit for bit."""
    mats = G.batched_uneven(count=40, config=21)
    o, a1 = capi.pack_batch([m["A"] for m in mats])
    a2 = a1 * 1.5  # the same layout with other values
    _, _, _, total, _ = capi.batch_layout(o)
    opts = MS.tier_options("bf16")
    eager = []
    for a in (a1, a2):
        A = dev(a)
        X = torch.zeros(total, dtype=torch.float64, device="cuda")
        reps = runtime.solve_batched(o, A, X, opts)
        eager.append((host(X).copy(), [r["corrections"] for r in reps]))
    A = dev(a1)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    runtime.set_graphs(True)
    try:
        for call, a in enumerate([a1, a1, a1, a2, a1, a2]):
            A.copy_(torch.from_numpy(a))
            X.zero_()
            reps = runtime.solve_batched(o, A, X, opts)
            k = 0 if a is a1 else 1
            got = host(X)
            say("graph call", call, "matches eager", np.array_equal(got, eager[k][0]))
            assert np.array_equal(got, eager[k][0]), call
            assert [r["corrections"] for r in reps] == eager[k][1]
            torch.cuda.synchronize()
            # unrelated eager work on the same runtime between replays
            runtime.orthogonality([16], torch.eye(32, device="cuda").reshape(-1))
    finally:
        runtime.set_graphs(False)


def test_graph_tail_steps_run_when_needed(runtime):
    """A captured solve holds the correction steps up to the last one that
    its eager call needed; later steps sit in a second graph that runs only
    while some matrix is still running.  Replays on harder contents of the
    same buffers (more corrections than the eager call needed) and on the
    easy ones again must both give the eager results bit for bit."""
    orders = [32, 48, 64, 96, 17]
    # Diagonal matrices of powers of four: V = I and exact square roots, so
    # the initial root already has a zero residual.
    easy = [np.diag(4.0 ** (np.arange(n) % 5 - 2)) for n in orders]
    hard = [G.single_matrix(0, n, 1e-5, index=k)[0] for k, n in enumerate(orders)]
    o, a1 = capi.pack_batch(easy)
    _, a2 = capi.pack_batch(hard)
    _, _, _, total, _ = capi.batch_layout(o)
    opts = MS.tier_options("bf16")
    eager = []
    for a in (a1, a2):
        X = torch.zeros(total, dtype=torch.float64, device="cuda")
        reps = runtime.solve_batched(o, dev(a), X, opts)
        eager.append((host(X).copy(), [r["corrections"] for r in reps]))
    say("eager corrections", eager[0][1], eager[1][1])
    assert max(eager[1][1]) > max(eager[0][1])
    A = dev(a1)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    runtime.set_graphs(True)
    try:
        for call, a in enumerate([a1, a1, a2, a1, a2, a2]):
            A.copy_(torch.from_numpy(a))
            X.zero_()
            reps = runtime.solve_batched(o, A, X, opts)
            k = 0 if a is a1 else 1
            got = host(X)
            say("graph call", call, "matches eager", np.array_equal(got, eager[k][0]))
            assert np.array_equal(got, eager[k][0]), call
            assert [r["corrections"] for r in reps] == eager[k][1]
    finally:
        runtime.set_graphs(False)


def test_autotuned_tiles_give_identical_roots(runtime):
    """The autotuner's table holds one of the kernel's tile edges per product
    type and order class; a tile edge only partitions the output, so solves
    with the tuned table, all-32 and all-128 tiles match the defaults bit for
    bit."""
    table = runtime.autotune(256)
    say("tile table\n", table)
    assert set(np.unique(table[:, :8])) <= {32, 64, 128}
    assert (table[:, 8] == 0).all()   # no class above 256 requested
    mats = G.batched_uneven(count=40, config=23)
    o, aflat = capi.pack_batch([m["A"] for m in mats])
    _, _, _, total, _ = capi.batch_layout(o)
    A = dev(aflat)
    results = []
    try:
        for t in (None, table, np.full_like(table, 32), np.full_like(table, 128)):
            runtime.set_tile_table(t)
            X = torch.zeros(total, dtype=torch.float64, device="cuda")
            runtime.solve_batched(o, A, X, MS.tier_options("bf16"))
            results.append(host(X).copy())
    finally:
        runtime.set_tile_table(None)
    for r in results[1:]:
        assert np.array_equal(r, results[0])
    with pytest.raises(capi.MsqError):
        runtime.set_tile_table(np.full_like(table, 48))
