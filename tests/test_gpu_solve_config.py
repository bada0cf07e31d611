"""The solver configuration of the boundary (msq_solve_options, ABI 5) through
the C ABI: basis escalation forced by the tolerance override, spectrum tiers
and their rejection, the frozen-factor certificate and the solve outputs,
unresolved spectra and their second Jacobi pass, the native-quantized
correction (the reference's bitwise 4 x 4 check and a native solve with the
stagnation stop), the n = 1024 BF16 configuration with exactly two
corrections, and the configuration checks of the reference."""
import numpy as np
import pytest
import torch

from gpu_util import dev, host, say
from msq import capi
from msq import generators as G
from oracle import codecs as C
from oracle import native as ON
from oracle import solver as S

pytestmark = pytest.mark.gpu


def batch(specs, config=5):
    """specs: (order, smallest eigenvalue) pairs; log-uniform spectra on
    [lam_min, 1] with Haar bases.  Returns (orders, host A list, roots)."""
    mats = [G.single_matrix(config, n, lam, index=i) for i, (n, lam) in enumerate(specs)]
    return [m[0].shape[0] for m in mats], [m[0] for m in mats], [m[1] for m in mats]


def solve(rt, orders, A_list, options, **outputs):
    o, flat = capi.pack_batch(A_list)
    _, _, _, total, vtotal = capi.batch_layout(o)
    A = dev(flat)
    X = torch.zeros(total, dtype=torch.float64, device="cuda")
    extra = {}
    if outputs.get("initial"):
        extra["initial_roots"] = torch.zeros(total, dtype=torch.float64, device="cuda")
    if outputs.get("factors"):
        extra["vectors"] = torch.zeros(total, dtype=torch.float32, device="cuda")
        extra["root_eigenvalues"] = torch.zeros(vtotal, dtype=torch.float32, device="cuda")
    reps = rt.solve_batched(o, A, X, options, **extra)
    out = {"X": capi.unpack_batch(o, host(X)), "reports": reps}
    if "initial_roots" in extra:
        out["X0"] = capi.unpack_batch(o, host(extra["initial_roots"]))
    if "vectors" in extra:
        out["V"] = capi.unpack_batch(o, host(extra["vectors"]))
        out["s"] = capi.unpack_vectors(o, host(extra["root_eigenvalues"]))
    return out


def errors(Xs, roots):
    return [float(S.relative_error(X, R)) for X, R in zip(Xs, roots)]


def test_forced_basis_escalation(runtime):
    """A basis tolerance below the BF16 rounding's orthogonality rejects the
    tier for half the batch: those matrices get X0 and the final root of the
    FP32-basis solve bit for bit, the others those of the default solve, and
    every one meets 1e-6 (src/metal_solver.cpp:217-236)."""
    orders, A, roots = batch([(40, 1e-2), (96, 1e-3), (150, 1e-2), (200, 1e-4), (64, 1e-3),
                              (256, 1e-2)])
    base = dict(min_corrections=0, max_corrections=12, residual_tolerance=0.0,
                forward_tolerance=0.5e-6,
                jacobi=capi.jacobi_options(off_tol=1e-7, correction_rate=0.5))
    forced = [capi.solve_options(capi.BF16, capi.BF16, basis_tolerance=(1e-9 if i % 2 else 0.0),
                                 **base) for i in range(len(orders))]
    say("escalation: forced")
    r_forced = solve(runtime, orders, A, forced, initial=True)
    r_default = solve(runtime, orders, A, capi.solve_options(capi.BF16, capi.BF16, **base),
                      initial=True)
    r_fp32 = solve(runtime, orders, A, capi.solve_options(capi.FP32, capi.BF16, **base),
                   initial=True)
    for i in range(len(orders)):
        rep = r_forced["reports"][i]
        assert rep["status"] == "converged", (i, rep["status"])
        ref = r_fp32 if i % 2 else r_default
        assert rep["basis_rejected"] == (1 if i % 2 else 0)
        assert rep["precision_escalations"] == (1 if i % 2 else 0)
        assert 0.0 < rep["basis_orthogonality"] < 5e-3
        assert np.array_equal(r_forced["X0"][i], ref["X0"][i]), i
        assert np.array_equal(r_forced["X"][i], ref["X"][i]), i
    assert max(errors(r_forced["X"], roots)) <= 1e-6


def quantize_decode(precision, s):
    """The oracle's storage of one spectrum vector (quantizeAndDecode)."""
    name = capi.PRECISION_NAMES[precision]
    x = np.asarray(s, np.float32)
    if precision in (capi.FP8_E4M3FN_LOG_S, capi.FP4_E2M1_LOG_S):
        e4 = precision == capi.FP8_E4M3FN_LOG_S
        codes, meta = C.encode_log(e4, x)
        return C.decode_log(e4, codes, meta, x.size)
    codes, scales = C.encode(name, x)
    return C.decode(name, codes, scales, x.size)


SPECTRA = [capi.BF16, capi.FP8_E4M3FN, capi.FP8_E5M2, capi.MXFP8_E4M3FN, capi.MXFP8_E5M2,
           capi.FP4_E2M1, capi.MXFP4_E2M1, capi.FP8_E4M3FN_LOG_S, capi.FP4_E2M1_LOG_S]


@pytest.mark.parametrize("precision", SPECTRA)
def test_spectrum_tier(runtime, precision):
    """X0 from the roots stored in the spectrum precision: X0 - X0(FP32 roots)
    = V diag(q - s) V^T, so its Frobenius norm is ||q - s||_2 for the oracle's
    storage q of the solve's own roots s (V orthonormal to float accuracy);
    the FP32 corrections then reach 1e-6."""
    orders, A, roots = batch([(24, 0.3), (64, 0.3), (131, 0.3), (200, 0.3)])
    opts = dict(max_corrections=16, residual_tolerance=0.0, forward_tolerance=0.5e-6)
    say(f"spectrum tier {capi.PRECISION_NAMES[precision]}")
    r = solve(runtime, orders, A, capi.solve_options(spectrum_precision=precision, **opts),
              initial=True, factors=True)
    r32 = solve(runtime, orders, A, capi.solve_options(**opts), initial=True)
    for i, n in enumerate(orders):
        rep = r["reports"][i]
        assert rep["spectrum_rejected"] == 0 and rep["precision_escalations"] == 0
        assert rep["status"] == "converged"
        s = r["s"][i]
        q = quantize_decode(precision, s).astype(np.float64)
        want = np.linalg.norm(q - s.astype(np.float64))
        got = np.linalg.norm(r["X0"][i] - r32["X0"][i])
        assert abs(got - want) <= 1e-5 * np.linalg.norm(r32["X0"][i]) + 1e-6 * want, (i, got, want)
        assert rep["frozen_sylvester_inverse_norm"] == pytest.approx(1.0 / (2.0 * q.min()),
                                                                    rel=1e-6)
    assert max(errors(r["X"], roots)) <= 1e-6


def test_spectrum_rejection(runtime):
    """Roots below 0.25 store as zero in FP4 E2M1: the spectrum tier is
    rejected (src/metal_solver.cpp:240-263) and X0 is the FP32-roots one bit
    for bit."""
    orders, A, roots = batch([(48, 1e-2), (100, 1e-3)])
    opts = dict(max_corrections=16, residual_tolerance=0.0, forward_tolerance=0.5e-6)
    r = solve(runtime, orders, A, capi.solve_options(spectrum_precision=capi.FP4_E2M1, **opts),
              initial=True)
    r32 = solve(runtime, orders, A, capi.solve_options(**opts), initial=True)
    for i in range(len(orders)):
        rep = r["reports"][i]
        assert rep["spectrum_rejected"] == 1 and rep["precision_escalations"] == 1
        assert np.array_equal(r["X0"][i], r32["X0"][i])
        assert np.array_equal(r["X"][i], r32["X"][i])
    assert max(errors(r["X"], roots)) <= 1e-6


def test_certificate_and_outputs(runtime):
    """certificate = 1: ||V^T V - I||_F / sqrt(n) of the FP32 basis, 1 / (2 min s)
    and usability (src/metal_solver.cpp:511-524) against the solve's own V and
    s; the contraction bound from the initial root; a BF16 basis' certificate
    rests on its tier orthogonality."""
    orders, A, roots = batch([(32, 1e-2), (120, 1e-3), (256, 1e-2)])
    r = solve(runtime, orders, A, capi.solve_options(max_corrections=12, residual_tolerance=0.0,
                                                     forward_tolerance=0.5e-6, certificate=1),
              initial=True, factors=True)
    for i, n in enumerate(orders):
        rep = r["reports"][i]
        V = r["V"][i].astype(np.float64)
        orth = np.linalg.norm(V.T @ V - np.eye(n)) / np.sqrt(n)
        # Measured from an FP32 Gram product, as measureF32Orthogonality does:
        # equal to the binary64 value to FP32 accumulation accuracy.
        assert rep["frozen_basis_orthogonality"] == pytest.approx(orth, rel=0.05, abs=1e-7)
        assert rep["certificate_usable"] == 1
        assert rep["frozen_sylvester_inverse_norm"] == pytest.approx(
            1.0 / (2.0 * float(r["s"][i].min())), rel=1e-6)
        # X0 is V diag(s) V^T to FP32 product accuracy.
        X0 = (V * r["s"][i].astype(np.float64)) @ V.T
        assert np.linalg.norm(r["X0"][i] - X0) <= 1e-5 * np.linalg.norm(X0)
        bound = 3.0 * rep["frozen_sylvester_inverse_norm"] * np.linalg.norm(r["X0"][i] - roots[i])
        assert np.isfinite(bound) and bound > 0
    rb = solve(runtime, orders, A, capi.solve_options(capi.BF16, capi.BF16, max_corrections=12,
                                                      residual_tolerance=0.0,
                                                      forward_tolerance=0.5e-6, certificate=1))
    for rep in rb["reports"]:
        assert rep["frozen_basis_orthogonality"] == pytest.approx(rep["basis_orthogonality"],
                                                                 rel=1e-6)
        assert rep["certificate_usable"] == (1 if rep["basis_orthogonality"] < 1e-3 else 0)


def test_unresolved_status_and_second_pass(runtime):
    """Spectra far below the FP32 Jacobi's resolution (condition 1e8): without
    the second pass no matrix reports converged with an error above 1e-6 (it
    reports unresolved); with it, every matrix converges within 1e-6."""
    orders, A, roots = batch([(64, 1e-8), (128, 1e-8), (200, 3e-9), (256, 1e-8)], config=6)
    opts = dict(max_corrections=30, residual_tolerance=0.0, forward_tolerance=0.5e-6,
                resolution=1e-6)
    plain = solve(runtime, orders, A, capi.solve_options(**opts))
    errs = errors(plain["X"], roots)
    for rep, e in zip(plain["reports"], errs):
        say(f"unresolved: status {rep['status']} error {e:.2e}")
        assert rep["unresolved"] == 1
        assert not (rep["status"] == "converged" and e > 1e-6)
    refined = solve(runtime, orders, A, capi.solve_options(refine_unresolved=1, **opts))
    errs = errors(refined["X"], roots)
    for rep, e in zip(refined["reports"], errs):
        say(f"refined: status {rep['status']} error {e:.2e} sweeps {rep['sweeps']}")
        assert rep["refined"] == 1 and rep["unresolved"] == 0
        assert rep["status"] == "converged"
    assert max(errs) <= 1e-6


NATIVE = ["fp8_e4m3fn", "fp8_e5m2", "mxfp8_e4m3fn", "mxfp8_e5m2", "fp4_e2m1", "mxfp4_e2m1"]


@pytest.mark.parametrize("fmt", NATIVE)
def test_native_staged_correction_reference_check(runtime, fmt):
    """src/metal_checks.cpp:603-735: identity basis, unit roots, the fixed 4 x 4
    residual; the decoded correction equals the host emulation bit for bit."""
    n = 4
    R = np.array([2.0, -0.5, 0.25, -0.125, -0.5, 1.75, -0.375, 0.25, 0.25, -0.375, 1.5, -0.25,
                  -0.125, 0.25, -0.25, 1.25]).reshape(n, n, order="F")
    V = np.eye(n, dtype=np.float32)
    roots = np.ones(n, np.float32)
    o, vflat = capi.pack_batch([V], np.float32)
    _, rflat = capi.pack_batch([R.astype(np.float32)], np.float32)
    _, _, voffs, total, vtotal = capi.batch_layout(o)
    sflat = np.zeros(vtotal, np.float32)
    sflat[:n] = roots
    out = torch.zeros(total, dtype=torch.float32, device="cuda")
    runtime.refine_native_staged(capi.PRECISION_BY_NAME[fmt], o, dev(vflat), dev(sflat),
                                 dev(rflat), out)
    got = capi.unpack_batch(o, host(out))[0]
    want = ON.staged_correction(fmt, V, roots, R).astype(np.float32)
    assert np.all(np.isfinite(got))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


@pytest.mark.parametrize("fmt,tolerance", [("fp8_e4m3fn", 1e-2), ("mxfp8_e5m2", 1e-2),
                                           ("mxfp4_e2m1", 0.25)])
def test_native_solve_with_stagnation_stop(runtime, fmt, tolerance):
    """Native-quantized corrections reach the storage format's verification
    tolerance (src/configuration.cpp:98-107: 8-bit 1e-2, 4-bit 0.25); with
    the stagnation policy the best iterate is returned
    (src/metal_solver.cpp:883-921)."""
    orders, A, roots = batch([(16, 0.3), (48, 0.2), (64, 0.3)])
    p = capi.PRECISION_BY_NAME[fmt]
    opts = capi.solve_options(p, p, spectrum_precision=p,
                              correction_mode=capi.CORRECTION_NATIVE_QUANTIZED,
                              kernel_path=capi.KERNEL_PATH_LEGACY,
                              stopping_policy=capi.STOPPING_STAGNATION, stagnation_window=4,
                              minimum_relative_improvement=1e-3, max_corrections=64,
                              residual_tolerance=1e-12, forward_tolerance=0.0)
    say(f"native solve {fmt}")
    r = solve(runtime, orders, A, opts)
    errs = errors(r["X"], roots)
    for rep, e in zip(r["reports"], errs):
        assert rep["status"] in ("stagnated", "budget_exhausted", "converged", "diverged"), \
            rep["status"]
        hist = rep["residuals"]
        assert rep["best_residual"] == pytest.approx(min(hist), rel=0, abs=0)
        assert rep["final_residual"] == rep["best_residual"]
        assert hist[rep["best_iteration"]] == rep["best_residual"]
        assert rep["certificate_usable"] == 0
        assert e <= tolerance, (fmt, e)


def test_configuration_checks(runtime):
    """The reference's configuration checks (src/configuration.cpp:290-346,
    src/metal_solver.cpp:359-370) as MSQ_ERROR_INVALID_ARGUMENT."""
    orders, A, _ = batch([(16, 0.3)])
    bad = [
        dict(correction_mode=capi.CORRECTION_NATIVE_QUANTIZED),  # fp32 storage, tiled path
        dict(stopping_policy=capi.STOPPING_STAGNATION),
        dict(dispatch_policy=capi.DISPATCH_PER_STAGE),
        dict(stagnation_window=0),
        dict(minimum_relative_improvement=-1.0),
        dict(basis=capi.FP8_E4M3FN_LOG_S),
    ]
    for fields in bad:
        basis = fields.pop("basis", capi.FP32)
        with pytest.raises(capi.MsqError):
            solve(runtime, orders, A, capi.solve_options(basis, capi.FP32, **fields))
    p = capi.FP8_E4M3FN
    with pytest.raises(capi.MsqError, match="legacy"):
        solve(runtime, orders, A, capi.solve_options(
            p, p, spectrum_precision=p, correction_mode=capi.CORRECTION_NATIVE_QUANTIZED))
    with pytest.raises(capi.MsqError, match="matching"):
        solve(runtime, orders, A, capi.solve_options(
            p, p, spectrum_precision=capi.FP32, kernel_path=capi.KERNEL_PATH_LEGACY,
            correction_mode=capi.CORRECTION_NATIVE_QUANTIZED))


def test_n1024_bf16_two_corrections(runtime):
    """Haar n = 1024, cond 1e4, BF16 storage and products with FP32
    accumulation, FP64 residual, exactly two corrections: the error per step
    falls and the second correction meets 1e-6."""
    A, X, _ = G.single_matrix(1, 1024, 1e-4)
    o, flat = capi.pack_batch([A])
    _, _, _, total, _ = capi.batch_layout(o)
    Ad = dev(flat)
    Xd = torch.zeros(total, dtype=torch.float64, device="cuda")
    errs = []
    for k in range(3):
        opts = capi.solve_options(capi.BF16, capi.BF16, capi.jacobi_options(off_tol=1e-7),
                                  min_corrections=k, max_corrections=k, residual_tolerance=0.0)
        rep = runtime.solve_batched(o, Ad, Xd, opts)[0]
        assert rep["corrections"] == k
        errs.append(float(S.relative_error(capi.unpack_batch(o, host(Xd))[0], X)))
    say(f"n=1024 bf16 error per step {errs}")
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] <= 1e-6
