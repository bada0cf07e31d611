"""The C++ mirror of the solver entry points (csrc/hip_runtime.*) and its
command-line driver: host-side checks (no GPU) and GPU runs of the
reference's controlled problems through it."""
import pathlib
import subprocess

import numpy as np
import pytest

from oracle import solver as S

ROOT = pathlib.Path(__file__).resolve().parent.parent
CSRC = ROOT / "mixed-mlx_amd" / "csrc"
MSQ = ROOT / "mixed-mlx_amd" / "msq"
CLI = MSQ / "msq_cli"


def _parse(text):
    """printResult output (space-separated key=value tokens; a key repeated on
    the history lines keeps its last value), cases separated by blank lines."""
    cases, cur = [], {}
    for line in text.splitlines():
        if not line.strip():
            if cur:
                cases.append(cur)
                cur = {}
            continue
        for tok in line.split():
            k, _, v = tok.partition("=")
            if k == "residual":
                cur.setdefault("residuals", []).append(float(v))
            cur[k] = v
    if cur:
        cases.append(cur)
    return cases


def _compile(tmp_path_factory, name, source):
    out = tmp_path_factory.mktemp("cxx") / name
    cmd = ["g++", "-O2", "-std=c++17", "-I", str(CSRC), str(ROOT / "tests" / source),
           "-o", str(out), "-L", str(MSQ), "-lmsq_mirror", "-lmsq_hip",
           f"-Wl,-rpath,{MSQ}", "-Wl,-rpath-link,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    return out


@pytest.fixture(scope="module")
def generator_binary(tmp_path_factory):
    return _compile(tmp_path_factory, "gen", "cxx_generator_check.cpp")


@pytest.mark.parametrize("n,cond", [(1, 5.0), (7, 10.0), (64, 100.0)])
def test_controlled_matrix_matches_oracle(generator_binary, n, cond):
    out = subprocess.run([str(generator_binary), str(n), str(cond)], check=True,
                         capture_output=True, text=True).stdout.split()
    vals = np.array([float(v) for v in out])
    A = vals[:n * n].reshape(n, n, order="F")
    X = vals[n * n:].reshape(n, n, order="F")
    # Same construction; numpy's sin/cos and summation order differ from the
    # C library's by ulps, so the comparison is normwise.
    ref = S.controlled_matrix(n, cond)
    assert np.linalg.norm(A - ref) <= 1e-13 * n * np.linalg.norm(ref)
    # the construction root squares to A
    np.testing.assert_allclose(X @ X, A, rtol=1e-10, atol=1e-10 * cond)


def test_cli_help_and_argument_errors():
    """The reference's option checks (src/application.cpp:79-260) on the HIP
    driver: each rejected combination fails with its message, before any GPU
    work."""
    assert CLI.exists(), "build the library first (mixed-mlx_amd/build.py)"
    r = subprocess.run([str(CLI), "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--paper-test-set" in r.stdout and "--native-stopping" in r.stdout
    bad = [
        (["--solver-backend", "metal"], "hip backend only"),
        (["--orders", "x"], "--orders"),
        (["--precision", "fp7"], "unknown precision"),
        (["--paper-test-set", "I"], "nonsymmetric paper test sets"),
        (["--paper-test-set", "VIII"], "unsupported paper test set"),
        (["--paper-test-set", "V", "--precision", "bf16"], "FP32 mixed correction mode"),
        (["--rounding-seed", "5"], "requires --rounding stochastic"),
        (["--rounding", "stochastic", "--correction-mode", "native_quantized"],
         "does not support native quantized"),
        (["--adaptive", "--correction-mode", "native_quantized"], "--adaptive does not support"),
        (["--adaptive", "--sweep"], "does not support sweep mode"),
        (["--kernel-path", "fast"], "unsupported kernel path"),
        (["--native-stopping", "never"], "unsupported native stopping policy"),
        (["--native-dispatch", "batched"], "unsupported native dispatch policy"),
        (["--native-stagnation-window", "0"], "positive integer"),
        (["--evaluations", "0"], "positive integer"),
        (["64", "128"], "at most one matrix order"),
    ]
    for args, message in bad:
        r = subprocess.run([str(CLI), *args], capture_output=True, text=True)
        assert r.returncode == 1 and message in r.stderr, (args, r.stderr)


@pytest.mark.gpu
def test_cli_controlled_problems_on_gpu(tmp_path):
    """A single enforced case (the reference's runMetal defaults: condition 4,
    8 corrections, residual 1e-12, forward error <= 1e-8 for FP32), then a
    BF16-product case at n = 1024 and a timed sweep that writes summary.csv /
    iterations.csv with the paired reference timing."""
    r = subprocess.run([str(CLI), "--solver-backend", "hip", "64"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    (c,) = _parse(r.stdout)
    assert c["termination_reason"] == "converged" and c["solver_backend"] == "hip"
    assert float(c["final_forward_error"]) <= 1e-8 and float(c["final_residual"]) <= 1e-12
    assert float(c["symmetry_error"]) <= 1e-12
    assert c["certificate_usable"] == "true" and float(c["frozen_contraction_bound"]) < 1.0
    assert c["selected_basis_stage"] == "fp32" and c["case"].startswith("spd_n64_k4_fp32")
    r = subprocess.run([str(CLI), "1024", "--condition", "1e4", "--precision", "bf16",
                        "--correction-precision", "bf16", "--correction-rate", "0.5",
                        "--forward-tolerance", "5e-7", "--tolerance", "0"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    (c,) = _parse(r.stdout)
    assert c["termination_reason"] == "converged" and float(c["final_forward_error"]) <= 1e-6
    assert c["selected_basis_stage"] == "bf16" and c["correction_stage_history"].startswith("bf16_frozen")
    r = subprocess.run([str(CLI), "--sweep", str(tmp_path), "--orders", "64,200",
                        "--conditions", "1e2", "--warmups", "1", "--evaluations", "3"],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "benchmark_failures=0" in r.stdout
    import csv
    import json
    golden = json.loads((ROOT / "tests" / "golden" / "reporting_columns.json").read_text())
    with open(tmp_path / "summary.csv", newline="") as f:
        summary = list(csv.reader(f))
    with open(tmp_path / "iterations.csv", newline="") as f:
        iterations = list(csv.reader(f))
    assert summary[0] == golden["summary"] and iterations[0] == golden["iterations"]
    assert len(summary) == 3
    for row in summary[1:]:
        row = dict(zip(summary[0], row))
        assert row["status"] == "converged" and row["evaluations"] == "3" and row["warmups"] == "1"
        # the paired protocol: reference timings and ratios measured
        assert float(row["reference_timing_mean_seconds"]) > 0.0
        assert float(row["paired_ratio_median"]) > 0.0
        assert float(row["paired_ratio_min"]) <= float(row["paired_ratio_median"]) <= float(row["paired_ratio_max"])
        assert float(row["timing_min_seconds"]) <= float(row["timing_median_seconds"]) <= float(row["timing_max_seconds"])
        assert float(row["final_forward_error"]) <= 1e-8
    assert len(iterations) - 1 == sum(int(dict(zip(summary[0], r))["iterations"]) + 1
                                      for r in summary[1:])


@pytest.mark.gpu
@pytest.mark.parametrize("paper,order", [("V", 512), ("VI", 512), ("VII", 256)])
def test_cli_paper_test_sets_on_gpu(paper, order):
    """Paper sets V / VI / VII (src/accelerate_backend.cpp:690-801) through
    runHipCase with the reference's enforce contract: converged to residual
    1e-12 within 12 corrections, forward error <= 1e-8, symmetric."""
    r = subprocess.run([str(CLI), "--paper-test-set", paper, "--paper-order", str(order)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    (c,) = _parse(r.stdout)
    assert c["paper_test_set"] == paper and c["case"] == f"paper_{paper}_hip_fp32"
    assert c["termination_reason"] == "converged"
    assert float(c["final_forward_error"]) <= 1e-8


@pytest.mark.gpu
def test_cli_adaptive_mode_on_gpu():
    """runAdaptiveHipCase: FP32 factors on the controlled problem, selection
    policy latency_first, enforced forward error <= 1e-9."""
    r = subprocess.run([str(CLI), "--adaptive", "--precision", "mxfp4_e2m1", "128"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    (c,) = _parse(r.stdout)
    assert c["selection_policy"] == "latency_first" and c["execution_model"].endswith("_adaptive")
    assert c["case"].startswith("spd_n128_k4_fp32")
    assert float(c["forward_error"]) <= 1e-9


@pytest.fixture(scope="module")
def csv_binary(tmp_path_factory):
    return _compile(tmp_path_factory, "csv", "cxx_csv_check.cpp")


def test_csv_columns_match_reference_schema(csv_binary):
    """summary.csv and iterations.csv carry the reference's columns in its
    order (fixture transcribed from src/reporting.cpp:124-166, :305-328), and
    every row has one value per column: one summary row per case, one
    iterations row per correction step."""
    import csv
    import io
    import json
    golden = json.loads((ROOT / "tests" / "golden" / "reporting_columns.json").read_text())
    text = subprocess.run([str(csv_binary)], check=True, capture_output=True, text=True).stdout
    summary_text, iterations_text = text.split("--\n")
    summary = list(csv.reader(io.StringIO(summary_text)))
    iterations = list(csv.reader(io.StringIO(iterations_text)))
    assert summary[0] == golden["summary"]
    assert iterations[0] == golden["iterations"]
    assert len(summary) == 2 and len(iterations) == 4
    assert all(len(row) == len(golden["summary"]) for row in summary)
    assert all(len(row) == len(golden["iterations"]) for row in iterations)
    row = dict(zip(summary[0], summary[1]))
    assert row["order"] == "64" and row["status"] == "converged" and row["converged"] == "1"
    assert row["basis_precision"] == "bf16" and row["warmups"] == "1" and row["evaluations"] == "3"
    assert float(row["timing_median_seconds"]) == 0.003
    assert float(row["timing_min_seconds"]) == 0.002 and float(row["timing_max_seconds"]) == 0.004
    assert float(row["reference_timing_mean_seconds"]) == 0.006
    assert float(row["paired_ratio_median"]) == 0.5 and float(row["mixed_to_reference"]) == 0.5
    assert float(row["best_residual"]) == 1e-9 and row["best_iteration"] == "2"
    assert row["iterations"] == "2" and float(row["initial_forward_error"]) == 2e-3
    assert row["certificate_usable"] == "1" and float(row["frozen_contraction_bound"]) == 1.5e-4
    steps = [dict(zip(iterations[0], r)) for r in iterations[1:]]
    assert [s["iteration"] for s in steps] == ["0", "1", "2"]
    assert [float(s["residual"]) for s in steps] == [1e-3, 1e-6, 1e-9]
    assert steps[1]["forward_error"] == "nan" and float(steps[2]["forward_error"]) == 2e-9


@pytest.fixture(scope="module")
def paper_binary(tmp_path_factory):
    return _compile(tmp_path_factory, "paper", "cxx_paper_check.cpp")


@pytest.mark.parametrize("paper,order", [("V", 37), ("VI", 64), ("VII", 50)])
def test_paper_matrix_matches_python_generator(paper_binary, paper, order):
    """preparePaperMatrix of the C++ mirror against msq.generators.paper_matrix
    (the same stream, column order and shift rule): entries equal up to the
    diagonal shift, which comes from two different LAPACK eigensolvers."""
    from msq import generators as G
    out = subprocess.run([str(paper_binary), paper, str(order)], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    n, cond, shift, seed = out[0].split()
    labels = out[1].split()
    vals = np.array([float(v) for v in out[2:] if v])
    A = vals.reshape(int(n), int(n), order="F")
    ref = G.paper_matrix(paper, order)
    assert labels == [paper, ref["family"], "symmetric", ref["shift_rule"]]
    assert int(seed) == ref["seed"]
    off = ~np.eye(order, dtype=bool)
    if paper == "VI":
        # Box-Muller through numpy's log / sin / cos and the C library's:
        # equal to a few ulps
        np.testing.assert_allclose(A[off], ref["A"][off], rtol=1e-14, atol=1e-15)
    else:
        assert np.array_equal(A[off], ref["A"][off])  # drawn values: bit for bit
    assert np.array_equal(A, A.T)
    assert float(shift) == pytest.approx(ref["shift"], rel=1e-12, abs=1e-12)
    assert float(shift) > 0.0
    assert float(cond) == pytest.approx(ref["condition"], rel=1e-9)
    lam = np.linalg.eigvalsh(A)
    if paper == "VII":
        i = np.arange(order)
        np.testing.assert_array_equal(A, 1.0 / (i[:, None] + i[None, :] + 1.0) + 1e-6 * np.eye(order))
    else:
        # the shift puts the smallest eigenvalue at max(1, 0.01 ||A0||_inf)
        A0 = A - float(shift) * np.eye(order)
        sep = max(1.0, 0.01 * np.abs(A0).sum(axis=1).max())
        assert lam[0] == pytest.approx(sep, rel=1e-10)


def test_paper_stream_first_draws_match_the_reference_generator():
    """The first uniforms of the paper-V stream from the scalar generator equal
    the vectorised xorshift64* streams of msq.generators (one seed), and the
    first column of set V is those draws."""
    from msq import generators as G
    seed = 0x243F6A8885A308D3 ^ ((0x9E3779B97F4A7C15 * 5) & (2**64 - 1))
    st = G.Streams([seed])
    draws = [float(st.uniform()[0]) for _ in range(10)]
    A = G.paper_matrix("V", 10)["A"]
    shift = G.paper_matrix("V", 10)["shift"]
    col = A[:, 0].copy()
    col[0] -= shift
    np.testing.assert_allclose(col, draws, rtol=0, atol=1e-15)


def test_timing_and_paired_ratio_summaries(paper_binary):
    """summarizeTimingSamples / summarizePairedRatios (src/benchmark.cpp:17-97):
    a hand-worked case and numpy's statistics on random samples."""
    out = subprocess.run([str(paper_binary), "summary", "3,1,2,4", "2,2,2,2"], check=True,
                         capture_output=True, text=True).stdout.split("\n")
    w, e, lo, hi, med, mean, var = out[0].split()
    assert (w, e) == ("3", "4")
    # sorted 1,2,3,4: median sorted[2] = 3 (upper middle), variance with n - 1
    assert [float(x) for x in (lo, hi, med, mean, var)] == [1.0, 4.0, 3.0, 2.5, 5.0 / 3.0]
    pm, pv, pmin, pmax, pmed = (float(x) for x in out[1].split())
    ratios = [float(x) for x in out[2:6]]
    assert ratios == [1.5, 0.5, 1.0, 2.0]  # evaluation order kept
    assert (pm, pmin, pmax, pmed) == (1.25, 0.5, 2.0, 1.25)  # even count: mean of middle two
    assert pv == pytest.approx(np.var(ratios, ddof=1), rel=1e-15)
    rng = np.random.default_rng(3)
    m, r = rng.uniform(0.1, 2.0, 9), rng.uniform(0.1, 2.0, 9)
    out = subprocess.run([str(paper_binary), "summary", ",".join(repr(float(x)) for x in m),
                          ",".join(repr(float(x)) for x in r)], check=True, capture_output=True,
                         text=True).stdout.split("\n")
    _, _, lo, hi, med, mean, var = (float(x) for x in out[0].split())
    assert (lo, hi, med) == (m.min(), m.max(), np.sort(m)[4])
    assert mean == pytest.approx(m.mean(), rel=1e-14) and var == pytest.approx(m.var(ddof=1), rel=1e-12)
    pm, pv, pmin, pmax, pmed = (float(x) for x in out[1].split())
    q = m / r
    assert pmed == np.median(q) and (pmin, pmax) == (q.min(), q.max())
    assert pm == pytest.approx(q.mean(), rel=1e-14) and pv == pytest.approx(q.var(ddof=1), rel=1e-12)
