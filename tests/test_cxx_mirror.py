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
    cases, cur = [], {}
    for line in text.splitlines():
        if not line.strip():
            if cur:
                cases.append(cur)
                cur = {}
            continue
        k, _, v = line.partition("=")
        cur[k] = v
    if cur:
        cases.append(cur)
    return cases


@pytest.fixture(scope="module")
def generator_binary(tmp_path_factory):
    out = tmp_path_factory.mktemp("cxx") / "gen"
    cmd = ["g++", "-O2", "-std=c++17", "-I", str(CSRC), str(ROOT / "tests" / "cxx_generator_check.cpp"),
           str(CSRC / "hip_runtime.cpp"), "-o", str(out), "-L", str(MSQ), "-lmsq_hip",
           f"-Wl,-rpath,{MSQ}", "-Wl,-rpath-link,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    return out


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
    assert CLI.exists(), "build the library first (mixed-mlx_amd/build.py)"
    r = subprocess.run([str(CLI), "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "--orders" in r.stdout
    r = subprocess.run([str(CLI), "--solver-backend", "metal"], capture_output=True, text=True)
    assert r.returncode == 1 and "hip backend only" in r.stderr
    r = subprocess.run([str(CLI), "--orders", "x"], capture_output=True, text=True)
    assert r.returncode == 1 and "--orders" in r.stderr
    r = subprocess.run([str(CLI), "--precision", "fp7"], capture_output=True, text=True)
    assert r.returncode == 1 and "unknown precision" in r.stderr


@pytest.mark.gpu
def test_cli_controlled_problems_on_gpu(tmp_path):
    """Config 0 (n = 64, cond 100, FP32, one correction) and n = 1024 BF16
    products through runHipCase, errors against the construction root; the
    second run also writes summary.csv / iterations.csv."""
    r = subprocess.run([str(CLI), "--solver-backend", "hip", "--orders", "64",
                        "--conditions", "100", "--precision", "fp32", "--max-iterations", "1",
                        "--enforce"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    (c,) = _parse(r.stdout)
    assert c["status"] == "converged" and c["solver_backend"] == "hip"
    assert float(c["final_forward_error"]) <= 1e-6
    assert float(c["symmetry_error"]) <= 1e-12
    r = subprocess.run([str(CLI), "--orders", "1024", "--conditions", "1e4", "--precision", "bf16",
                        "--correction-rate", "0.5", "--enforce", "--output-dir", str(tmp_path)],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    (c,) = _parse(r.stdout.split("summary_csv=")[0])
    assert c["status"] == "converged" and float(c["final_forward_error"]) <= 1e-6
    import csv
    import json
    golden = json.loads((ROOT / "tests" / "golden" / "reporting_columns.json").read_text())
    with open(tmp_path / "summary.csv", newline="") as f:
        summary = list(csv.reader(f))
    with open(tmp_path / "iterations.csv", newline="") as f:
        iterations = list(csv.reader(f))
    assert summary[0] == golden["summary"] and iterations[0] == golden["iterations"]
    row = dict(zip(summary[0], summary[1]))
    assert row["order"] == "1024" and row["status"] == "converged"
    assert float(row["final_forward_error"]) == pytest.approx(float(c["final_forward_error"]), rel=1e-5)
    assert len(iterations) - 1 == int(row["iterations"]) + 1


@pytest.fixture(scope="module")
def csv_binary(tmp_path_factory):
    out = tmp_path_factory.mktemp("cxx") / "csv"
    cmd = ["g++", "-O2", "-std=c++17", "-I", str(CSRC), str(ROOT / "tests" / "cxx_csv_check.cpp"),
           str(CSRC / "hip_runtime.cpp"), "-o", str(out), "-L", str(MSQ), "-lmsq_hip",
           f"-Wl,-rpath,{MSQ}", "-Wl,-rpath-link,/opt/rocm/lib"]
    subprocess.run(cmd, check=True)
    return out


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
    assert row["basis_precision"] == "bf16" and row["warmups"] == "1"
    assert float(row["timing_median_seconds"]) == 0.003
    assert float(row["timing_min_seconds"]) == 0.002 and float(row["timing_max_seconds"]) == 0.004
    assert float(row["best_residual"]) == 1e-9 and row["best_iteration"] == "2"
    steps = [dict(zip(iterations[0], r)) for r in iterations[1:]]
    assert [s["iteration"] for s in steps] == ["0", "1", "2"]
    assert [float(s["residual"]) for s in steps] == [1e-3, 1e-6, 1e-9]
