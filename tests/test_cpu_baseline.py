"""The C++ CPU baseline (oracle/cpu_baseline.cpp, the reference's CPU plain
path 1_baseline/spd.cpp:348-430 on host LAPACK) against the numpy restatement
in oracle/solver.py and the construction-known roots, on the CPU."""
import numpy as np
import pytest

from msq import generators as G
from oracle import cpu_baseline as CB
from oracle import solver as S


@pytest.fixture(scope="module")
def sample():
    return G.batched_uneven(count=48)


def test_matches_the_numpy_restatement(sample):
    """Same LAPACK/BLAS library, single-threaded: the same roots, bit for
    bit, and the same number of corrections; only the residual norm's
    accumulation differs (hypot against a sum of squares)."""
    from threadpoolctl import threadpool_limits
    b = CB.Batch([m["A"] for m in sample])
    _, corr, res = CB.solve_batch(b, 1)
    for i in range(0, len(sample), 7):
        with threadpool_limits(limits=1):
            X, steps = S.accelerate_mixed(sample[i]["A"], None)
        assert corr[i] == len(steps) - 1
        assert np.array_equal(b.root(i), X)
        assert res[i] == pytest.approx(steps[-1][0], rel=1e-6, abs=1e-18)


def test_threads_do_not_change_results_and_roots_are_accurate(sample):
    b1 = CB.Batch([m["A"] for m in sample])
    b4 = CB.Batch([m["A"] for m in sample])
    _, c1, r1 = CB.solve_batch(b1, 1)
    _, c4, r4 = CB.solve_batch(b4, 4)
    assert np.array_equal(b1.X, b4.X) and np.array_equal(c1, c4)
    assert (c1 >= 0).all() and (r1 <= 1e-12).all()
    errs = [np.linalg.norm(b1.root(i) - m["X"]) / np.linalg.norm(m["X"])
            for i, m in enumerate(sample)]
    assert max(errs) <= 1e-8


def test_indefinite_matrix_is_reported():
    A = np.diag([1.0, -1.0, 2.0])
    b = CB.Batch([A, np.eye(4) * 4.0])
    _, corr, res = CB.solve_batch(b, 2)
    assert corr[0] == -1 and np.isnan(res[0])
    assert corr[1] == 0 and np.allclose(b.root(1), 2.0 * np.eye(4))


def test_report_fields():
    assert "OpenBLAS" in CB.blas_config()
    assert 1 <= CB.host_threads() <= 16
    assert CB.cpu_model()
