"""ctypes wrapper of oracle/libcpu_baseline.so, the C++ restatement of the
reference's CPU plain path (1_baseline/spd.cpp:348-430) run one matrix per
host thread (test infrastructure: bench.py's cpu_baseline leg and tests/)."""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

LIB = pathlib.Path(__file__).resolve().parent / "libcpu_baseline.so"
_lib = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} is not built (make -C oracle)")
        lib = ctypes.CDLL(str(LIB))
        lib.cpu_accelerate_mixed_batch.restype = ctypes.c_double
        lib.cpu_accelerate_mixed_batch.argtypes = [
            ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
            ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p]
        lib.cpu_blas_config.restype = ctypes.c_char_p
        _lib = lib
    return _lib


def host_threads(limit: int = 16) -> int:
    """Threads for one matrix per core: the CPUs this process may run on,
    capped at `limit` (a one-GPU box's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(limit, n))


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def blas_config() -> str:
    return _load().cpu_blas_config().decode()


class Batch:
    """Matrices packed dense and column-major, one after another."""

    def __init__(self, mats):
        self.orders = np.array([m.shape[0] for m in mats], dtype=np.uint32)
        sizes = self.orders.astype(np.uint64) ** 2
        self.offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.uint64)
        self.A = np.concatenate([np.asarray(m, np.float64).reshape(-1, order="F") for m in mats])
        self.X = np.zeros_like(self.A)

    def root(self, i):
        n = int(self.orders[i])
        o = int(self.offsets[i])
        return self.X[o:o + n * n].reshape(n, n, order="F")


def solve_batch(batch: Batch, threads: int, max_corrections: int = 12, tolerance: float = 1e-12):
    """(seconds, corrections per matrix (-1: failed eigensolve), final
    residual per matrix)."""
    lib = _load()
    count = len(batch.orders)
    corr = np.zeros(count, np.int32)
    res = np.zeros(count, np.float64)
    secs = lib.cpu_accelerate_mixed_batch(
        count, batch.orders.ctypes.data, batch.offsets.ctypes.data, batch.A.ctypes.data,
        batch.X.ctypes.data, max_corrections, tolerance, threads, corr.ctypes.data,
        res.ctypes.data)
    if secs < 0:
        raise RuntimeError("cpu_accelerate_mixed_batch failed")
    return secs, corr, res
