"""Host emulation of the reference's native staged correction (test
infrastructure only).

Restates the independent host reference of src/metal_checks.cpp:603-735,
which checks refineNativeStaged (src/metal_runtime.cpp:312-411) bit for bit:
basis, transposed basis, roots and residual stored with the reference codecs
(copyToStorage / copyResidualToStorage on the linear n x n layout), then

    U = Q_t R,  C = U Q,  Y_ij = C_ij / (s_i + s_j),  V2 = Q Y,  E = V2 Q_t,

each product accumulated in float with fused multiply-adds over the inner
index in order and stored in the format (quantize_host), the Sylvester step
in binary64 rounded to float.  Only tests/ load this module.
"""
from __future__ import annotations

import numpy as np

from . import codecs as C


def _fma32(a, b, c):
    """float32 fma: a*b is exact in binary64 for float operands, and the sum
    rounds once to binary64, then to float (exact for the test operands, whose
    products and sums are exactly representable)."""
    return np.float32(float(a) * float(b) + float(c))


def _store(fmt, values):
    """Values (float32, linear column-major) stored in the format and decoded."""
    codes, scales = C.encode(fmt, values)
    return C.decode(fmt, codes, scales, values.size).astype(np.float64)


def _matmul(fmt, left, right, n):
    out = np.zeros(n * n, np.float32)
    for col in range(n):
        for row in range(n):
            acc = np.float32(0.0)
            for k in range(n):
                acc = _fma32(np.float32(left[row + k * n]), np.float32(right[k + col * n]), acc)
            out[row + col * n] = acc
    return _store(fmt, out)


def staged_correction(fmt, V, roots, R):
    """Decoded E of one native staged correction (V, R: n x n; roots: n)."""
    n = V.shape[0]
    lin = lambda M: np.asarray(M, np.float32).reshape(-1, order="F")  # noqa: E731
    q = _store(fmt, lin(V))
    qt = _store(fmt, lin(np.asarray(V).T))
    s = _store(fmt, np.asarray(roots, np.float32))
    r = _store(fmt, lin(np.asarray(R, np.float64).astype(np.float32)))
    u = _matmul(fmt, qt, r, n)
    c = _matmul(fmt, u, q, n)
    y = np.zeros(n * n, np.float32)
    for col in range(n):
        for row in range(n):
            y[row + col * n] = np.float32(c[row + col * n] / (s[row] + s[col]))
    y = _store(fmt, y)
    v = _matmul(fmt, q, y, n)
    e = _matmul(fmt, v, qt, n)
    return e.reshape(n, n, order="F")
