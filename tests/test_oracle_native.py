"""Pin of the native staged correction's host emulation (oracle/native.py).

With the identity basis and unit roots every stage but the Sylvester
division maps its stored input to itself (storing a stored value is the
identity, MX scales included), so the correction is E = store(store(R) / 2):
a closed form that does not go through the emulation's product chain."""
import numpy as np
import pytest

from oracle import codecs as C
from oracle import native as ON

FORMATS = ["fp8_e4m3fn", "fp8_e5m2", "mxfp8_e4m3fn", "mxfp8_e5m2", "fp4_e2m1", "mxfp4_e2m1"]


def store(fmt, M):
    x = np.asarray(M, np.float32).reshape(-1, order="F")
    codes, scales = C.encode(fmt, x)
    return C.decode(fmt, codes, scales, x.size).reshape(M.shape, order="F")


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("n", [4, 7])
def test_identity_basis_closed_form(fmt, n):
    rng = np.random.default_rng(n)
    R = rng.standard_normal((n, n)) * 0.7
    E = ON.staged_correction(fmt, np.eye(n), np.ones(n), R)
    want = store(fmt, store(fmt, R).astype(np.float64) / 2.0)
    np.testing.assert_array_equal(E.astype(np.float32), want.astype(np.float32))
