"""Pins of the oracle's stochastic-rounding and log-affine spectrum codecs
(oracle/codecs_ref.c) by the reference's own known answers and by properties
the formats fix, on the CPU.

* Stochastic rounding (metal_mixed_sqrt.metal:394-405, :462-483, :545-567,
  :740-762, :1888-1910): the reference's check (src/metal_checks.cpp:409-535)
  -- reproducible for a seed, changed by another seed, sample mean within
  5e-5 (BF16), 1e-2 (4-bit) or 1e-3 (8-bit) of the source; every result is
  one of the two codes around the value; values the nearest mode maps
  exactly, zeros, NaN and the saturating range keep their nearest code; the
  round-up frequency matches the position between the two codes.  The hash
  constants themselves are pinned only through these statistics (parity of
  individual draws with the reference is unpinned).
* Log-affine spectrum (:1689-1817; log2 rounded toward -inf, so that the
  top code's decode exp2(min + span) never exceeds the block maximum, which
  the reference's one-ulp span step aims at): the reference's edge check
  (src/metal_checks.cpp:360-408) and storage check (:537-571, relative error
  below 0.5 for FP8 and 4.0 for FP4), and closed forms on blocks of powers of
  two, where log2 is exact.
"""
import math

import numpy as np
import pytest

from oracle import codecs as C

FORMATS = ["bf16", "fp8_e4m3fn", "fp8_e5m2", "mxfp8_e4m3fn", "mxfp8_e5m2", "fp4_e2m1",
           "mxfp4_e2m1"]
TOLERANCE = {"bf16": 5e-5, "fp4_e2m1": 1e-2, "mxfp4_e2m1": 1e-2}


@pytest.mark.parametrize("fmt", FORMATS)
def test_stochastic_rounding_reference_check(fmt):
    """src/metal_checks.cpp:409-535 on the oracle."""
    n = 65536
    src = np.where(np.arange(n) < n // 2, 0.3, -0.3).astype(np.float32)
    c1, s1 = C.encode_stochastic(fmt, src, 0x53520000, 0)
    c1b, _ = C.encode_stochastic(fmt, src, 0x53520000, 0)
    c2, _ = C.encode_stochastic(fmt, src, 0x53520001, 0)
    assert np.array_equal(c1, c1b)
    assert not np.array_equal(c1, c2)
    d = C.decode(fmt, c1, s1, n).astype(np.float64)
    err = max(abs(d[:n // 2].mean() - 0.3), abs(d[n // 2:].mean() + 0.3))
    assert err <= TOLERANCE.get(fmt, 1e-3)


def _neighbours(fmt, x):
    """The two adjacent decoded values around |x| (scalar formats)."""
    c, s = C.encode(fmt, np.array([x], np.float32))
    near = float(C.decode(fmt, c, s, 1)[0])
    return near


@pytest.mark.parametrize("fmt", ["fp8_e4m3fn", "fp8_e5m2", "fp4_e2m1", "bf16"])
def test_stochastic_results_are_neighbouring_codes(fmt):
    rng = np.random.default_rng(3)
    x = (rng.standard_normal(20000) * rng.choice([1e-3, 1.0, 30.0], 20000)).astype(np.float32)
    cs, _ = C.encode_stochastic(fmt, x, 7, 2)
    ds = C.decode(fmt, cs, None, x.size).astype(np.float64)
    cn, _ = C.encode(fmt, x)
    dn = C.decode(fmt, cn, None, x.size).astype(np.float64)
    # Either the nearest code, or the code on the other side of x (sign kept,
    # never farther than the gap between the two).
    other = ds != dn
    assert np.all(np.signbit(ds[other]) == np.signbit(x[other]))
    lo = np.minimum(ds, dn)
    hi = np.maximum(ds, dn)
    xs = x.astype(np.float64)
    assert np.all((lo[other] <= xs[other]) & (xs[other] <= hi[other]))


@pytest.mark.parametrize("fmt", ["fp8_e4m3fn", "fp8_e5m2", "fp4_e2m1", "bf16"])
def test_stochastic_keeps_exact_and_special_values(fmt):
    # Values every one of these formats represents exactly, and the specials.
    special = np.array([0.0, -0.0, np.nan, np.inf, -np.inf, 1e9, -1e9, 1.0, -0.5, 2.0, 4.0],
                       np.float32)
    cs, _ = C.encode_stochastic(fmt, special, 11, 0)
    cn, _ = C.encode(fmt, special)
    assert np.array_equal(cs, cn)


def test_stochastic_round_up_frequency():
    """0.3 lies 60% of the way from 0.28125 to 0.3125 in E4M3FN."""
    n = 200000
    src = np.full(n, 0.3, np.float32)
    c, _ = C.encode_stochastic("fp8_e4m3fn", src, 1234, 5)
    d = C.decode("fp8_e4m3fn", c, None, n)
    assert set(np.unique(d).tolist()) == {0.28125, 0.3125}
    assert abs(np.mean(d == 0.3125) - 0.6) < 0.005


@pytest.mark.parametrize("e4m3", [True, False])
def test_log_spectrum_edge_check(e4m3):
    """src/metal_checks.cpp:360-408."""
    src = np.array([np.finfo(np.float32).tiny, 2.0 ** -100, 2.0 ** -64, 2.0 ** -32, 1.0,
                    2.0 ** 32, 2.0 ** 64, np.finfo(np.float32).max] + [3.0] * 8, np.float32)
    for stochastic in (False, True):
        codes, meta = C.encode_log(e4m3, src, stochastic, 0x4c4f4753, 0)
        d = C.decode_log(e4m3, codes, meta, src.size)
        assert d[0] > 0 and np.isfinite(d[0])
        assert d[7] > 0 and np.isfinite(d[7])
        assert np.all(np.abs(d[8:16] / 3.0 - 1.0) <= 1e-6)


@pytest.mark.parametrize("e4m3,bound", [(True, 0.5), (False, 4.0)])
def test_log_spectrum_storage_check(e4m3, bound):
    """src/metal_checks.cpp:537-571: relative error of a decoded spectrum."""
    roots = np.array([1.0, 1.5, 2.0, 3.0, 10.0, 31.622776, 100.0, 1000.0], np.float32)
    codes, meta = C.encode_log(e4m3, roots)
    d = C.decode_log(e4m3, codes, meta, roots.size)
    assert np.all(np.isfinite(d) & (d > 0))
    assert np.max(np.abs(d / roots - 1.0)) < bound


def test_log_spectrum_powers_of_two_closed_form():
    """Block of 2^0 .. 2^7: log2 exact, min 0, span 7 (min + span equals the
    maximum, no ulp step).  E4M3: codes of 64 k, decode exactly 2^k.  E2M1:
    codes of 6 k / 7 rounded to the nearest E2M1 value, decode 2^(7 c / 6)."""
    src = (2.0 ** np.arange(8)).astype(np.float32)
    codes, meta = C.encode_log(True, src)
    assert meta.tolist() == [0.0, 7.0]
    assert np.array_equal(C.decode_log(True, codes, meta, 8), src)
    codes, meta = C.encode_log(False, src)
    grid = np.array([0, 0.5, 1, 1.5, 2, 3, 4, 6])
    want_codes = [grid[np.argmin(np.abs(grid - 6 * k / 7))] for k in range(8)]
    got = C.decode_log(False, codes, meta, 8)
    want = np.array([np.float32(2.0 ** (np.float32(7.0) * np.float32(c) / np.float32(6.0)))
                     for c in want_codes], np.float32)
    assert np.array_equal(got, want)


def test_log_spectrum_metadata_rounds_span_down():
    """min + span must not exceed the block maximum (the decode of the top code
    is exp2(min + span)); when the float sum rounds up the span steps down."""
    rng = np.random.default_rng(5)
    for _ in range(2000):
        v = np.exp2(rng.uniform(-20, 20, 8)).astype(np.float32)
        _, meta = C.encode_log(True, v)
        lo, span = np.float32(meta[0]), np.float32(meta[1])
        top = np.float32(max(math.log2(float(x)) for x in v))
        assert np.float32(lo + span) <= top
