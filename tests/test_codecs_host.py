"""CPU checks of the codec oracle and of the device codec logic.

The oracle (oracle/codecs_ref.c) restates the reference's host codecs
(src/storage_codecs.cpp).  It is pinned here by:
  * the reference's own known answers (checkLowPrecisionCodecs,
    src/storage_codecs.cpp:200-247) and edge tables (src/metal_checks.cpp:195-203,
    :286-292) and boundary values (tests/matrix_codec.cpp:59-76);
  * torch's float8_e4m3fn / float8_e5m2 / bfloat16 casts, an independent
    round-to-nearest-even implementation, on in-range values;
  * exhaustive decode tables against the formats' closed-form definitions;
  * an exact brute force in rational arithmetic for FP4 on every quarter
    grid point.
The device codecs (mixed-mlx_amd/csrc/device_codecs.hpp) are then compiled for
the host and compared with the oracle over millions of float bit patterns.
"""
import fractions
import math
import pathlib
import subprocess

import numpy as np
import pytest
import torch

from oracle import codecs as C

ROOT = pathlib.Path(__file__).resolve().parent.parent
NAN = float("nan")
INF = float("inf")


def test_reference_known_answers():
    # src/storage_codecs.cpp:201-211
    codes, _ = C.encode("fp8_e4m3fn", [448.0, NAN, 1.0625, 1.1875])
    assert codes[0] == 0x7E
    assert np.isnan(C.decode("fp8_e4m3fn", codes[1:2], None, 1)[0])
    assert codes[2] == 0x38 and codes[3] == 0x3A
    e5, _ = C.encode("fp8_e5m2", [INF])
    assert e5[0] == 0x7C and np.isinf(C.decode("fp8_e5m2", e5, None, 1)[0])
    # MXFP4 block {6.1, 0 ...}, 33 elements: scales 127, 127 and code 7 (:213-224)
    v, s = C.encode("mxfp4_e2m1", [6.1] + [0.0] * 32)
    assert list(s) == [127, 127] and (v[0] & 0xF) == 0x7
    # MXFP8 E4M3 block {512, 0 ...}: scale 128, decode 512 (:226-236)
    v, s = C.encode("mxfp8_e4m3fn", [512.0] + [0.0] * 32)
    assert list(s) == [128, 127]
    assert C.decode("mxfp8_e4m3fn", v, s, 33)[0] == 512.0
    # MXFP8 E5M2 all-ones block: scale 112 (:238-246)
    v, s = C.encode("mxfp8_e5m2", [1.0] * 32)
    assert s[0] == 112 and C.decode("mxfp8_e5m2", v, s, 32)[0] == 1.0


@pytest.mark.parametrize("fmt,minimum,maximum,inf_preserved,nan_preserved", [
    ("bf16", np.finfo(np.float32).tiny, math.ldexp(255.0, 120), True, True),
    ("fp8_e4m3fn", 0.001953125, 448.0, False, True),
    ("fp8_e5m2", 0.0000152587890625, 57344.0, True, True),
    ("fp4_e2m1", 0.5, 6.0, False, False),
])
def test_scalar_edge_table(fmt, minimum, maximum, inf_preserved, nan_preserved):
    """checkQuantizationEdgeBehavior scalar table, src/metal_checks.cpp:195-280."""
    fmax = float(np.finfo(np.float32).max)
    src = np.array([0.0, -0.0, minimum, -minimum, 0.25 * minimum, -0.25 * minimum,
                    maximum, -maximum, fmax, -fmax, INF, -INF, NAN], np.float32)
    codes, _ = C.encode(fmt, src)
    d = C.decode(fmt, codes, None, src.size)
    assert d[0] == 0 and not np.signbit(d[0])
    assert d[1] == 0 and np.signbit(d[1])
    assert d[2] == minimum and d[3] == -minimum
    # The Metal table allows 0 or the minimum (its GPU flushes denormals); the
    # reference host codec keeps a BF16 subnormal such as 2^-128 exactly.
    under = (0.0, minimum) if fmt != "bf16" else (0.0, minimum, np.float32(0.25 * minimum))
    assert abs(d[4]) in under and not np.signbit(d[4])
    assert abs(d[5]) in under and np.signbit(d[5])
    assert d[6] == maximum and d[7] == -maximum
    if fmt == "bf16":
        assert d[8] == maximum or np.isinf(d[8])
    else:
        assert d[8] == maximum and d[9] == -maximum
    if inf_preserved:
        assert np.isinf(d[10]) and d[10] > 0 and np.isinf(d[11]) and d[11] < 0
    else:
        assert d[10] == maximum and d[11] == -maximum
    assert np.isnan(d[12]) if nan_preserved else d[12] == 0


@pytest.mark.parametrize("fmt,first_scale,first_min", [
    ("mxfp8_e4m3fn", -8, math.ldexp(1.0, -17)),
    ("mxfp8_e5m2", -15, math.ldexp(1.0, -31)),
    ("mxfp4_e2m1", -2, 0.125),
])
def test_block_edge_table(fmt, first_scale, first_min):
    """checkQuantizationEdgeBehavior block table, src/metal_checks.cpp:282-330."""
    src = np.zeros(64, np.float32)
    src[0] = 1.0
    src[2] = -0.0
    src[3] = first_min
    src[4] = 0.25 * first_min
    src[5] = INF
    src[6] = NAN
    src[32] = 8.0
    src[34] = -0.0
    src[35] = math.ldexp(first_min, 3)
    src[36] = math.ldexp(0.25 * first_min, 3)
    v, s = C.encode(fmt, src)
    d = C.decode(fmt, v, s, 64)
    assert d[0] == 1.0 and d[32] == 8.0
    assert s[0] == first_scale + 127 and s[1] == first_scale + 3 + 127
    assert d[1] == 0 and not np.signbit(d[1])
    assert d[2] == 0 and np.signbit(d[2])
    assert d[3] == first_min
    assert d[35] == math.ldexp(first_min, 3)


def _boundaries():
    # tests/matrix_codec.cpp:59-76
    return np.array([-0.0, 2.0**-25, 3 * 2.0**-25, 2.0**-17, 3 * 2.0**-17, 2.0**-10,
                     3 * 2.0**-10, 448.0, 464.0, -464.0, 57344.0, 61440.0, 65504.0,
                     65520.0], np.float32)


@pytest.mark.parametrize("fmt,tdtype,maximum", [
    ("fp8_e4m3fn", torch.float8_e4m3fn, 448.0),
    ("fp8_e5m2", torch.float8_e5m2, 57344.0),
    ("bf16", torch.bfloat16, None),
])
def test_against_torch_in_range(fmt, tdtype, maximum):
    rng = np.random.default_rng(7)
    x = np.concatenate([
        rng.standard_normal(200_000).astype(np.float32) * np.float32(2.0) **
        rng.integers(-20, 16, 200_000).astype(np.float32),
        _boundaries(),
        # every tie point of the format
        np.array([], np.float32),
    ]).astype(np.float32)
    if maximum is not None:
        x = x[np.abs(x) <= maximum]
    ours, _ = C.encode(fmt, x)
    ref = torch.from_numpy(x).to(tdtype).view(torch.uint8 if fmt != "bf16" else torch.int16)
    ref = ref.numpy().view(np.uint8 if fmt != "bf16" else np.uint16)
    np.testing.assert_array_equal(ours, ref)


def test_decode_tables_closed_form():
    for c in range(256):
        s = -1.0 if c & 0x80 else 1.0
        e, f = (c >> 3) & 0xF, c & 7
        want = NAN if (e == 15 and f == 7) else s * (f * 2.0**-9 if e == 0 else (1 + f / 8) * 2.0**(e - 7))
        got = C.decode("fp8_e4m3fn", np.array([c], np.uint8), None, 1)[0]
        assert (np.isnan(want) and np.isnan(got)) or got == want
        e, f = (c >> 2) & 0x1F, c & 3
        if e == 31:
            want = s * INF if f == 0 else NAN
        else:
            want = s * (f * 2.0**-16 if e == 0 else (1 + f / 4) * 2.0**(e - 15))
        got = C.decode("fp8_e5m2", np.array([c], np.uint8), None, 1)[0]
        assert (np.isnan(want) and np.isnan(got)) or got == want
    table = [0, 0.5, 1, 1.5, 2, 3, 4, 6]
    for c in range(16):
        want = (-1 if c & 8 else 1) * table[c & 7]
        got = C.decode("fp4_e2m1", np.array([c], np.uint8), None, 1)[0]
        assert got == want


def test_fp4_rational_brute_force():
    """Nearest E2M1 code with ties to even, in exact rationals."""
    grid = [fractions.Fraction(k, 4) for k in range(0, 31)]
    codes = [fractions.Fraction(v) for v in (0, 0.5, 1, 1.5, 2, 3, 4, 6)]
    for x in grid:
        best = min(range(8), key=lambda c: (abs(x - codes[c]), c & 1))
        if x > 6:
            best = 7
        got = C.encode("fp4_e2m1", np.array([float(x)], np.float32))[0][0] & 0xF
        assert got == best, x


def test_fp8_scaled_oracle():
    x = np.random.default_rng(3).standard_normal(4096).astype(np.float32) * 1e-3
    codes, sigma = C.fp8_scaled(x)
    assert sigma == np.float32(448.0 / float(np.abs(x).max()))
    back = C.decode_fp8_scaled(codes, sigma, x.size)
    assert np.max(np.abs(back - x)) <= np.abs(x).max() * 2.0**-4


def test_device_codecs_match_oracle_on_host(tmp_path):
    exe = tmp_path / "host_codec_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-I", str(ROOT / "mixed-mlx_amd" / "csrc"),
                    "-x", "c++", str(ROOT / "tests" / "host_codec_check.cpp"),
                    "-x", "c", str(ROOT / "oracle" / "codecs_ref.c"), "-o", str(exe), "-lm"],
                   check=True)
    # "full": every float bit pattern against the generic formulation too.
    out = subprocess.run([str(exe), "99991", "997", "full"], capture_output=True, text=True,
                         timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
