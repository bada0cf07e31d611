"""HIP codec kernels against the codec oracle, bit for bit.

Inputs: the reference's edge tables and boundary values, seeded random floats
across all magnitudes (with zeros, subnormals, infinities and NaNs), odd
lengths, exhaustive decode tables, and the n = 4096 basis shape
(16,777,216 elements).
"""
import math

import numpy as np
import pytest
import torch

from gpu_util import dev, empty, host, say
from oracle import codecs as C

pytestmark = pytest.mark.gpu

FORMATS = ["bf16", "fp8_e4m3fn", "fp8_e5m2", "fp4_e2m1", "mxfp8_e4m3fn", "mxfp8_e5m2",
           "mxfp4_e2m1"]


def mixed_floats(n, seed):
    rng = np.random.default_rng(seed)
    x = (rng.standard_normal(n) * 2.0 ** rng.integers(-40, 40, n)).astype(np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 2.0**-149, -2.0**-149, 2.0**-126,
                        448.0, 464.0, -464.0, 57344.0, 61440.0, 65504.0, 65520.0, 6.0, 6.1,
                        5.0, 0.25, 0.75, 1.0625, 1.1875, 2.0**-10, 3 * 2.0**-11], np.float32)
    idx = rng.integers(0, n, special.size * 50)
    x[idx] = np.resize(special, idx.size)
    # bit patterns straight from the integers, including NaN payloads
    raw = rng.integers(0, 2**32, n // 8, dtype=np.uint64).astype(np.uint32).view(np.float32)
    x[: raw.size] = raw
    return x


def run_codec(runtime, fmt, x):
    from msq import capi
    p = capi.PRECISION_BY_NAME[fmt]
    n = x.size
    src = dev(x)
    vals = empty(capi.storage_bytes(p, n))
    scales = empty(capi.scale_bytes(p, n)) if fmt.startswith("mx") else None
    dec = torch.empty(max(n, 1), dtype=torch.float32, device="cuda")
    runtime.quantize(p, src, vals, scales)
    runtime.decode(p, vals, scales, dec[:n] if n else dec)
    nb = capi.storage_bytes(p, n)
    codes = host(vals)[:nb]
    sc = host(scales)[: capi.scale_bytes(p, n)] if scales is not None else None
    return codes, sc, host(dec)[:n]


def compare(fmt, x, codes, sc, dec):
    ref_codes, ref_sc = C.encode(fmt, x)
    ref_codes = ref_codes.view(np.uint8)
    bad = np.nonzero(codes != ref_codes)[0]
    assert bad.size == 0, f"{fmt}: {bad.size} code mismatches, first at byte {bad[0]}"
    if ref_sc is not None:
        np.testing.assert_array_equal(sc, ref_sc)
    ref_dec = C.decode(fmt, ref_codes, ref_sc, x.size)
    np.testing.assert_array_equal(dec.view(np.uint32), ref_dec.view(np.uint32))


@pytest.mark.parametrize("fmt", FORMATS)
def test_codec_random_and_edges(runtime, fmt):
    for n, seed in ((1_000_003, 1), (33, 2), (65, 3), (1, 4), (2, 5), (31, 6), (4097, 7)):
        say(f"codec {fmt} n={n} seed={seed}")
        x = mixed_floats(n, seed)
        compare(fmt, x, *run_codec(runtime, fmt, x))


@pytest.mark.parametrize("fmt", ["mxfp8_e4m3fn", "mxfp8_e5m2", "mxfp4_e2m1"])
def test_mx_block_edge_table(runtime, fmt):
    """src/metal_checks.cpp:282-330 block table through the device kernels."""
    first = {"mxfp8_e4m3fn": (-8, 2.0**-17), "mxfp8_e5m2": (-15, 2.0**-31),
             "mxfp4_e2m1": (-2, 0.125)}[fmt]
    x = np.zeros(64, np.float32)
    x[0], x[2], x[3], x[4], x[5], x[6] = 1.0, -0.0, first[1], 0.25 * first[1], np.inf, np.nan
    x[32], x[34], x[35], x[36] = 8.0, -0.0, math.ldexp(first[1], 3), math.ldexp(0.25 * first[1], 3)
    codes, sc, dec = run_codec(runtime, fmt, x)
    compare(fmt, x, codes, sc, dec)
    assert sc[0] == first[0] + 127 and sc[1] == first[0] + 130
    assert dec[0] == 1.0 and dec[32] == 8.0


@pytest.mark.parametrize("fmt", ["mxfp8_e4m3fn", "mxfp8_e5m2", "mxfp4_e2m1"])
def test_mx_decode_exhaustive(runtime, fmt):
    """Every code under every scale byte (E8M0 0..255)."""
    from msq import capi
    p = capi.PRECISION_BY_NAME[fmt]
    if fmt == "mxfp4_e2m1":
        codes = np.tile(np.arange(16, dtype=np.uint8), 256 * 2)  # 32 nibbles per block
        packed = (codes[0::2] | (codes[1::2] << 4)).astype(np.uint8)
        n = codes.size
    else:
        codes = np.tile(np.arange(256, dtype=np.uint8), 256 // 8 * 8)[: 256 * 32]
        packed = codes
        n = codes.size
    scales = (np.arange(n // 32) % 256).astype(np.uint8)
    out = torch.empty(n, dtype=torch.float32, device="cuda")
    runtime.decode(p, dev(packed), dev(scales), out)
    ref = C.decode(fmt, packed, scales, n)
    got = host(out)
    same = (got.view(np.uint32) == ref.view(np.uint32)) | (np.isnan(got) & np.isnan(ref))
    assert same.all()


def test_fp8_scaled_bit_exact(runtime):
    for n, seed in ((1 << 20, 11), (100, 12)):
        say(f"fp8 scaled n={n}")
        x = (np.random.default_rng(seed).standard_normal(n) * 3e-4).astype(np.float32)
        vals = empty(n)
        sigma = torch.zeros(1, dtype=torch.float32, device="cuda")
        runtime.quantize_fp8_scaled(dev(x), vals, sigma)
        ref, rs = C.fp8_scaled(x)
        assert host(sigma)[0] == rs
        np.testing.assert_array_equal(host(vals)[:n], ref)
        dec = torch.empty(n, dtype=torch.float32, device="cuda")
        runtime.decode_fp8_scaled(vals, sigma, dec)
        np.testing.assert_array_equal(host(dec), C.decode_fp8_scaled(ref, rs, n))


@pytest.mark.parametrize("fmt", ["bf16", "fp8_e4m3fn", "mxfp4_e2m1"])
def test_codec_n4096_basis_shape(runtime, fmt):
    """All 16,777,216 elements of an n = 4096 orthogonal-like basis."""
    n = 4096 * 4096
    say(f"codec {fmt} n={n}")
    rng = np.random.default_rng(99)
    x = (rng.standard_normal(n) / 64.0).astype(np.float32)
    compare(fmt, x, *run_codec(runtime, fmt, x))
