"""Stochastic rounding and the log-affine spectrum codec on the GPU, through
the C ABI (msq_runtime_set_rounding, msq_quantize_stream, msq_decode): codes,
scales, metadata and decoded values bit for bit against the oracle
(oracle/codecs_ref.c), and the reference's own checks on the GPU's output
(src/metal_checks.cpp:360-571)."""
import numpy as np
import pytest
import torch

from gpu_util import dev, empty, host, say
from msq import capi
from oracle import codecs as C
from test_gpu_codecs import FORMATS, mixed_floats

pytestmark = pytest.mark.gpu


def gpu_stochastic(runtime, fmt, x, seed, stream):
    p = capi.PRECISION_BY_NAME[fmt]
    n = x.size
    vals = empty(capi.storage_bytes(p, n))
    scales = empty(capi.scale_bytes(p, n)) if fmt.startswith("mx") else None
    dec = torch.empty(max(n, 1), dtype=torch.float32, device="cuda")
    runtime.set_rounding(capi.ROUNDING_STOCHASTIC, seed)
    try:
        runtime.quantize_stream(p, dev(x), vals, scales, stream)
    finally:
        runtime.set_rounding(capi.ROUNDING_NEAREST, 0)
    runtime.decode(p, vals, scales, dec[:n])
    codes = host(vals)[:capi.storage_bytes(p, n)]
    sc = host(scales)[:capi.scale_bytes(p, n)] if scales is not None else None
    return codes, sc, host(dec)[:n]


@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("n", [1, 31, 1000003])
def test_stochastic_codes_match_oracle(runtime, fmt, n):
    x = mixed_floats(n, 17 + n)
    seed, stream = 0xC0FFEE + n, 3
    say(f"stochastic {fmt} n={n}")
    codes, sc, dec = gpu_stochastic(runtime, fmt, x, seed, stream)
    ref, ref_sc = C.encode_stochastic(fmt, x, seed, stream)
    ref = ref.view(np.uint8)
    bad = np.nonzero(codes != ref)[0]
    assert bad.size == 0, f"{fmt}: {bad.size} mismatches, first at byte {bad[0]}"
    if ref_sc is not None:
        np.testing.assert_array_equal(sc, ref_sc)
    ref_dec = C.decode(fmt, ref, ref_sc, n)
    np.testing.assert_array_equal(dec.view(np.uint32), ref_dec.view(np.uint32))


@pytest.mark.parametrize("fmt", FORMATS)
def test_stochastic_reference_check_on_gpu(runtime, fmt):
    """src/metal_checks.cpp:409-535 on the GPU's codes."""
    n = 65536
    src = np.where(np.arange(n) < n // 2, 0.3, -0.3).astype(np.float32)
    c1, _, d1 = gpu_stochastic(runtime, fmt, src, 0x53520000, 0)
    c1b, _, _ = gpu_stochastic(runtime, fmt, src, 0x53520000, 0)
    c2, _, _ = gpu_stochastic(runtime, fmt, src, 0x53520001, 0)
    assert np.array_equal(c1, c1b), "not reproducible"
    assert not np.array_equal(c1, c2), "the seed does not change the codes"
    d = d1.astype(np.float64)
    err = max(abs(d[:n // 2].mean() - 0.3), abs(d[n // 2:].mean() + 0.3))
    tol = 5e-5 if fmt == "bf16" else (1e-2 if "fp4" in fmt else 1e-3)
    assert err <= tol


def gpu_log(runtime, e4m3, x, stochastic=False, seed=0, stream=1):
    p = capi.FP8_E4M3FN_LOG_S if e4m3 else capi.FP4_E2M1_LOG_S
    n = x.size
    codes = empty(capi.storage_bytes(p, n))
    meta = torch.zeros(max(1, capi.spectrum_metadata_bytes(p, n) // 4), dtype=torch.float32,
                       device="cuda")
    dec = torch.empty(max(n, 1), dtype=torch.float32, device="cuda")
    if stochastic:
        runtime.set_rounding(capi.ROUNDING_STOCHASTIC, seed)
    try:
        runtime.quantize_stream(p, dev(x), codes, meta, stream)
    finally:
        runtime.set_rounding(capi.ROUNDING_NEAREST, 0)
    runtime.decode(p, codes, meta, dec[:n])
    return (host(codes)[:capi.storage_bytes(p, n)], host(meta)[:2 * ((n + 7) // 8)],
            host(dec)[:n])


@pytest.mark.parametrize("e4m3", [True, False])
@pytest.mark.parametrize("stochastic", [False, True])
def test_log_spectrum_matches_oracle(runtime, e4m3, stochastic):
    rng = np.random.default_rng(23)
    n = 100003
    x = np.exp2(rng.uniform(-60, 60, n)).astype(np.float32)
    x[:16] = 3.0  # constant blocks
    x[16:24] = 2.0 ** np.arange(8)
    say(f"log spectrum e4m3={e4m3} stochastic={stochastic}")
    codes, meta, dec = gpu_log(runtime, e4m3, x, stochastic, 0x4c4f4753, 1)
    rc, rm = C.encode_log(e4m3, x, stochastic, 0x4c4f4753, 1)
    np.testing.assert_array_equal(meta.view(np.uint32), rm.view(np.uint32))
    np.testing.assert_array_equal(codes, rc)
    rd = C.decode_log(e4m3, rc, rm, n)
    np.testing.assert_array_equal(dec.view(np.uint32), rd.view(np.uint32))


@pytest.mark.parametrize("e4m3,bound", [(True, 0.5), (False, 4.0)])
def test_log_spectrum_reference_checks_on_gpu(runtime, e4m3, bound):
    edge = np.array([np.finfo(np.float32).tiny, 2.0 ** -100, 2.0 ** -64, 2.0 ** -32, 1.0,
                     2.0 ** 32, 2.0 ** 64, np.finfo(np.float32).max] + [3.0] * 8, np.float32)
    for stochastic in (False, True):
        _, _, d = gpu_log(runtime, e4m3, edge, stochastic, 0x4c4f4753, 0)
        assert d[0] > 0 and np.isfinite(d[0])
        assert d[7] > 0 and np.isfinite(d[7])
        assert np.all(np.abs(d[8:16] / 3.0 - 1.0) <= 1e-6)
    roots = np.array([1.0, 1.5, 2.0, 3.0, 10.0, 31.622776, 100.0, 1000.0], np.float32)
    _, _, d = gpu_log(runtime, e4m3, roots)
    assert np.all(np.isfinite(d) & (d > 0))
    assert np.max(np.abs(d / roots - 1.0)) < bound
