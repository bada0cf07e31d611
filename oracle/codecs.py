"""ctypes wrapper of the C codec oracle (oracle/codecs_ref.c).

Test infrastructure only: the checker the HIP codec kernels are compared with.
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "liboracle_codecs.so"


def build():
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


if not LIB.exists():
    build()
_lib = ctypes.CDLL(str(LIB))
_fp = ctypes.POINTER(ctypes.c_float)
_u8 = ctypes.POINTER(ctypes.c_uint8)
_u16 = ctypes.POINTER(ctypes.c_uint16)
_sz = ctypes.c_size_t
for name, args in [
    ("oracle_encode_bf16", [_fp, _u16, _sz]),
    ("oracle_encode_fp8_e4m3fn", [_fp, _u8, _sz]),
    ("oracle_encode_fp8_e5m2", [_fp, _u8, _sz]),
    ("oracle_encode_fp4_e2m1", [_fp, _u8, _sz]),
    ("oracle_encode_mxfp8_e4m3fn", [_fp, _u8, _u8, _sz]),
    ("oracle_encode_mxfp8_e5m2", [_fp, _u8, _u8, _sz]),
    ("oracle_encode_mxfp4_e2m1", [_fp, _u8, _u8, _sz]),
    ("oracle_decode", [ctypes.c_int, _u8, _u8, _fp, _sz]),
    ("oracle_encode_fp8_e4m3fn_scaled", [_fp, _u8, _sz, ctypes.c_float]),
]:
    getattr(_lib, name).argtypes = args
    getattr(_lib, name).restype = None
_lib.oracle_decode.restype = ctypes.c_int
_lib.oracle_fp8_amax_scale.argtypes = [_fp, _sz]
_lib.oracle_fp8_amax_scale.restype = ctypes.c_float

# Format ids of oracle_decode (the reference's HostStorageCodec order).
DECODE_ID = {"bf16": 1, "fp8_e4m3fn": 2, "fp8_e5m2": 3, "mxfp8_e4m3fn": 4,
             "mxfp8_e5m2": 5, "fp4_e2m1": 6, "mxfp4_e2m1": 7}


def _f(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    return x, x.ctypes.data_as(_fp)


def encode(fmt: str, values):
    """Returns (codes uint8 / uint16, scales uint8 or None)."""
    x, p = _f(values)
    n = x.size
    if fmt == "bf16":
        out = np.zeros(n, np.uint16)
        _lib.oracle_encode_bf16(p, out.ctypes.data_as(_u16), n)
        return out, None
    if fmt in ("fp8_e4m3fn", "fp8_e5m2"):
        out = np.zeros(n, np.uint8)
        getattr(_lib, "oracle_encode_" + fmt)(p, out.ctypes.data_as(_u8), n)
        return out, None
    if fmt == "fp4_e2m1":
        out = np.zeros((n + 1) // 2, np.uint8)
        _lib.oracle_encode_fp4_e2m1(p, out.ctypes.data_as(_u8), n)
        return out, None
    if fmt in ("mxfp8_e4m3fn", "mxfp8_e5m2", "mxfp4_e2m1"):
        nbytes = (n + 1) // 2 if fmt == "mxfp4_e2m1" else n
        out = np.zeros(nbytes, np.uint8)
        sc = np.zeros((n + 31) // 32, np.uint8)
        getattr(_lib, "oracle_encode_" + fmt)(p, out.ctypes.data_as(_u8),
                                              sc.ctypes.data_as(_u8), n)
        return out, sc
    raise ValueError(fmt)


def decode(fmt: str, codes, scales, n):
    out = np.zeros(n, np.float32)
    codes = np.ascontiguousarray(codes).view(np.uint8)
    sc = None if scales is None else np.ascontiguousarray(scales, np.uint8)
    rc = _lib.oracle_decode(DECODE_ID[fmt], codes.ctypes.data_as(_u8),
                            None if sc is None else sc.ctypes.data_as(_u8),
                            out.ctypes.data_as(_fp), n)
    if rc != 0:
        raise ValueError(fmt)
    return out


def fp8_scaled(values):
    """(codes, sigma): per-tensor amax-scaled E4M3FN."""
    x, p = _f(values)
    sigma = _lib.oracle_fp8_amax_scale(p, x.size)
    out = np.zeros(x.size, np.uint8)
    _lib.oracle_encode_fp8_e4m3fn_scaled(p, out.ctypes.data_as(_u8), x.size,
                                         ctypes.c_float(sigma))
    return out, np.float32(sigma)


def decode_fp8_scaled(codes, sigma, n):
    vals = decode("fp8_e4m3fn", codes, None, n)
    return (vals / np.float32(sigma)).astype(np.float32)


_lib.oracle_stochastic_hash.argtypes = [ctypes.c_uint32] * 3
_lib.oracle_stochastic_hash.restype = ctypes.c_uint32
_lib.oracle_encode_stochastic.argtypes = [ctypes.c_int, _fp, _u8, _u8, _sz, ctypes.c_uint32,
                                          ctypes.c_uint32]
_lib.oracle_encode_stochastic.restype = ctypes.c_int
_lib.oracle_encode_log.argtypes = [ctypes.c_int, _fp, _u8, _fp, _sz, ctypes.c_int,
                                   ctypes.c_uint32, ctypes.c_uint32]
_lib.oracle_encode_log.restype = None
_lib.oracle_decode_log.argtypes = [ctypes.c_int, _u8, _fp, _fp, _sz]
_lib.oracle_decode_log.restype = None


def stochastic_hash(index, seed, stream):
    return int(_lib.oracle_stochastic_hash(index, seed, stream))


def encode_stochastic(fmt: str, values, seed, stream):
    """Stochastically rounded codes (and MX scales or None)."""
    x, p = _f(values)
    n = x.size
    fid = DECODE_ID[fmt]
    nbytes = {1: 2 * n, 6: (n + 1) // 2, 7: (n + 1) // 2}.get(fid, n)
    out = np.zeros(max(nbytes, 1), np.uint8)
    sc = np.zeros(max((n + 31) // 32, 1), np.uint8)
    rc = _lib.oracle_encode_stochastic(fid, p, out.ctypes.data_as(_u8), sc.ctypes.data_as(_u8),
                                       n, seed, stream)
    if rc != 0:
        raise ValueError(fmt)
    codes = out[:nbytes].view(np.uint16) if fid == 1 else out[:nbytes]
    return codes, (sc if fmt.startswith("mx") else None)


def encode_log(e4m3: bool, values, stochastic=False, seed=0, stream=0):
    """Log-affine spectrum codes and metadata (min, span per block of 8)."""
    x, p = _f(values)
    n = x.size
    codes = np.zeros(n if e4m3 else (n + 1) // 2, np.uint8)
    meta = np.zeros(2 * ((n + 7) // 8), np.float32)
    _lib.oracle_encode_log(1 if e4m3 else 0, p, codes.ctypes.data_as(_u8),
                           meta.ctypes.data_as(_fp), n, 1 if stochastic else 0, seed, stream)
    return codes, meta


def decode_log(e4m3: bool, codes, metadata, n):
    out = np.zeros(n, np.float32)
    c = np.ascontiguousarray(codes, np.uint8)
    m = np.ascontiguousarray(metadata, np.float32)
    _lib.oracle_decode_log(1 if e4m3 else 0, c.ctypes.data_as(_u8), m.ctypes.data_as(_fp),
                           out.ctypes.data_as(_fp), n)
    return out
