"""The C-ABI library loads and exports every symbol include/msq_hip.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import pathlib
import re

ROOT = pathlib.Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "msq_hip.h").read_text()
    return sorted(set(re.findall(r"\b(msq_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from msq import capi
    declared = declared_symbols()
    assert len(declared) >= 20
    assert sorted(capi.EXPORTED_SYMBOLS) == declared
    for name in declared:
        assert hasattr(capi.lib, name), name
    assert capi.lib.msq_abi_version() == 4


def test_storage_bytes_match_reference_formulas():
    """storageBytes / scaleBytes, src/configuration.cpp:188-213."""
    from msq import capi
    n = 33
    assert capi.storage_bytes(capi.FP32, n) == 132
    assert capi.storage_bytes(capi.BF16, n) == 66
    assert capi.storage_bytes(capi.FP8_E4M3FN, n) == 33
    assert capi.storage_bytes(capi.MXFP4_E2M1, n) == 17
    assert capi.scale_bytes(capi.MXFP4_E2M1, n) == 2
    assert capi.scale_bytes(capi.FP8_E5M2, n) == 0


def test_batch_layout():
    from msq import capi
    lds, offs, voffs, total, vtotal = capi.batch_layout([16, 33, 64])
    assert list(lds) == [32, 64, 64]
    assert list(offs) == [0, 1024, 1024 + 4096]
    assert list(voffs) == [0, 32, 96]
    assert total == 1024 + 4096 + 4096 and vtotal == 160
    orders, flat = capi.pack_batch([np.eye(2) * 3 for np in [__import__("numpy")]])
    assert orders == [2] and flat[0] == 3 and flat[33] == 3
