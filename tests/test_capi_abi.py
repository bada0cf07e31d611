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
    assert capi.lib.msq_abi_version() == capi.ABI_VERSION == 6


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


def test_report_records_match_the_c_struct():
    """The record array that msq_solve_batched fills has the C struct's layout,
    and a record reads back as the same dict as the ctypes structure."""
    import ctypes

    import numpy as np

    from msq import capi
    dt = capi.REPORT_DTYPE
    assert dt.itemsize == ctypes.sizeof(capi.SolveReport)
    for name, _ in capi.SolveReport._fields_:
        assert dt.fields[name][1] == getattr(capi.SolveReport, name).offset, name
    r = capi.SolveReport()
    r.order, r.sweeps, r.corrections, r.status = 40, 3, 2, 1
    r.matrix_norm, r.final_residual = 2.5, 1e-9
    for k in range(3):
        r.residuals[k] = 10.0 ** -k
        r.off_history[k] = 0.5 ** k
    rec = np.frombuffer(ctypes.string_at(ctypes.addressof(r), ctypes.sizeof(r)), dtype=dt)
    assert capi.report_dict(rec[0]) == capi.report_dict(r)
    reps = capi.Reports(np.concatenate([rec, rec]))
    reps.records["status"][1] = 2
    assert list(reps.converged) == [True, False] and reps[1]["status"] == "budget_exhausted"
    assert [d["corrections"] for d in reps] == [2, 2]
    reps.replace(1, capi.Reports(rec.copy()), 0)
    assert reps.converged.all() and reps[1]["fallback_from"]["status"] == "budget_exhausted"
    assert reps[0]["residuals"] == [1.0, 0.1, 0.01] and reps[-1]["off_history"] == [1.0, 0.5, 0.25]


def test_batch_layout_validates_orders():
    """Every batched entry point validates its batch through the layout:
    an empty batch and orders outside [1, 65504] are errors, order 1 is a
    matrix like any other (a 32 x 32 tile)."""
    import pytest

    from msq import capi
    for bad in ([], [0], [16, 0, 32], [65505]):
        with pytest.raises(capi.MsqError):
            capi.batch_layout(bad)
    lds, offs, voffs, total, vtotal = capi.batch_layout([1, 65504])
    assert list(lds) == [32, 65504] and list(offs) == [0, 1024]
    assert total == 1024 + 65504 ** 2 and vtotal == 32 + 65504
