"""ctypes binding of libmsq_hip.so (include/msq_hip.h).

Mirrors the reference's device interface, MetalRuntime
(src/metal_runtime.hpp:69-176), as a Python class: one runtime per device,
methods named after the reference's (quantizeAndDecode -> quantize_and_decode,
symmetricJacobi -> jacobi_batched, ...).  Device memory is passed as torch
tensors (or raw addresses); torch is plumbing here, every operation runs in
the HIP library.  Importing this module fails loudly when the library is
missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

LIB_PATH = pathlib.Path(os.environ.get(
    "MSQ_HIP_LIBRARY", pathlib.Path(__file__).resolve().parent / "libmsq_hip.so"))

if not LIB_PATH.exists():
    raise ImportError(
        f"libmsq_hip.so not found at {LIB_PATH}; build it with "
        "`python mixed-mlx_amd/build.py` (or __graft_entry__.build())")

# torch ships its own HIP runtime under the same soname as the system one the
# library links against.  The process must load torch's first: if the
# library's runtime is loaded first, torch's device initialisation fails.
try:
    import torch  # noqa: F401
except ImportError:  # the library itself does not need torch
    pass

lib = ctypes.CDLL(str(LIB_PATH))

# Precision values (src/precision.hpp:9-20 order) and the scaled FP8 tier.
FP32, BF16, FP8_E4M3FN, FP8_E5M2, MXFP8_E4M3FN, MXFP8_E5M2, FP4_E2M1, \
    MXFP4_E2M1, FP8_E4M3FN_LOG_S, FP4_E2M1_LOG_S = range(10)
FP8_E4M3FN_SCALED = 16
PRECISION_NAMES = {
    FP32: "fp32", BF16: "bf16", FP8_E4M3FN: "fp8_e4m3fn", FP8_E5M2: "fp8_e5m2",
    MXFP8_E4M3FN: "mxfp8_e4m3fn", MXFP8_E5M2: "mxfp8_e5m2", FP4_E2M1: "fp4_e2m1",
    MXFP4_E2M1: "mxfp4_e2m1", FP8_E4M3FN_LOG_S: "fp8_e4m3fn_log_s",
    FP4_E2M1_LOG_S: "fp4_e2m1_log_s", FP8_E4M3FN_SCALED: "fp8_e4m3fn_scaled",
}
PRECISION_BY_NAME = {v: k for k, v in PRECISION_NAMES.items()}

SOLVE_STATUS = {0: "running", 1: "converged", 2: "budget_exhausted",
                3: "diverged", 4: "nonfinite", 5: "invalid_spectrum", 6: "stagnated",
                7: "unresolved"}
JACOBI_HISTORY = 32
SOLVE_HISTORY = 65
ABI_VERSION = 6
# KernelPath / CorrectionMode / NativeStoppingPolicy / NativeDispatchPolicy /
# RoundingMode (src/precision.hpp:22, :126-146), the reference's order.
KERNEL_PATH_LEGACY, KERNEL_PATH_TILED = 0, 1
CORRECTION_FP32_REORTHOGONALIZED, CORRECTION_NATIVE_QUANTIZED = 0, 1
STOPPING_FIXED_BUDGET, STOPPING_STAGNATION = 0, 1
DISPATCH_PER_STAGE, DISPATCH_FUSED = 0, 1
ROUNDING_NEAREST, ROUNDING_STOCHASTIC = 0, 1
# msq_jacobi_method: eigensolver of the orders up to 256.
JACOBI_SCALAR, JACOBI_BLOCKED_MFMA = 0, 1
JACOBI_DEFAULT = JACOBI_BLOCKED_MFMA
OWN_STREAM = 2**64 - 1  # MSQ_OWN_STREAM
MLP_INPUTS = 16
MLP_HIDDEN = 64
MLP_OUTPUTS = 8
MLP_PARAMS = 5768
FEATURE_MAX_ORDER = 512
NOALLOC_MAX_MATRICES = 4096

# Every symbol include/msq_hip.h declares.
EXPORTED_SYMBOLS = (
    "msq_last_error", "msq_abi_version", "msq_runtime_create", "msq_runtime_destroy",
    "msq_runtime_set_stream", "msq_synchronize", "msq_runtime_set_graphs",
    "msq_runtime_set_rounding", "msq_runtime_set_jacobi_method", "msq_alloc", "msq_free",
    "msq_copy_to_device", "msq_copy_to_host", "msq_memset", "msq_storage_bytes",
    "msq_scale_bytes", "msq_spectrum_metadata_bytes", "msq_batch_layout", "msq_quantize",
    "msq_quantize_stream", "msq_decode",
    "msq_quantize_and_decode", "msq_quantize_fp8_scaled", "msq_decode_fp8_scaled",
    "msq_jacobi_batched", "msq_jacobi_batched_noalloc", "msq_reconstruct", "msq_residual_fp64",
    "msq_correction", "msq_refine_native_staged",
    "msq_orthogonality", "msq_solve_batched", "msq_solve_batched_ex", "msq_solve_options_init",
    "msq_last_factors",
    "msq_matrix_features", "msq_matrix_features_subset", "msq_mlp_forward",
    "msq_mlp_forward_bf16", "msq_mlp_backward",
    "msq_mlp_adam", "msq_autotune", "msq_set_tile_table",
)
TILE_TYPES = ("fp64_residual", "fp32", "bf16", "fp8_scaled", "mxfp4")
TILE_CLASSES = (32, 64, 96, 128, 160, 192, 224, 256, 512)


class JacobiOptions(ctypes.Structure):
    _fields_ = [("fixed_sweeps", ctypes.c_uint32), ("max_sweeps", ctypes.c_uint32),
                ("off_tol", ctypes.c_float), ("off_tol_lambda", ctypes.c_float),
                ("stagnation", ctypes.c_float), ("floor_tol", ctypes.c_float),
                ("correction_rate", ctypes.c_float)]


class JacobiReport(ctypes.Structure):
    _fields_ = [("sweeps", ctypes.c_uint32), ("invalid_spectrum", ctypes.c_uint32),
                ("off_norm", ctypes.c_double), ("matrix_norm", ctypes.c_double),
                ("min_diagonal", ctypes.c_float),
                ("off_history", ctypes.c_float * JACOBI_HISTORY),
                ("correction_rate", ctypes.c_float)]


class SolveOptions(ctypes.Structure):
    _fields_ = [("basis_precision", ctypes.c_int32), ("correction_precision", ctypes.c_int32),
                ("jacobi", JacobiOptions), ("min_corrections", ctypes.c_uint32),
                ("max_corrections", ctypes.c_uint32), ("residual_tolerance", ctypes.c_double),
                ("forward_tolerance", ctypes.c_double),
                ("spectrum_precision", ctypes.c_int32), ("kernel_path", ctypes.c_int32),
                ("correction_mode", ctypes.c_int32), ("stopping_policy", ctypes.c_int32),
                ("dispatch_policy", ctypes.c_int32), ("stagnation_window", ctypes.c_uint32),
                ("minimum_relative_improvement", ctypes.c_double),
                ("basis_tolerance", ctypes.c_double), ("resolution", ctypes.c_float),
                ("refine_unresolved", ctypes.c_uint32), ("certificate", ctypes.c_uint32)]


class SolveReport(ctypes.Structure):
    _fields_ = [("order", ctypes.c_uint32), ("sweeps", ctypes.c_uint32),
                ("corrections", ctypes.c_uint32), ("status", ctypes.c_uint32),
                ("matrix_norm", ctypes.c_double), ("min_root", ctypes.c_double),
                ("off_norm", ctypes.c_double), ("final_residual", ctypes.c_double),
                ("final_forward_estimate", ctypes.c_double),
                ("residuals", ctypes.c_double * SOLVE_HISTORY),
                ("forward_estimates", ctypes.c_double * SOLVE_HISTORY),
                ("off_history", ctypes.c_float * JACOBI_HISTORY),
                ("correction_rate", ctypes.c_float),
                ("basis_orthogonality", ctypes.c_float),
                ("precision_escalations", ctypes.c_uint32),
                ("basis_rejected", ctypes.c_uint32), ("spectrum_rejected", ctypes.c_uint32),
                ("frozen_sylvester_inverse_norm", ctypes.c_double),
                ("frozen_basis_orthogonality", ctypes.c_double),
                ("certificate_usable", ctypes.c_uint32), ("unresolved", ctypes.c_uint32),
                ("refined", ctypes.c_uint32), ("best_iteration", ctypes.c_uint32),
                ("stagnation_steps", ctypes.c_uint32), ("reserved", ctypes.c_uint32),
                ("best_residual", ctypes.c_double)]


class SolveOutputs(ctypes.Structure):
    _fields_ = [("roots", ctypes.c_void_p), ("initial_roots", ctypes.c_void_p),
                ("vectors", ctypes.c_void_p), ("root_eigenvalues", ctypes.c_void_p)]


_p = ctypes.c_void_p
_u32 = ctypes.c_uint32
_sz = ctypes.c_size_t
_i = ctypes.c_int
_u32p = ctypes.POINTER(ctypes.c_uint32)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_f64p = ctypes.POINTER(ctypes.c_double)


def _sig(name, restype, *argtypes):
    f = getattr(lib, name)
    f.restype = restype
    f.argtypes = list(argtypes)
    return f


_sig("msq_last_error", ctypes.c_char_p)
_sig("msq_abi_version", _i)
_sig("msq_runtime_create", _i, _i, ctypes.POINTER(_p))
_sig("msq_runtime_destroy", _i, _p)
_sig("msq_runtime_set_stream", _i, _p, _p)
_sig("msq_synchronize", _i, _p)
_sig("msq_runtime_set_graphs", _i, _p, _i)
_sig("msq_autotune", _i, _p, _u32, _p)
_sig("msq_set_tile_table", _i, _p, _p)
_sig("msq_alloc", _i, _p, _sz, ctypes.POINTER(_p))
_sig("msq_free", _i, _p, _p)
_sig("msq_copy_to_device", _i, _p, _p, _p, _sz)
_sig("msq_copy_to_host", _i, _p, _p, _p, _sz)
_sig("msq_memset", _i, _p, _p, _i, _sz)
_sig("msq_storage_bytes", _sz, _i, _sz)
_sig("msq_scale_bytes", _sz, _i, _sz)
_sig("msq_batch_layout", _i, _u32, _u32p, _u32p, _u64p, _u64p, _u64p, _u64p)
_sig("msq_quantize", _i, _p, _i, _p, _p, _p, _sz)
_sig("msq_decode", _i, _p, _i, _p, _p, _p, _sz)
_sig("msq_quantize_and_decode", _i, _p, _i, _p, _p, _p, _p, _sz)
_sig("msq_quantize_fp8_scaled", _i, _p, _p, _p, _p, _sz)
_sig("msq_decode_fp8_scaled", _i, _p, _p, _p, _p, _sz)
_sig("msq_jacobi_batched", _i, _p, _u32, _u32p, _p, _p, _p, _p,
     ctypes.POINTER(JacobiOptions), _u32, ctypes.POINTER(JacobiReport))
_sig("msq_jacobi_batched_noalloc", _i, _p, _u32, _u32p, _p, _p, _p, _p,
     ctypes.POINTER(JacobiOptions), _u32, ctypes.POINTER(JacobiReport))
_sig("msq_reconstruct", _i, _p, _i, _u32, _u32p, _p, _p, _p)
_sig("msq_residual_fp64", _i, _p, _u32, _u32p, _p, _p, _p, _p, _f64p)
_sig("msq_correction", _i, _p, _i, _u32, _u32p, _p, _p, _p, _p)
_sig("msq_orthogonality", _i, _p, _u32, _u32p, _p, _f64p)
_sig("msq_solve_batched", _i, _p, _u32, _u32p, _p, _p, ctypes.POINTER(SolveOptions), _u32,
     ctypes.POINTER(SolveReport))
_sig("msq_solve_batched_ex", _i, _p, _u32, _u32p, _p, ctypes.POINTER(SolveOutputs),
     ctypes.POINTER(SolveOptions), _u32, ctypes.POINTER(SolveReport))
_sig("msq_solve_options_init", None, ctypes.POINTER(SolveOptions))
_sig("msq_refine_native_staged", _i, _p, _i, _u32, _u32p, _p, _p, _p, _p)
_sig("msq_runtime_set_rounding", _i, _p, _i, _u32)
_sig("msq_runtime_set_jacobi_method", _i, _p, _i)
_sig("msq_quantize_stream", _i, _p, _i, _p, _p, _p, _sz, _u32)
_sig("msq_spectrum_metadata_bytes", _sz, _i, _sz)
_sig("msq_last_factors", _i, _p, ctypes.POINTER(_p), ctypes.POINTER(_p))
_sig("msq_matrix_features", _i, _p, _u32, _u32p, _p, _p)
_sig("msq_matrix_features_subset", _i, _p, _u32, _u32p, _p, _u32, _u32p, _p)
_sig("msq_mlp_forward", _i, _p, _p, _p, _u32, _p, _p, _p)
_sig("msq_mlp_forward_bf16", _i, _p, _p, _p, _u32, _p)
_sig("msq_mlp_backward", _i, _p, _p, _p, _p, _p, _p, _p, _p, _u32, ctypes.c_float, _p, _p)
_sig("msq_mlp_adam", _i, _p, _p, _p, _p, _p, _u32, ctypes.c_float, ctypes.c_float,
     ctypes.c_float, ctypes.c_float, ctypes.c_float, _u32)


class MsqError(RuntimeError):
    pass


def check(status):
    if status != 0:
        raise MsqError(f"msq status {status}: {lib.msq_last_error().decode()}")


def storage_bytes(precision, elements):
    return lib.msq_storage_bytes(precision, elements)


def scale_bytes(precision, elements):
    return lib.msq_scale_bytes(precision, elements)


def spectrum_metadata_bytes(precision, elements):
    return lib.msq_spectrum_metadata_bytes(precision, elements)


def _addr(x):
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def _u32_array(values):
    # The pointer keeps a reference to the array for the call's duration.
    return np.ascontiguousarray(values, dtype=np.uint32).ctypes.data_as(_u32p)


# ---------------------------------------------------------------- layout
def batch_layout(orders):
    """ld = roundup(n, 32), offsets of ld x ld tiles and of ld-vectors."""
    orders = [int(n) for n in orders]
    count = len(orders)
    lds = (ctypes.c_uint32 * count)()
    offs = (ctypes.c_uint64 * count)()
    voffs = (ctypes.c_uint64 * count)()
    total = ctypes.c_uint64()
    vtotal = ctypes.c_uint64()
    check(lib.msq_batch_layout(count, _u32_array(orders), lds, offs, voffs,
                               ctypes.byref(total), ctypes.byref(vtotal)))
    return (np.array(lds[:], dtype=np.int64), np.array(offs[:], dtype=np.int64),
            np.array(voffs[:], dtype=np.int64), int(total.value), int(vtotal.value))


def pack_batch(matrices, dtype=np.float64):
    """Host arrays (n_i x n_i) -> (orders, flat column-major batch layout)."""
    orders = [m.shape[0] for m in matrices]
    lds, offs, _, total, _ = batch_layout(orders)
    flat = np.zeros(total, dtype=dtype)
    for m, ld, off in zip(matrices, lds, offs):
        n = m.shape[0]
        tile = np.zeros((ld, ld), dtype=dtype)
        tile[:n, :n] = m
        flat[off:off + ld * ld] = tile.reshape(-1, order="F")
    return orders, flat


def unpack_batch(orders, flat):
    lds, offs, _, _, _ = batch_layout(orders)
    out = []
    for n, ld, off in zip(orders, lds, offs):
        out.append(np.asarray(flat[off:off + ld * ld]).reshape(ld, ld, order="F")[:n, :n].copy())
    return out


def unpack_vectors(orders, flat):
    lds, _, voffs, _, _ = batch_layout(orders)
    return [np.asarray(flat[v:v + n]).copy() for n, v in zip(orders, voffs)]


# ---------------------------------------------------------------- options
def jacobi_options(fixed_sweeps=0, max_sweeps=30, off_tol=1e-7, off_tol_lambda=0.0,
                   stagnation=0.5, floor_tol=1e-5, correction_rate=0.0):
    return JacobiOptions(fixed_sweeps, max_sweeps, off_tol, off_tol_lambda, stagnation,
                         floor_tol, correction_rate)


def solve_options(basis=FP32, correction=FP32, jacobi=None, min_corrections=0,
                  max_corrections=12, residual_tolerance=1e-12, forward_tolerance=0.0,
                  **fields):
    """An msq_solve_options record: msq_solve_options_init's reference
    defaults, then these values; `fields` sets any other field by name
    (spectrum_precision, kernel_path, correction_mode, stopping_policy,
    dispatch_policy, stagnation_window, minimum_relative_improvement,
    basis_tolerance, resolution, refine_unresolved, certificate)."""
    o = SolveOptions()
    lib.msq_solve_options_init(ctypes.byref(o))
    o.basis_precision = basis
    o.correction_precision = correction
    o.jacobi = jacobi if jacobi is not None else jacobi_options()
    o.min_corrections = min_corrections
    o.max_corrections = max_corrections
    o.residual_tolerance = residual_tolerance
    o.forward_tolerance = forward_tolerance
    for k, v in fields.items():
        if k not in dict(SolveOptions._fields_):
            raise TypeError(f"no solve option {k}")
        setattr(o, k, v)
    return o


def _struct_dtype(struct):
    """numpy dtype with the field offsets of a ctypes structure."""
    names, formats, offsets = [], [], []
    for name, ctype in struct._fields_:
        names.append(name)
        offsets.append(getattr(struct, name).offset)
        if issubclass(ctype, ctypes.Array):
            formats.append((np.dtype(ctype._type_), (ctype._length_,)))
        else:
            formats.append(np.dtype(ctype))
    return np.dtype({"names": names, "formats": formats, "offsets": offsets,
                     "itemsize": ctypes.sizeof(struct)})


REPORT_DTYPE = _struct_dtype(SolveReport)
# One raw msq_solve_options record (per-matrix options as a numpy array).
OPTIONS_RECORD = np.dtype((np.void, ctypes.sizeof(SolveOptions)))


def options_array(records, index):
    """Per-matrix options as one array of raw msq_solve_options records:
    records[index[i]] for matrix i.  (A ctypes array built from a list of
    structures costs about 0.35 us a matrix, 1.4 ms for 4096; this is one
    gather.)"""
    table = np.frombuffer(b"".join(bytes(r) for r in records), dtype=OPTIONS_RECORD)
    return table[np.asarray(index, dtype=np.intp)]
CONVERGED = 1


def report_dict(r):
    """One report as a dict (r: a SolveReport or a REPORT_DTYPE record)."""
    corrections = int(r["corrections"] if isinstance(r, np.void) else r.corrections)
    get = (lambda k: r[k]) if isinstance(r, np.void) else (lambda k: getattr(r, k))
    steps = corrections + 1
    sweeps = int(get("sweeps"))
    status = int(get("status"))
    return {
        "order": int(get("order")), "sweeps": sweeps, "corrections": corrections,
        "status": SOLVE_STATUS.get(status, str(status)),
        "matrix_norm": float(get("matrix_norm")), "min_root": float(get("min_root")),
        "off_norm": float(get("off_norm")),
        "final_residual": float(get("final_residual")),
        "final_forward_estimate": float(get("final_forward_estimate")),
        "residuals": [float(x) for x in get("residuals")[:steps]],
        "forward_estimates": [float(x) for x in get("forward_estimates")[:steps]],
        "off_history": [float(x) for x in get("off_history")[:sweeps]],
        "correction_rate": float(get("correction_rate")),
        "basis_orthogonality": float(get("basis_orthogonality")),
        "precision_escalations": int(get("precision_escalations")),
        "basis_rejected": int(get("basis_rejected")),
        "spectrum_rejected": int(get("spectrum_rejected")),
        "frozen_sylvester_inverse_norm": float(get("frozen_sylvester_inverse_norm")),
        "frozen_basis_orthogonality": float(get("frozen_basis_orthogonality")),
        "certificate_usable": int(get("certificate_usable")),
        "unresolved": int(get("unresolved")), "refined": int(get("refined")),
        "best_iteration": int(get("best_iteration")),
        "stagnation_steps": int(get("stagnation_steps")),
        "best_residual": float(get("best_residual")),
    }


class Reports:
    """The per-matrix reports of one batched solve.

    A sequence of dicts to the caller (``reports[i]["status"]``, iteration),
    stored as one numpy record array: building a dict per matrix costs more
    host time than the solve of a small matrix, so it happens only for the
    entries that are read.  Whole-batch fields are numpy arrays:
    ``status``, ``converged``, ``sweeps``, ``corrections``."""

    def __init__(self, records):
        self.records = records
        self.fallback_from = {}

    def __len__(self):
        return len(self.records)

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[k] for k in range(*i.indices(len(self)))]
        i = int(i)
        if i < 0:
            i += len(self)
        d = report_dict(self.records[i])
        if i in self.fallback_from:
            d["fallback_from"] = self.fallback_from[i]
        return d

    def __iter__(self):
        return (self[i] for i in range(len(self)))

    @property
    def status(self):
        return self.records["status"]

    @property
    def converged(self):
        return self.records["status"] == CONVERGED

    @property
    def sweeps(self):
        return self.records["sweeps"]

    @property
    def corrections(self):
        return self.records["corrections"]

    def replace(self, i, other, k):
        """Entry i becomes entry k of ``other`` (a re-solve of the matrix);
        the replaced report is kept as its ``fallback_from``."""
        self.fallback_from[int(i)] = self[int(i)]
        self.records[int(i)] = other.records[int(k)]


# ---------------------------------------------------------------- runtime
class Runtime:
    """HIP backend runtime (MetalRuntime's role, src/metal_runtime.hpp:69-176)."""

    def __init__(self, device: int = 0, stream="current"):
        """stream: "current" enqueues on torch's current stream (ordered with
        the torch operations that fill and read the buffers), None on the
        runtime's own stream, else a torch stream or HIP stream address."""
        h = _p()
        check(lib.msq_runtime_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device
        if stream == "current":
            import torch
            stream = torch.cuda.current_stream(device)
        self.set_stream(stream)

    def close(self):
        if self.handle:
            check(lib.msq_runtime_destroy(self.handle))
            self.handle = None

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                lib.msq_runtime_destroy(self.handle)
        except Exception:
            pass

    def set_stream(self, stream=None):
        """Enqueue on a torch stream (or HIP stream address, 0 = the null
        stream); None: the runtime's own stream."""
        if stream is None:
            addr = OWN_STREAM
        else:
            addr = stream if isinstance(stream, int) else stream.cuda_stream
        check(lib.msq_runtime_set_stream(self.handle, addr))

    def synchronize(self):
        check(lib.msq_synchronize(self.handle))

    def autotune(self, max_order=256):
        """Times tile edges 32 / 64 / 128 per product type and order class and
        keeps the fastest; returns the (5, 9) table of edges (0: default)."""
        table = (ctypes.c_uint8 * (len(TILE_TYPES) * len(TILE_CLASSES)))()
        check(lib.msq_autotune(self.handle, max_order, table))
        return np.array(table[:], dtype=np.uint8).reshape(len(TILE_TYPES), len(TILE_CLASSES))

    def set_tile_table(self, table=None):
        if table is None:
            check(lib.msq_set_tile_table(self.handle, None))
            return
        t = np.ascontiguousarray(table, dtype=np.uint8).reshape(-1)
        check(lib.msq_set_tile_table(self.handle, t.ctypes.data))

    def set_graphs(self, enable=True):
        """Record each batch signature's solve in a HIP graph and replay it."""
        check(lib.msq_runtime_set_graphs(self.handle, 1 if enable else 0))

    def set_rounding(self, mode=ROUNDING_NEAREST, seed=0):
        """MetalRuntime's rounding mode and seed (nearest or stochastic)."""
        check(lib.msq_runtime_set_rounding(self.handle, mode, seed))

    def set_jacobi_method(self, method=JACOBI_BLOCKED_MFMA):
        """Eigensolver of the orders up to 256: JACOBI_SCALAR (the reference's
        rotations on the padded circle schedule) or JACOBI_BLOCKED_MFMA
        (8-index blocks, pair subproblems, FP32 MFMA products)."""
        check(lib.msq_runtime_set_jacobi_method(self.handle, method))

    # codecs -----------------------------------------------------------
    def quantize_stream(self, precision, src, values, scales=None, stream=0):
        """quantize with the quantizeAndDecode rounding stream; the log-spectrum
        precisions take their metadata plane as `scales`."""
        check(lib.msq_quantize_stream(self.handle, precision, _addr(src), _addr(values),
                                      _addr(scales), src.numel(), stream))

    def quantize(self, precision, src, values, scales=None):
        check(lib.msq_quantize(self.handle, precision, _addr(src), _addr(values),
                               _addr(scales), src.numel()))

    def decode(self, precision, values, scales, dst):
        check(lib.msq_decode(self.handle, precision, _addr(values), _addr(scales),
                             _addr(dst), dst.numel()))

    def quantize_and_decode(self, precision, src, values, scales, decoded):
        check(lib.msq_quantize_and_decode(self.handle, precision, _addr(src), _addr(values),
                                          _addr(scales), _addr(decoded), src.numel()))

    def quantize_fp8_scaled(self, src, values, sigma):
        check(lib.msq_quantize_fp8_scaled(self.handle, _addr(src), _addr(values),
                                          _addr(sigma), src.numel()))

    def decode_fp8_scaled(self, values, sigma, dst):
        check(lib.msq_decode_fp8_scaled(self.handle, _addr(values), _addr(sigma),
                                        _addr(dst), dst.numel()))

    # solver pieces ------------------------------------------------------
    def jacobi_batched(self, orders, matrices, vectors, diagonal, roots, options):
        opts = options if isinstance(options, (list, tuple)) else [options]
        arr = (JacobiOptions * len(opts))(*opts)
        reports = (JacobiReport * len(orders))()
        check(lib.msq_jacobi_batched(self.handle, len(orders), _u32_array(orders),
                                     _addr(matrices), _addr(vectors), _addr(diagonal),
                                     _addr(roots), arr, len(opts), reports))
        return list(reports)

    def jacobi_batched_noalloc(self, orders, matrices, vectors, diagonal, roots, options):
        """jacobi_batched through the allocation-free entry point (orders <= 256,
        at most NOALLOC_MAX_MATRICES matrices)."""
        opts = options if isinstance(options, (list, tuple)) else [options]
        arr = (JacobiOptions * len(opts))(*opts)
        reports = (JacobiReport * len(orders))()
        check(lib.msq_jacobi_batched_noalloc(self.handle, len(orders), _u32_array(orders),
                                             _addr(matrices), _addr(vectors), _addr(diagonal),
                                             _addr(roots), arr, len(opts), reports))
        return list(reports)

    def reconstruct(self, precision, orders, vectors, roots, out):
        check(lib.msq_reconstruct(self.handle, precision, len(orders), _u32_array(orders),
                                  _addr(vectors), _addr(roots), _addr(out)))

    def residual_fp64(self, orders, matrices, root, residual, residual_f32):
        rel = (ctypes.c_double * len(orders))()
        check(lib.msq_residual_fp64(self.handle, len(orders), _u32_array(orders),
                                    _addr(matrices), _addr(root), _addr(residual),
                                    _addr(residual_f32), rel))
        return np.array(rel[:])

    def correction(self, precision, orders, vectors, roots, residual_f32, out):
        check(lib.msq_correction(self.handle, precision, len(orders), _u32_array(orders),
                                 _addr(vectors), _addr(roots), _addr(residual_f32), _addr(out)))

    def refine_native_staged(self, precision, orders, vectors, roots, residual_f32, out):
        """refineNativeStaged: one native-quantized correction, decoded into out."""
        check(lib.msq_refine_native_staged(self.handle, precision, len(orders),
                                           _u32_array(orders), _addr(vectors), _addr(roots),
                                           _addr(residual_f32), _addr(out)))

    def orthogonality(self, orders, vectors):
        res = (ctypes.c_double * len(orders))()
        check(lib.msq_orthogonality(self.handle, len(orders), _u32_array(orders),
                                    _addr(vectors), res))
        return np.array(res[:])

    def solve_batched(self, orders, matrices, roots, options, reports=True,
                      initial_roots=None, vectors=None, root_eigenvalues=None):
        """Reports: a ``Reports`` sequence (one dict-like entry per matrix).
        initial_roots / vectors / root_eigenvalues: optional device outputs
        (msq_solve_batched_ex)."""
        if isinstance(options, np.ndarray):  # options_array records
            if options.dtype != OPTIONS_RECORD:
                raise TypeError("per-matrix options must be an options_array")
            rec = np.ascontiguousarray(options)
            arr, nopts = rec.ctypes.data_as(ctypes.POINTER(SolveOptions)), len(rec)
        else:
            opts = options if isinstance(options, (list, tuple)) else [options]
            arr, nopts = (SolveOptions * len(opts))(*opts), len(opts)
        # The library writes every field of every record (no zeroing here: 1.3 KB
        # a record).
        recs = np.empty(len(orders), dtype=REPORT_DTYPE) if reports else None
        outs = SolveOutputs(_addr(roots), _addr(initial_roots), _addr(vectors),
                            _addr(root_eigenvalues))
        check(lib.msq_solve_batched_ex(self.handle, len(orders), _u32_array(orders),
                                       _addr(matrices), ctypes.byref(outs), arr, nopts,
                                       recs.ctypes.data_as(ctypes.POINTER(SolveReport))
                                       if reports else None))
        return Reports(recs) if reports else None

    def last_factors(self):
        v = _p()
        s = _p()
        check(lib.msq_last_factors(self.handle, ctypes.byref(v), ctypes.byref(s)))
        return v.value, s.value

    # tier selector ------------------------------------------------------
    def matrix_features(self, orders, matrices, features, members=None):
        """Features of every matrix, or of the batch positions in `members`
        only (their rows of `features`; the other rows untouched)."""
        if members is None:
            check(lib.msq_matrix_features(self.handle, len(orders), _u32_array(orders),
                                          _addr(matrices), _addr(features)))
        else:
            check(lib.msq_matrix_features_subset(self.handle, len(orders), _u32_array(orders),
                                                 _addr(matrices), len(members),
                                                 _u32_array(members), _addr(features)))

    def mlp_forward(self, params, x, logits, h1=None, h2=None):
        batch = x.numel() // MLP_INPUTS
        check(lib.msq_mlp_forward(self.handle, _addr(params), _addr(x), batch, _addr(logits),
                                  _addr(h1), _addr(h2)))

    def mlp_forward_bf16(self, params16, x, logits):
        """params16: device uint8/int16 tensor of MLP_PARAMS BF16 codes."""
        batch = x.numel() // MLP_INPUTS
        check(lib.msq_mlp_forward_bf16(self.handle, _addr(params16), _addr(x), batch,
                                       _addr(logits)))

    def mlp_backward(self, params, x, h1, h2, logits, targets, mask, mask_total, grads,
                     losses=None):
        batch = x.numel() // MLP_INPUTS
        check(lib.msq_mlp_backward(self.handle, _addr(params), _addr(x), _addr(h1), _addr(h2),
                                   _addr(logits), _addr(targets), _addr(mask), batch,
                                   float(mask_total), _addr(grads), _addr(losses)))

    def mlp_adam(self, params, grads, m1, m2, lr, step, beta1=0.9, beta2=0.999, eps=1e-8,
                 decay=0.0):
        check(lib.msq_mlp_adam(self.handle, _addr(params), _addr(grads), _addr(m1), _addr(m2),
                               params.numel(), lr, beta1, beta2, eps, decay, step))
