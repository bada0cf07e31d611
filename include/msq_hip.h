/*
 * msq_hip.h -- C ABI of libmsq_hip.so, the MI355X (gfx950) backend of the
 * mixed-precision SPD square root X = A^(1/2) = V diag(sqrt(lambda)) V^T with
 * frozen-factor Sylvester correction.
 *
 * The reference exposes this path as a C++ class, MetalRuntime
 * (src/metal_runtime.hpp:69-176), driven by solveGpuResidentSpd and
 * runMetalCase (src/metal_solver.cpp:92-604).  A HIP backend is selected like
 * the Metal and Accelerate ones (SolverBackend, src/precision.hpp:130).  This
 * header is what a binding of that backend calls: plain pointers and sizes,
 * int status codes, no C++ or torch types.  The C++ mirror of MetalRuntime
 * (mixed-mlx_amd/csrc/hip_runtime.hpp) is written on top of it.
 *
 * Conventions
 *  - Matrices are column-major.  Batched calls use the batch layout of
 *    msq_batch_layout: matrix i occupies an ld_i x ld_i tile at element offset
 *    offsets[i], ld_i = order_i rounded up to a multiple of 32, zero outside
 *    the order; per-matrix vectors (roots) occupy ld_i entries at
 *    vector_offsets[i].
 *  - Device pointers are plain addresses in the runtime's device memory; all
 *    work is enqueued on the runtime's stream and every call returns after
 *    its work completed unless its comment says otherwise.
 *  - Status: 0 on success; msq_last_error() describes the last failure of the
 *    calling thread.
 *  - Calls on one runtime are not reentrant (the reference's MetalRuntime is
 *    single-threaded too, src/metal_runtime.cpp:1050-1058).
 */
#ifndef MSQ_HIP_H
#define MSQ_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSQ_ABI_VERSION 6

enum msq_status {
  MSQ_OK = 0,
  MSQ_ERROR_INVALID_ARGUMENT = 1,
  MSQ_ERROR_HIP = 2,
  MSQ_ERROR_OUT_OF_MEMORY = 3,
  MSQ_ERROR_UNSUPPORTED = 4,
};

/* Precision, in the order and meaning of the reference enum
 * (src/precision.hpp:9-20, spec table :64-90). */
enum msq_precision {
  MSQ_FP32 = 0,
  MSQ_BF16 = 1,
  MSQ_FP8_E4M3FN = 2,
  MSQ_FP8_E5M2 = 3,
  MSQ_MXFP8_E4M3FN = 4,
  MSQ_MXFP8_E5M2 = 5,
  MSQ_FP4_E2M1 = 6,
  MSQ_MXFP4_E2M1 = 7,
  MSQ_FP8_E4M3FN_LOG_S = 8,
  MSQ_FP4_E2M1_LOG_S = 9,
  /* No reference counterpart: FP8 E4M3FN with one per-tensor amax scale
   * sigma = 448 / amax (the reference's FP8 is unscaled). */
  MSQ_FP8_E4M3FN_SCALED = 16,
};

/* RoundingMode (src/precision.hpp:22): how quantization rounds. */
enum msq_rounding_mode {
  MSQ_ROUNDING_NEAREST = 0,
  MSQ_ROUNDING_STOCHASTIC = 1,
};

typedef struct msq_runtime msq_runtime;

/* Solver entry state of one matrix (solve status): the reference's
 * termination reasons (src/metal_solver.cpp:922-934) plus two of this
 * backend's. */
enum msq_solve_status {
  MSQ_SOLVE_RUNNING = 0,
  MSQ_SOLVE_CONVERGED = 1,
  MSQ_SOLVE_BUDGET_EXHAUSTED = 2,
  MSQ_SOLVE_DIVERGED = 3,
  MSQ_SOLVE_NONFINITE = 4,
  /* a Jacobi eigenvalue that is not positive and finite */
  MSQ_SOLVE_INVALID_SPECTRUM = 5,
  /* MSQ_STOPPING_STAGNATION stopped the corrections (the best iterate is
   * returned) */
  MSQ_SOLVE_STAGNATED = 6,
  /* The residual met its target, but the smallest Jacobi eigenvalue is below
   * what the FP32 eigensolve resolves (msq_solve_options.resolution), so the
   * forward-error estimate that stopped the corrections is not reliable and
   * the error may exceed the target.  refine_unresolved = 1 resolves it. */
  MSQ_SOLVE_UNRESOLVED = 7,
};

/* KernelPath, CorrectionMode, NativeStoppingPolicy and NativeDispatchPolicy
 * (src/precision.hpp:126-146), in the reference's order. */
enum msq_kernel_path { MSQ_KERNEL_PATH_LEGACY = 0, MSQ_KERNEL_PATH_TILED = 1 };
enum msq_correction_mode {
  MSQ_CORRECTION_FP32_REORTHOGONALIZED = 0,
  MSQ_CORRECTION_NATIVE_QUANTIZED = 1,
};
enum msq_stopping_policy { MSQ_STOPPING_FIXED_BUDGET = 0, MSQ_STOPPING_STAGNATION = 1 };
enum msq_dispatch_policy { MSQ_DISPATCH_PER_STAGE = 0, MSQ_DISPATCH_FUSED = 1 };

#define MSQ_JACOBI_HISTORY 32
#define MSQ_SOLVE_HISTORY 65

/* Jacobi options of one matrix.  The reference runs a fixed 8 sweeps
 * (src/metal_solver.cpp:210-213); fixed_sweeps = 8 reproduces that.  With
 * fixed_sweeps = 0 the sweeps stop at the first sweep whose FP64 off-diagonal
 * norm off(A) satisfies off <= off_tol ||A||_F, or off <= off_tol_lambda *
 * min diag(A), or (stagnation) off > stagnation * previous off while
 * off <= floor_tol ||A||_F, or once the frozen-factor correction is predicted
 * to contract by correction_rate per step, or after max_sweeps.
 *
 * correction_rate (0: off; no counterpart in the reference): with F the
 * off-diagonal part of the rotated A and s_i = sqrt(a_ii), one Sylvester
 * correction in the frozen basis contracts the root's error, to first order,
 * by at most ||G||_F / min_i s_i with G_ij = F_ij / (s_i + s_j).  Stopping at
 * that bound leaves the rest of the accuracy to the corrections. */
typedef struct msq_jacobi_options {
  uint32_t fixed_sweeps;
  uint32_t max_sweeps;
  float off_tol;
  float off_tol_lambda;
  float stagnation;
  float floor_tol;
  float correction_rate;
} msq_jacobi_options;

typedef struct msq_jacobi_report {
  uint32_t sweeps;
  uint32_t invalid_spectrum; /* a diagonal entry <= 0 or non-finite */
  double off_norm;           /* off(A) after the last sweep */
  double matrix_norm;        /* ||A||_F (binary64) */
  float min_diagonal;
  float off_history[MSQ_JACOBI_HISTORY]; /* off(A) / ||A||_F per sweep */
  float correction_rate;                 /* ||G||_F / min s after the last sweep */
} msq_jacobi_report;

/* Options of one matrix of msq_solve_batched: MetalSolverConfiguration and
 * MetalProblem (src/metal_solver.hpp:12-31) for this backend.
 * msq_solve_options_init fills the reference's defaults.
 *
 * Products: reconstruction X0 = G G^T with G = V diag(sqrt(s)) rounded to
 * basis_precision, s the roots rounded to spectrum_precision (both decided per
 * matrix: a basis whose ||V_t^T V_t - I||_F / sqrt(n) exceeds basis_tolerance,
 * or a spectrum with a decoded root that is not positive and finite, falls
 * back to FP32, as src/metal_solver.cpp:217-263 does); every correction product
 * W = V^T R, Y = (W V) o H, W2 = V Y, E = W2 V^T in correction_precision, on
 * the FP32 basis and roots.  MSQ_FP32 for all three is the reference's
 * GPU-resident FP32 path.  basis / correction precisions: MSQ_FP32, MSQ_BF16,
 * MSQ_FP8_E4M3FN_SCALED, MSQ_MXFP4_E2M1; spectrum: every msq_precision but
 * MSQ_FP8_E4M3FN_SCALED.
 *
 * correction_mode MSQ_CORRECTION_NATIVE_QUANTIZED is the reference's native
 * path (refineNativeStaged, src/metal_runtime.cpp:312-411): basis, roots,
 * residual and every intermediate product are stored in basis_precision (FP8
 * E4M3FN / E5M2, MXFP8, FP4 E2M1, MXFP4, the reference's unscaled codecs on
 * its linear n x n layout), and the update is X = sym(X + E).  As in the
 * reference (validateCorrectionMode, src/configuration.cpp:322-346) it needs
 * basis_precision == spectrum_precision == correction_precision, one of those
 * six, and kernel_path MSQ_KERNEL_PATH_LEGACY; stopping_policy
 * MSQ_STOPPING_STAGNATION and dispatch_policy MSQ_DISPATCH_PER_STAGE need it.
 * On this backend both kernel paths and both dispatch policies run the same
 * device kernels on the runtime's stream (the reference's Legacy / Tiled and
 * PerStage / Fused differ in kernel granularity and command buffers, not in
 * the values computed).
 *
 * Corrections stop when the relative FP64 residual ||A - X^2||_F / ||A||_F <=
 * residual_tolerance or the forward-error estimate ||R||_F / (2 s_min ||X||_F)
 * <= forward_tolerance, after at least min_corrections and at most
 * max_corrections (< MSQ_SOLVE_HISTORY); with MSQ_STOPPING_STAGNATION also
 * after stagnation_window corrections in a row that improve the best residual
 * by no more than minimum_relative_improvement (src/metal_solver.cpp:889-902),
 * and the best iterate is returned. */
typedef struct msq_solve_options {
  int32_t basis_precision;
  int32_t correction_precision;
  msq_jacobi_options jacobi;
  uint32_t min_corrections;
  uint32_t max_corrections;
  double residual_tolerance;
  double forward_tolerance;
  int32_t spectrum_precision;
  int32_t kernel_path;      /* enum msq_kernel_path */
  int32_t correction_mode;  /* enum msq_correction_mode */
  int32_t stopping_policy;  /* enum msq_stopping_policy */
  int32_t dispatch_policy;  /* enum msq_dispatch_policy */
  uint32_t stagnation_window;
  double minimum_relative_improvement;
  /* Basis acceptance tolerance; 0: the tier's basisOrthogonalityTolerance
   * (src/configuration.cpp:109-118: BF16 5e-3, 8-bit 0.15, 4-bit 0.5). */
  double basis_tolerance;
  /* No reference counterpart.  A matrix whose smallest Jacobi eigenvalue is
   * below resolution * (largest) is unresolved: its forward-error estimate is
   * not reliable (0: no test).  With refine_unresolved = 1 such a matrix gets
   * a second Jacobi pass on V^T A V formed in binary64, whose small
   * eigenvalues FP32 resolves to their own relative accuracy, and V <- V W;
   * otherwise a matrix that stops unresolved reports MSQ_SOLVE_UNRESOLVED. */
  float resolution;
  uint32_t refine_unresolved;
  /* 1: measure ||V^T V - I||_F / sqrt(n) of the FP32 basis for the certificate
   * (runMetalCase measures it, src/metal_solver.cpp:511-524). */
  uint32_t certificate;
} msq_solve_options;

typedef struct msq_solve_report {
  uint32_t order;
  uint32_t sweeps;
  uint32_t corrections;
  uint32_t status; /* enum msq_solve_status */
  double matrix_norm;
  double min_root;
  double off_norm;
  double final_residual;
  double final_forward_estimate;
  double residuals[MSQ_SOLVE_HISTORY];         /* X0, then each correction */
  double forward_estimates[MSQ_SOLVE_HISTORY];
  float off_history[MSQ_JACOBI_HISTORY];
  float correction_rate; /* Jacobi's ||G||_F / min s (see msq_jacobi_options) */
  /* Basis tier acceptance (src/metal_solver.cpp:217-236): ||V_t^T V_t - I||_F /
   * sqrt(n) of V rounded to the basis tier (0 for an FP32 basis). */
  float basis_orthogonality;
  /* Tiers rejected for FP32: basis_rejected + spectrum_rejected. */
  uint32_t precision_escalations;
  uint32_t basis_rejected;
  uint32_t spectrum_rejected;
  /* Frozen-factor certificate (src/metal_solver.cpp:511-524):
   * 1 / (2 min s) over the frozen roots X0 was formed with; the basis
   * orthogonality it rests on (the selected tier's, or the FP32 basis' when
   * options.certificate = 1, else NaN); usable when that is below 1e-3.  The
   * contraction bound 3 ||X0 - X*||_F / (2 min s) needs the FP64 root X*: the
   * caller forms it from msq_solve_outputs.initial_roots. */
  double frozen_sylvester_inverse_norm;
  double frozen_basis_orthogonality;
  uint32_t certificate_usable;
  /* Smallest Jacobi eigenvalue below options.resolution (after the second
   * pass, when it ran); refined: the second pass ran (its sweeps are in
   * `sweeps` too). */
  uint32_t unresolved;
  uint32_t refined;
  /* Best residual iterate (MSQ_STOPPING_STAGNATION returns it). */
  uint32_t best_iteration;
  uint32_t stagnation_steps;
  uint32_t reserved;
  double best_residual;
} msq_solve_report;

/* Outputs of msq_solve_batched_ex: `roots` (X, binary64, batch layout) is
 * required; the others may be NULL.  initial_roots: X0 (binary64, batch
 * layout) before the first correction; vectors: the FP32 eigenvectors the
 * corrections used (float, batch layout); root_eigenvalues: their roots s =
 * sqrt(diag) (float, vector layout). */
typedef struct msq_solve_outputs {
  double *roots;
  double *initial_roots;
  float *vectors;
  float *root_eigenvalues;
} msq_solve_outputs;

/* ---------------------------------------------------------------- runtime */

/* Thread-local description of the last failure. */
const char *msq_last_error(void);
int msq_abi_version(void);

/* MetalRuntime::MetalRuntime (src/metal_runtime.cpp:41-126): one device, one
 * stream (created here, default priority). */
int msq_runtime_create(int device, msq_runtime **out);
/* ~MetalRuntime. */
int msq_runtime_destroy(msq_runtime *runtime);
/* Enqueue on a caller's stream instead: any hipStream_t, NULL being HIP's
 * null stream; MSQ_OWN_STREAM restores the runtime's own stream.  Work that a
 * call spreads over the runtime's internal streams is forked from and joined
 * back into this stream with events. */
#define MSQ_OWN_STREAM ((void *)(~(uintptr_t)0))
int msq_runtime_set_stream(msq_runtime *runtime, void *hip_stream);
int msq_synchronize(msq_runtime *runtime);
/* The rounding mode and seed of MetalRuntime's constructor
 * (src/metal_runtime.cpp:41-126; default nearest, seed 0).  With
 * MSQ_ROUNDING_STOCHASTIC every quantization of this runtime (the codec calls
 * and a solve's basis and spectrum tiers) rounds each element up or down with
 * probability given by its distance to the two neighbouring codes, decided by
 * a hash of (element index, seed, rounding stream)
 * (metal_mixed_sqrt.metal:394-405): results are reproducible for a seed. */
int msq_runtime_set_rounding(msq_runtime *runtime, int mode, uint32_t seed);
/* No counterpart in the reference.  enable != 0: msq_solve_batched records
 * the device work of each batch signature (orders, options, matrices and
 * roots pointers) in a HIP graph -- the first call runs eagerly, the second
 * is captured, later ones replay it (the plans and uploads stay in an arena
 * of the signature's own).  The capture takes two graphs: the work up to
 * the last correction step that the eager call needed, and the steps after
 * it, which a replay launches only if the host, after synchronising, finds
 * a matrix still running (solves under the stagnation policy keep one
 * graph).  Batches with orders above 256 always run eagerly.  At most 4
 * signatures are kept; enable = 0 releases them. */
int msq_runtime_set_graphs(msq_runtime *runtime, int enable);

/* Eigensolver of the orders up to 256 (no counterpart in the reference, whose
 * Jacobi is one scalar kernel for every order):
 *  - MSQ_JACOBI_SCALAR: one rotation at a time per index pair, the reference's
 *    element formulas on the padded circle schedule (orders 129..256 in one
 *    persistent launch, smaller ones one launch per padded order);
 *  - MSQ_JACOBI_BLOCKED_MFMA: blocks of 8 indices paired on the circle, each
 *    16 x 16 pair subproblem swept once by scalar rotations, the rest of the
 *    round applied as FP32 MFMA products (V in registers); same stop rules,
 *    a sweep being one pass over the block pairs; orders 17..256 in one
 *    persistent launch (orders up to 16 as with MSQ_JACOBI_SCALAR).
 * Orders above 256 (the blocked 64 x 64 solver) are not affected.  Changing the method
 * releases the runtime's HIP graphs.  Default: MSQ_JACOBI_BLOCKED_MFMA. */
enum msq_jacobi_method { MSQ_JACOBI_SCALAR = 0, MSQ_JACOBI_BLOCKED_MFMA = 1 };
int msq_runtime_set_jacobi_method(msq_runtime *runtime, int method);

/* makeSharedBuffer / releaseBuffers (src/metal_runtime.cpp:1426-1432).  No
 * unified memory here: device allocations plus explicit copies. */
int msq_alloc(msq_runtime *runtime, size_t bytes, void **device_ptr);
int msq_free(msq_runtime *runtime, void *device_ptr);
int msq_copy_to_device(msq_runtime *runtime, void *dst, const void *src, size_t bytes);
int msq_copy_to_host(msq_runtime *runtime, void *dst, const void *src, size_t bytes);
int msq_memset(msq_runtime *runtime, void *dst, int value, size_t bytes);

/* storageBytes / scaleBytes (src/configuration.cpp:188-213). */
size_t msq_storage_bytes(int precision, size_t elements);
size_t msq_scale_bytes(int precision, size_t elements);
/* spectrumMetadataBytes (src/storage_codecs.cpp:387-398): two floats per block
 * of 8 elements for the log-spectrum precisions, else msq_scale_bytes. */
size_t msq_spectrum_metadata_bytes(int precision, size_t elements);

/* Batch layout: ld_i = roundup(order_i, 32); offsets[i] = sum_{j<i} ld_j^2;
 * vector_offsets[i] = sum_{j<i} ld_j.  Totals in elements.  Every batched
 * call validates its orders this way: count >= 1 and 1 <= order_i <= 65504,
 * else MSQ_ERROR_INVALID_ARGUMENT. */
int msq_batch_layout(uint32_t count, const uint32_t *orders, uint32_t *lds,
                     uint64_t *offsets, uint64_t *vector_offsets,
                     uint64_t *total_elements, uint64_t *total_vector_elements);

/* ----------------------------------------------------------------- codecs */

/* The quantize half of quantizeAndDecode (src/metal_runtime.cpp:672-740),
 * reference linear layout (MX blocks of 32 consecutive elements).  Supports
 * MSQ_BF16, MSQ_FP8_E4M3FN, MSQ_FP8_E5M2, MSQ_MXFP8_E4M3FN, MSQ_MXFP8_E5M2,
 * MSQ_FP4_E2M1, MSQ_MXFP4_E2M1; bit-exact with the reference host codecs
 * (src/storage_codecs.cpp:16-198).  scales may be NULL for scalar formats. */
int msq_quantize(msq_runtime *runtime, int precision, const float *src, void *values,
                 void *scales, size_t elements);
/* msq_quantize with the quantizeAndDecode rounding stream argument
 * (src/metal_runtime.cpp:672-740): stochastic rounding (msq_runtime_set_rounding)
 * hashes (element index, seed, stream).  Also takes the log-spectrum
 * precisions MSQ_FP8_E4M3FN_LOG_S and MSQ_FP4_E2M1_LOG_S, whose `scales` is the
 * metadata plane (msq_spectrum_metadata_bytes: per block of 8 elements the
 * floats min log2 v and the log2 span; log_spectrum_metadata and
 * quantize_*_log_s, metal_mixed_sqrt.metal:1689-1799). */
int msq_quantize_stream(msq_runtime *runtime, int precision, const float *src, void *values,
                        void *scales, size_t elements, uint32_t rounding_stream);
/* The decode half (decode_*_to_f32, metal_mixed_sqrt.metal:1661-1685,
 * :1801-1817; for the log-spectrum precisions `scales` is the metadata). */
int msq_decode(msq_runtime *runtime, int precision, const void *values,
               const void *scales, float *dst, size_t elements);
/* quantizeAndDecode as one call. */
int msq_quantize_and_decode(msq_runtime *runtime, int precision, const float *src,
                            void *values, void *scales, float *decoded, size_t elements);
/* MSQ_FP8_E4M3FN_SCALED: sigma (one device float) = 448 / amax, codes =
 * E4M3FN(src * sigma); decode = E4M3FN^-1(code) / sigma. */
int msq_quantize_fp8_scaled(msq_runtime *runtime, const float *src, void *values,
                            float *sigma, size_t elements);
int msq_decode_fp8_scaled(msq_runtime *runtime, const void *values, const float *sigma,
                          float *dst, size_t elements);

/* ----------------------------------------------------------------- Jacobi */

/* symmetricJacobi (src/metal_runtime.cpp:584-646) for a batch: the FP64
 * matrices (batch layout) are rounded to float and diagonalised; vectors
 * (float, batch layout) receive V, diagonal and roots (float, vector layout)
 * receive diag(V^T A V) and its square roots.  options has 1 or count
 * entries; reports (host, count entries) may be NULL.  Orders up to 256 run
 * one workgroup per matrix, all sweeps in LDS (msq_runtime_set_jacobi_method
 * picks the eigensolver and which orders share one persistent launch); larger
 * orders run the blocked solver (64 x 64 pair subproblems, rotations applied
 * as MFMA products), the matrices of one padded order advancing together. */
int msq_jacobi_batched(msq_runtime *runtime, uint32_t count, const uint32_t *orders,
                       const double *matrices, float *vectors, float *diagonal,
                       float *roots, const msq_jacobi_options *options,
                       uint32_t options_count, msq_jacobi_report *reports);

/* No counterpart in the reference.  msq_jacobi_batched for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  Same arguments, results and reports as
 * msq_jacobi_batched (bit for bit), for orders 1..256 and at most
 * MSQ_NOALLOC_MAX_MATRICES matrices per call; larger orders or counts are
 * rejected with MSQ_ERROR_INVALID_ARGUMENT.  It allocates nothing: its launch
 * plan and per-matrix status live in one static device array of
 * MSQ_NOALLOC_SCRATCH_BYTES per GPU, which the library holds until the process
 * ends.  Every call on a GPU shares that GPU's array, so calls that overlap on
 * one GPU (two runtimes on it, or two host threads) race; calls in sequence do
 * not (each call rewrites the entries it reads). */
#define MSQ_NOALLOC_MAX_MATRICES 4096
#define MSQ_NOALLOC_SCRATCH_BYTES 1703936
int msq_jacobi_batched_noalloc(msq_runtime *runtime, uint32_t count, const uint32_t *orders,
                               const double *matrices, float *vectors, float *diagonal,
                               float *roots, const msq_jacobi_options *options,
                               uint32_t options_count, msq_jacobi_report *reports);

/* --------------------------------------------------------------- products */

/* reconstructTiled / reconstructLowFactorTiled (src/metal_runtime.cpp:429-534)
 * for a batch: X0 = G G^T, G = V diag(sqrt(roots)) rounded to precision,
 * FP32 accumulate; root (float, batch layout). */
int msq_reconstruct(msq_runtime *runtime, int precision, uint32_t count,
                    const uint32_t *orders, const float *vectors, const float *roots,
                    float *root);

/* computeWorkingResidual (src/metal_runtime.cpp:789-834) in binary64:
 * R = A - X X for symmetric X; residual (double) and residual_f32 (float)
 * in batch layout; relative_norms (host, count) = ||R||_F / ||A||_F. */
int msq_residual_fp64(msq_runtime *runtime, uint32_t count, const uint32_t *orders,
                      const double *matrices, const double *root, double *residual,
                      float *residual_f32, double *relative_norms);

/* refineTiled / refineLowFactorTiled (src/metal_runtime.cpp:461-582): one
 * frozen-factor correction E = V [ (V^T R V) o H ] V^T with H_ij =
 * 1 / (s_i + s_j), products in precision, FP32 accumulate; correction (float,
 * batch layout). */
int msq_correction(msq_runtime *runtime, int precision, uint32_t count,
                   const uint32_t *orders, const float *vectors, const float *roots,
                   const float *residual_f32, float *correction);

/* refineNativeStaged (src/metal_runtime.cpp:312-411): one native-quantized
 * correction.  V (and V^T, stored separately), the roots and the residual
 * (float, batch / vector layout) are stored in `precision` (FP8 E4M3FN / E5M2,
 * MXFP8 E4M3FN / E5M2, FP4 E2M1, MXFP4 E2M1; the reference's host codecs on
 * its linear n x n layout); then U = Q_t R, C = U Q, Y_ij = C_ij / (s_i + s_j),
 * V2 = Q Y, E = V2 Q_t, each formed in FP32 from decoded operands and stored
 * in the format; correction (float, batch layout) receives decoded E. */
int msq_refine_native_staged(msq_runtime *runtime, int precision, uint32_t count,
                             const uint32_t *orders, const float *vectors, const float *roots,
                             const float *residual_f32, float *correction);

/* measureF32Orthogonality (src/metal_runtime.cpp:742-776):
 * ||V^T V - I||_F / sqrt(n) per matrix into results (host, count). */
int msq_orthogonality(msq_runtime *runtime, uint32_t count, const uint32_t *orders,
                      const float *vectors, double *results);

/* ------------------------------------------------------------------ solve */

/* solveGpuResidentSpd (src/metal_solver.cpp:92-336) for a batch of uneven
 * SPD matrices: Jacobi, reconstruction, binary64 residual and corrections,
 * one fixed launch sequence on the device; each matrix stops on its own
 * (converged matrices skip the remaining launches).  matrices and roots are
 * device pointers in batch layout (binary64); options has 1 or count entries;
 * reports (host, count) may be NULL.  Returns MSQ_OK when the launches ran;
 * per-matrix outcomes are in reports[i].status. */
int msq_solve_batched(msq_runtime *runtime, uint32_t count, const uint32_t *orders,
                      const double *matrices, double *roots,
                      const msq_solve_options *options, uint32_t options_count,
                      msq_solve_report *reports);
/* msq_solve_batched with the outputs of an msq_solve_outputs record; in graph
 * mode the output pointers are part of the batch signature. */
int msq_solve_batched_ex(msq_runtime *runtime, uint32_t count, const uint32_t *orders,
                         const double *matrices, const msq_solve_outputs *outputs,
                         const msq_solve_options *options, uint32_t options_count,
                         msq_solve_report *reports);
/* The reference's defaults (MetalSolverConfiguration, MetalProblem and
 * src/metal_backend.cpp:105-113): FP32 tiers, Tiled, FP32-reorthogonalized,
 * fixed budget, Fused, stagnation window 4 and improvement 1e-3, Jacobi to
 * off(A) <= 1e-7 ||A||_F, at most 12 corrections, residual tolerance 1e-12. */
void msq_solve_options_init(msq_solve_options *options);

/* Eigenvectors and roots of the last msq_solve_batched call on this runtime
 * (device pointers into the runtime's scratch, float, batch / vector layout).
 * Valid until the next call on this runtime that uses scratch (any solve,
 * Jacobi, product or feature call) or that releases graphs; msq_solve_batched_ex
 * copies them to caller memory instead. */
int msq_last_factors(msq_runtime *runtime, const float **vectors, const float **roots);

/* -------------------------------------------------------------- autotuner */
/* No counterpart in the reference.  Times the batched GEMM kernel with tile
 * edges 32, 64 and 128 (HIP events, synthetic batches of about 2^25
 * elements) for each product type (FP64 residual, FP32, BF16, scaled FP8,
 * MXFP4) and each order class (ld 32, 64, ..., 256 up to max_order, and one
 * class above 256), and keeps the fastest edge per (type, class) in the
 * runtime: later plans use it.  table (host, MSQ_TILE_TABLE_BYTES, may be
 * NULL) receives the choices, row-major [type][class], 0 = default.  A tile
 * edge only partitions the output: results are bit-identical for every
 * choice. */
#define MSQ_TILE_TABLE_BYTES 45
int msq_autotune(msq_runtime *runtime, uint32_t max_order, uint8_t *table);
/* Loads a table from msq_autotune (NULL: back to the defaults). */
int msq_set_tile_table(msq_runtime *runtime, const uint8_t *table);

/* ------------------------------------------------------- learned selector */
/* No counterpart in the reference: the tier selector of the batched solver.
 * A small MLP (MSQ_MLP_INPUTS -> MSQ_MLP_HIDDEN -> MSQ_MLP_HIDDEN ->
 * MSQ_MLP_OUTPUTS, ReLU, float32) reads per-matrix features and predicts,
 * per solver configuration, whether it meets the accuracy target.  Parameter
 * vector (MSQ_MLP_PARAMS floats): W1[64][16], b1[64], W2[64][64], b2[64],
 * W3[8][64], b3[8], each W row-major (W[j][i]: input i into output j).
 * The four calls below only enqueue their work on the runtime's stream and
 * return before it completes (training issues thousands of them). */
#define MSQ_MLP_INPUTS 16
#define MSQ_MLP_HIDDEN 64
#define MSQ_MLP_OUTPUTS 8
#define MSQ_MLP_PARAMS 5768
#define MSQ_FEATURE_MAX_ORDER 512

/* Features of each SPD matrix (binary64, batch layout, orders <= 512) into
 * features (device, count x MSQ_MLP_INPUTS floats): order, power-iteration
 * lambda_max / lambda_min estimates (16 / 24 steps), Gershgorin bounds,
 * diagonal spread and dominance, norms. */
int msq_matrix_features(msq_runtime *runtime, uint32_t count, const uint32_t *orders,
                        const double *matrices, float *features);
/* msq_matrix_features for the batch positions in members[0 .. member_count)
 * only (each < count): row members[q] of features receives matrix
 * members[q]'s features, the other rows are left as they are.  The selector
 * estimates its batch gate from a sample this way. */
int msq_matrix_features_subset(msq_runtime *runtime, uint32_t count, const uint32_t *orders,
                               const double *matrices, uint32_t member_count,
                               const uint32_t *members, float *features);

/* logits (device, batch x MSQ_MLP_OUTPUTS) of the network at params for the
 * inputs x (device, batch x MSQ_MLP_INPUTS).  h1 / h2 (device, batch x
 * MSQ_MLP_HIDDEN, post-ReLU) are written when non-null, for the backward. */
int msq_mlp_forward(msq_runtime *runtime, const float *params, const float *x, uint32_t batch,
                    float *logits, float *h1, float *h2);

/* BF16 inference: logits (device, batch x MSQ_MLP_OUTPUTS, float) of the
 * network whose weights `params` (device, MSQ_MLP_PARAMS BF16 codes, e.g. from
 * msq_quantize with MSQ_BF16) are stored in BF16, with the inputs and the
 * hidden activations rounded to BF16 (nearest even, after the ReLU) and float
 * accumulation in index order.  No reference counterpart (the learned tier
 * selector is new). */
int msq_mlp_forward_bf16(msq_runtime *runtime, const uint16_t *params, const float *x,
                         uint32_t batch, float *logits);

/* Gradient (device, MSQ_MLP_PARAMS) of the masked binary cross-entropy
 *   L = sum_{b,k} mask_bk (softplus(z_bk) - y_bk z_bk) / mask_total
 * at the forward pass (x, h1, h2, logits) with targets y and mask (device,
 * batch x MSQ_MLP_OUTPUTS); per-sample losses (device, batch, unnormalised)
 * into losses when non-null.  Deterministic (fixed summation order). */
int msq_mlp_backward(msq_runtime *runtime, const float *params, const float *x, const float *h1,
                     const float *h2, const float *logits, const float *targets,
                     const float *mask, uint32_t batch, float mask_total, float *grads,
                     float *losses);

/* One Adam step (step >= 1, bias-corrected moments) with an L2 term added
 * to the gradient, g = grads + decay params; m1, m2 the first and second
 * moments (device, count floats). */
int msq_mlp_adam(msq_runtime *runtime, float *params, const float *grads, float *m1, float *m2,
                 uint32_t count, float lr, float beta1, float beta2, float eps, float decay,
                 uint32_t step);

#ifdef __cplusplus
}
#endif

#endif /* MSQ_HIP_H */
