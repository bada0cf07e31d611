// Kernel launch interface shared by the runtime and the kernel translation
// units.  Nothing here is exported from the library; the C ABI is
// include/msq_hip.h.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace msq {

// Element formats of the codec kernels, in the order of the reference's
// HostStorageCodec (src/precision.hpp:24-33).
enum class Format : int {
  Fp32 = 0,
  Bf16 = 1,
  Fp8E4M3Fn = 2,
  Fp8E5M2 = 3,
  Mxfp8E4M3Fn = 4,
  Mxfp8E5M2 = 5,
  Fp4E2M1 = 6,
  Mxfp4E2M1 = 7,
};

hipError_t launch_encode(Format fmt, const float* src, void* values, uint8_t* scales,
                         size_t n, hipStream_t stream);
hipError_t launch_decode(Format fmt, const void* values, const uint8_t* scales,
                         float* dst, size_t n, hipStream_t stream);
hipError_t launch_fp8_amax_sigma(const float* src, size_t n, uint32_t* amax_bits,
                                 float* sigma, hipStream_t stream);
hipError_t launch_encode_fp8_scaled(const float* src, uint8_t* values,
                                    const float* sigma, size_t n, hipStream_t stream);
hipError_t launch_decode_fp8_scaled(const uint8_t* values, const float* sigma,
                                    float* dst, size_t n, hipStream_t stream);
// Stochastic rounding of the linear formats (the hash of element index, seed
// and stream decides each rounding direction).
hipError_t launch_encode_stochastic(Format fmt, const float* src, void* values, uint8_t* scales,
                                   size_t n, uint32_t seed, uint32_t stream, hipStream_t s);
// Log-affine spectrum codec (E4M3FN or E2M1 codes, two floats of metadata per
// block of 8 elements), nearest or stochastic.
hipError_t launch_encode_log(bool e4m3, const float* src, uint8_t* values, float* metadata,
                             size_t n, bool stochastic, uint32_t seed, uint32_t stream,
                             hipStream_t s);
hipError_t launch_decode_log(bool e4m3, const uint8_t* values, const float* metadata, float* dst,
                             size_t n, hipStream_t s);

// ------------------------------------------------------------------ Jacobi
constexpr uint32_t kBatchedJacobiMaxOrder = 256;
constexpr uint32_t kJacobiHistory = 32;
constexpr uint32_t kJacobiInvalidSpectrum = 1u;

// One matrix of a batched Jacobi launch.  Offsets are in elements.
struct JacobiJob {
  uint64_t a_offset;       // binary64 input, column-major, pitch ld
  uint64_t v_offset;       // float eigenvectors, column-major, pitch ld
  uint64_t s_offset;       // float diagonal / roots, ld entries
  uint32_t order;
  uint32_t ld;
  uint32_t max_sweeps;     // cap when converging by tolerance
  uint32_t fixed_sweeps;   // >0: run exactly this many sweeps, no test
  float off_tol;           // stop when off(A) <= off_tol * ||A||_F ...
  float off_tol_lambda;    // ... or <= off_tol_lambda * min diag(A) (0: off)
  float stagnation;        // ... or off > stagnation * previous off and
  float floor_tol;         //     off <= floor_tol * ||A||_F (FP32 floor)
  uint32_t status_index;   // slot of this matrix in the status array
  float correction_rate;   // ... or ||G||_F / min s <= correction_rate (0: off)
};

struct JacobiStatus {
  uint32_t sweeps;
  uint32_t flags;
  double off_norm;
  double matrix_norm;
  float min_diagonal;
  float off_history[kJacobiHistory];  // off(A)/||A||_F after each sweep
  float correction_rate;              // ||G||_F / min s after the last sweep
};

// One launch covers matrices sharing the padded order (order rounded up to
// 16: the launch's block size and LDS size follow it).
uint32_t batched_jacobi_padded_order(uint32_t order);
int batched_jacobi_threads(uint32_t order, uint32_t count);
size_t batched_jacobi_lds_bytes(uint32_t order);
hipError_t launch_batched_jacobi(const JacobiJob* jobs, uint32_t count, uint32_t order,
                                 const double* a64, float* vectors, float* diagonal,
                                 float* roots, JacobiStatus* status, hipStream_t stream);
// Orders 129..256 in one persistent launch: `grid` workgroups (one per CU)
// take jobs[0, count) in order from `counter` (one uint32 of device memory,
// cleared here by a kernel before the launch; one launch at a time per
// counter).  Every job's order must be in [kPersistentMinOrder, 256]; the
// results equal launch_batched_jacobi's, bit for bit.
constexpr uint32_t kPersistentMinOrder = 129;
hipError_t launch_persistent_jacobi(const JacobiJob* jobs, uint32_t count, uint32_t* counter,
                                    int grid, const double* a64, float* vectors,
                                    float* diagonal, float* roots, JacobiStatus* status,
                                    hipStream_t stream);

// Orders 17..256 (padded from tile_jacobi_min_padded()) on the blocked LDS
// solver, one persistent launch like launch_persistent_jacobi: blocks of 8
// indices paired on the circle, 16 x 16 pair subproblems diagonalised by
// scalar rotations and the rest of each round applied as FP32 MFMA products,
// V in registers.  Same persistent launch shape, counter use, outputs and
// stop rules as launch_persistent_jacobi; a sweep is N / 8 - 1 block rounds.
// Smallest padded order of the blocked solver's persistent launch.
uint32_t tile_jacobi_min_padded();
hipError_t launch_tile_jacobi(const JacobiJob* jobs, uint32_t count, uint32_t* counter, int grid,
                              const double* a64, float* vectors, float* diagonal, float* roots,
                              JacobiStatus* status, hipStream_t stream);

// ------------------------------------------------------ learned selector
constexpr int kMlpInputs = 16;
constexpr int kMlpHidden = 64;
constexpr int kMlpOutputs = 8;
constexpr int kMlpParams = kMlpHidden * kMlpInputs + kMlpHidden + kMlpHidden * kMlpHidden +
                           kMlpHidden + kMlpOutputs * kMlpHidden + kMlpOutputs;  // 5768
constexpr int kMaxFeatureOrder = 512;
constexpr int kPowerIterationsMax = 16;  // lambda_max: power iteration on A
constexpr int kPowerIterationsMin = 24;  // lambda_min: on (c I - A), c > lambda_max

struct Segment;
// Features of the segments of one order range: rows = 64 (orders 1..64),
// 128 (65..128), 192 (129..192), 256 (193..256) or 0 (257..512); segments
// outside the range are skipped.  Row sg.index of `features` receives segment sg's 16 values.
hipError_t launch_features(int rows, const Segment* segs, uint32_t count, const double* a64,
                           float* features, hipStream_t stream);
int feature_rows_for_order(uint32_t order);
hipError_t launch_mlp_forward_bf16(const uint16_t* params, const float* x, int batch,
                                   float* logits, hipStream_t stream);
hipError_t launch_mlp_forward(const float* params, const float* x, int batch, float* logits,
                              float* h1, float* h2, hipStream_t stream);
hipError_t launch_mlp_backward(const float* params, const float* x, const float* h1,
                               const float* h2, const float* logits, const float* targets,
                               const float* mask, int batch, float mask_total, float* dz,
                               float* dh1, float* dh2, float* losses, float* grads,
                               hipStream_t stream);
hipError_t launch_mlp_adam(float* params, const float* grads, float* m1, float* m2, int count,
                           float lr, float beta1, float beta2, float eps, float decay, int step,
                           hipStream_t stream);

// Blocked Jacobi (orders above kBatchedJacobiMaxOrder): 32-wide blocks paired
// on the circle, 64 x 64 subproblems, rotations applied as MFMA products.  The
// matrices of one padded order N form a class; member k's working copies are
// A + k N^2 and V + k N^2 (float, pitch N).
constexpr int kBlockedWidth = 32;
uint32_t blocked_padded_order(uint32_t order);  // multiple of 64
struct BlockedMember {
  uint64_t offset;    // binary64 input and float V output (pitch ld)
  uint64_t s_offset;  // diagonal / roots (ld entries)
  uint32_t order;
  uint32_t ld;
};
// Partial results of the per-member reductions below: part k of member j at
// [j * kOffnormParts + k], blocked_offnorm_parts(count) parts per member,
// added in part order by the caller.
constexpr int kOffnormParts = 64;
uint32_t blocked_offnorm_parts(uint32_t count);
// Working copies of every member and ||A_k||_F^2 (binary64) in partial sums.
hipError_t launch_blocked_init(const BlockedMember* members, uint32_t count, const double* a64,
                               float* A, float* V, uint32_t N, double* norm2, hipStream_t stream);
// One block round of the `count` members listed in `active`: gather each
// member's pair subproblems into staging (binary64, count x N / 64 matrices of
// 64 x 64, in active-list order), diagonalise them with the batched kernel
// (jobs prepared by the caller: order 64, ld 64, one per subproblem) into U,
// then apply the rotations to A and V.
hipError_t launch_blocked_round(const uint32_t* active, uint32_t count, float* A, float* V,
                                uint32_t N, uint32_t round, double* staging, float* U,
                                const JacobiJob* jobs, float* pair_diag, float* pair_roots,
                                JacobiStatus* pair_status, hipStream_t stream);
// off(A)^2, ||G||_F^2 and min diag of each active member in partial results.
hipError_t launch_blocked_offnorm(const uint32_t* active, uint32_t count,
                                  const BlockedMember* members, const float* A, uint32_t N,
                                  double* off2, double* g2, float* dmin, hipStream_t stream);
// V (pitch ld), diagonal and roots of every member; bad[k] |= 1 for a
// diagonal entry that is not positive and finite (bad zeroed by the caller).
hipError_t launch_blocked_finish(const BlockedMember* members, uint32_t count, const float* A,
                                 const float* Vw, uint32_t N, float* V, float* diag, float* roots,
                                 uint32_t* bad, hipStream_t stream);
hipError_t launch_scatter_status(const JacobiStatus* src, const uint32_t* index, uint32_t count,
                                 JacobiStatus* dst, hipStream_t stream);

// -------------------------------------------------------------------- GEMM
// Operand element types of the batched square GEMMs.
enum class GemmType : int { F64 = 0, F32 = 1, Bf16 = 2, Fp8 = 3, Mxfp4 = 4 };

// Epilogues.
enum class Epilogue : int {
  Store = 0,     // C (or C^T) as float (double for F64), times alpha
  Hadamard = 1,  // C o H, H_ij = 1 / (s_i + s_j) inside the order, 0 outside
  Residual = 2,  // F64 only: R = A - C (double) and R as float, ||R||_F^2
  AddSym = 3,    // X += C on and above the diagonal, mirrored below (X binary64
                 // in c, inside the order); the plan lists the tiles with tile
                 // row <= tile column
  Gram = 4,      // no store: ||C - I||_F^2 inside the order into norms (with
                 // mirror, off-diagonal tiles of the upper list count twice)
  Accumulate = 5,  // C += c_scale C_new (float), times alpha
};

// One matrix of a batched GEMM launch: C = A B with both operands stored
// K-contiguous: element (m, k) of A at a[k + m ld] (that is, A^T column-major)
// and element (k, n) of B at b[k + n ld].  Offsets are in elements (bytes for
// the packed FP4 codes are offset / 2; MX scale bytes are offset / 32).
struct GemmJob {
  uint64_t a_offset;
  uint64_t b_offset;
  uint64_t c_offset;
  uint64_t aux_offset;   // residual: the binary64 matrix; Hadamard: roots (ld)
  uint32_t order;
  uint32_t ld;           // multiple of 32 (of 128 for MXFP4)
  uint32_t reserved;
  uint32_t status_index; // index into active[] / norms[]
};

struct GemmArgs {
  const GemmJob* jobs;
  const uint32_t* tile_job;  // job of each tile
  const uint32_t* tile_pos;  // (tile_row << 16) | tile_col
  const void* a;
  const void* b;
  const uint8_t* a_scales;   // MXFP4 block scales (E8M0) of A and B
  const uint8_t* b_scales;
  const float* alpha;        // per status_index multiplier (FP8 dequant), or null
  const int* active;         // per status_index flag, or null (all active)
  void* c;                   // float (double for F64)
  const void* aux;           // see GemmJob::aux_offset
  float* c_float;            // Residual: R as float
  double* norms;             // Residual: ||R||^2 accumulators per status_index
  int transpose_c;           // store C^T (row-major C) instead of C
  uint16_t* c_bf16;          // non-null: store C (Residual: R) as BF16 codes
                             // here instead of float / double (RNE, the
                             // codec of msq_quantize)
  int mirror;                // Residual: the plan lists tiles with tile row <=
                             // tile column; off-diagonal tiles also store their
                             // transpose and count twice in the norm.
                             // Hadamard: the same tile list; entries on and
                             // above the diagonal are stored at (i, j) and
                             // (j, i) (a symmetric product)
  float c_scale;             // Accumulate: C += c_scale * alpha * A B
  // Residual / Gram: each wave's share of ||.||^2 goes to partials[partial_base
  // + tile * waves + wave] (tile: the position in this launch's tile list,
  // waves: gemm_waves(tile edge)); launch_gemm_norms then adds each matrix's
  // slots in order, so the norms do not depend on which tile finishes first.
  double* partials;
  uint32_t partial_base;
};

hipError_t launch_gemm(GemmType type, Epilogue epi, int tile, const GemmArgs& args,
                       uint32_t tiles, hipStream_t stream);
// Waves per workgroup of a tile edge (32: 1, 64 and 128: 4).
int gemm_waves(int tile);
// norms[jobs[k].status_index] = sum of partials[first[k] .. first[k] + count[k])
// in order, for the jobs whose matrix is active (active null: all).
hipError_t launch_gemm_norms(const GemmJob* jobs, const uint32_t* first, const uint32_t* count,
                             uint32_t njobs, const double* partials, const int* active,
                             double* norms, hipStream_t stream);

// ------------------------------------------------------ elementwise helpers
// All per-matrix arrays of a batch share one segment descriptor: an ld x ld
// column-major tile at `offset` (zero outside the order) and ld-vectors at
// `vec_offset`.
struct Segment {
  uint64_t offset;
  uint64_t vec_offset;
  uint32_t order;
  uint32_t ld;
  uint32_t index;     // matrix index: slot in per-matrix arrays (active, sigma, ...)
  uint32_t reserved;  // tile-pair passes: pairs of the segments before this one
};

// Tile-pair passes: one workgroup per pair {(ti, tj), (tj, ti)}, ti <= tj, of
// 32 x 32 tiles; `pairs` is the total over the list, and each segment's
// `reserved` field the number of pairs before it (segments in list order).
hipError_t launch_symmetrize_to_f64(const Segment* segs, uint32_t count, uint32_t pairs,
                                    const float* c, double* x, const int* active,
                                    hipStream_t stream);
// The same as BF16 codes.
hipError_t launch_scale_rows_transpose_bf16(const Segment* segs, uint32_t count, uint32_t pairs,
                                            const float* v, const float* roots, int use_sqrt,
                                            uint16_t* out, const int* active, hipStream_t stream);
// active (may be null): matrices whose flag is 0 are skipped (out untouched).
hipError_t launch_scale_rows_transpose(const Segment* segs, uint32_t count, uint32_t pairs,
                                       const float* v, const float* roots, int use_sqrt,
                                       float* out, const int* active, hipStream_t stream);
hipError_t launch_transpose(const Segment* segs, uint32_t count, uint32_t pairs,
                            const float* v, float* out, hipStream_t stream);
// BF16 codes of v (cv) and of its per-matrix transpose (cvt), same layout.
hipError_t launch_transpose_encode_bf16(const Segment* segs, uint32_t count, uint32_t pairs,
                                        const float* v, uint16_t* cv, uint16_t* cvt,
                                        hipStream_t stream);
hipError_t launch_zero_f64(double* p, size_t n, hipStream_t stream);
hipError_t launch_subtract(const float* a, const float* b, float* out, size_t n,
                           hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t value, hipStream_t stream);
hipError_t launch_sumsq_f64(const Segment* segs, uint32_t count, uint32_t max_ld, const double* a,
                            double* out, hipStream_t stream);
hipError_t launch_orthogonality(const Segment* segs, uint32_t count, uint32_t max_ld,
                                const float* gram, double* out, hipStream_t stream);
hipError_t launch_segment_fp8(const Segment* segs, uint32_t count, uint32_t max_ld,
                              const float* src, uint32_t* amax, float* sigma, uint8_t* dst,
                              hipStream_t stream);
hipError_t launch_alpha(const float* sa, const float* sb, float* alpha, uint32_t count,
                        hipStream_t stream);
// Compact matrix storage of the native-quantized correction: segment k's
// n x n elements in column-major linear order at compact element offset
// coff[k] (a multiple of 64), in an msq_precision codec (the reference's host
// codecs, nearest rounding); the float side is the batch layout, optionally
// transposed.  active: per matrix index, or null.
hipError_t launch_compact_quantize(const Segment* segs, const uint64_t* coff, uint32_t count,
                                  uint32_t max_order, int precision, const float* src,
                                  bool src_transposed, uint8_t* codes, uint8_t* scales,
                                  const int* active, hipStream_t s);
hipError_t launch_compact_dequantize(const Segment* segs, const uint64_t* coff, uint32_t count,
                                    uint32_t max_ld, int precision, const uint8_t* codes,
                                    const uint8_t* scales, float* dst, bool dst_transposed,
                                    const int* active, hipStream_t s);
// Spectrum tier of a solve: each segment's roots (order entries at
// vec_offset) rounded to `precision` (an msq_precision value) and decoded into
// qroots; rejected[index] = 1 (and qroots = roots) when a decoded root is not
// positive and finite.  One workgroup per segment.
hipError_t launch_segment_spectrum(const Segment* segs, uint32_t count, int precision,
                                   const float* roots, float* qroots, int* rejected,
                                   const int* active, bool stochastic, uint32_t seed,
                                   uint32_t stream, hipStream_t s);

// ------------------------------------------------------------ solve state
constexpr uint32_t kSolveHistory = 65;
enum : uint32_t {
  kSolveRunning = 0,
  kSolveConverged = 1,
  kSolveBudgetExhausted = 2,
  kSolveDiverged = 3,
  kSolveNonfinite = 4,
  kSolveInvalidSpectrum = 5,
  kSolveStagnated = 6,
  kSolveUnresolved = 7,
};

struct SolveLimits {
  double residual_tolerance;  // stop when ||A - X^2||_F / ||A||_F <= this ...
  double forward_tolerance;   // ... or the forward-error estimate is <= this
  double minimum_relative_improvement;  // stagnation policy
  double basis_tolerance;     // basis acceptance (0: the group's tier default)
  uint32_t min_corrections;
  uint32_t max_corrections;
  uint32_t stagnation_window;  // 0: fixed budget
  float resolution;            // unresolved: min diag < resolution * max diag
  uint32_t refine;             // the second Jacobi pass runs for unresolved ones
  uint32_t reserved;
};

struct SolveState {
  double matrix_norm;            // ||A||_F
  double root_norm;              // sqrt(trace A) = ||A^(1/2)||_F
  double min_root;               // sqrt of the smallest Jacobi diagonal entry
  double residual_tolerance;
  double forward_tolerance;
  double minimum_relative_improvement;
  uint32_t min_corrections;
  uint32_t max_corrections;
  uint32_t stagnation_window;
  uint32_t steps;                // residuals recorded (initial root + corrections)
  uint32_t corrections;          // corrections applied
  uint32_t status;
  uint32_t unresolved;           // smallest eigenvalue below the resolution
  uint32_t best_iteration;
  uint32_t stagnation_steps;
  uint32_t escalated;            // basis tier rejected, X0 rebuilt in FP32
  uint32_t spectrum_rejected;    // spectrum tier rejected, X0 from FP32 roots
  uint32_t reserved;
  double best_residual;
  double final_residual;
  double final_forward_estimate;
  double basis_orthogonality;    // ||V_t^T V_t - I||_F / sqrt(n) of the basis-tier V
  double frozen_min_root;        // min of the roots X0 was formed with
  double fp32_orthogonality;     // ||V^T V - I||_F / sqrt(n) of the FP32 basis (NaN: not measured)
  double residuals[kSolveHistory];          // after X0, then after each correction
  double forward_estimates[kSolveHistory];  // ||R||_F / (2 s_min ||X||_F)
};

// allow_refine: matrices unresolved with refine set stay running even with a
// diagonal entry that is not positive (the second Jacobi pass decides them).
hipError_t launch_solve_state(const Segment* segs, const JacobiStatus* js,
                              const float* diagonal, const SolveLimits* limits, SolveState* st,
                              int* active, uint32_t count, bool allow_refine,
                              hipStream_t stream);
// After R = A - X^2 of step `step` (0: X0): history, stop rules, next active
// flag; save[i] = 1 when matrix i's residual is its best so far under the
// stagnation policy (the caller keeps that iterate).
hipError_t launch_convergence(SolveState* st, const double* norms, int* active, int* save,
                              uint32_t count, uint32_t step, hipStream_t stream);
// Basis escalation of the segments' matrices: orthogonality o = sqrt(orth2 /
// n) into the state, and esc = active && !(o <= tolerance), the tolerance
// limits[i].basis_tolerance or else `tolerance`.
hipError_t launch_basis_escalation(const Segment* segs, uint32_t count, const double* orth2,
                                   double tolerance, const SolveLimits* limits, const int* active,
                                   int* esc, SolveState* st, hipStream_t stream);
// ||V^T V - I||_F / sqrt(n) = sqrt(orth2 / n) of each segment's FP32 basis
// into the state's fp32_orthogonality.
hipError_t launch_record_orthogonality(const Segment* segs, uint32_t count, const double* orth2,
                                       SolveState* st, hipStream_t stream);
// Native-quantized matrices whose frozen roots were rejected stop as an
// invalid spectrum (the reference fails, src/metal_solver.cpp:720-727).
hipError_t launch_reject_spectrum(const Segment* segs, uint32_t count, const int* rejected,
                                  SolveState* st, int* active, hipStream_t stream);
// Frozen roots of each segment's matrix: min over qroots (the spectrum tier's,
// or the FP32 roots when that was rejected) into the state; rejected[i] into
// the state's spectrum_rejected.
hipError_t launch_frozen_roots(const Segment* segs, uint32_t count, const float* qroots,
                               const int* rejected, SolveState* st, hipStream_t stream);
// Copies the ld x ld tiles of the segments whose flag is set from src to dst
// (binary64); flag null: all.
hipError_t launch_copy_tiles_f64(const Segment* segs, uint32_t count, uint32_t max_ld,
                                 const double* src, double* dst, const int* flag,
                                 hipStream_t stream);
// x = ((x + d) + (x + d)^T) / 2 over each active segment's order (the
// reference's addCorrectionFromStorage then symmetrizeInPlace); d float.
hipError_t launch_add_symmetrize_f64(const Segment* segs, uint32_t count, uint32_t max_ld,
                                     double* x, const float* d, const int* active,
                                     hipStream_t stream);
// Native-quantized stages on float tiles (zero outside the order):
// out_ij = in_ij / (s_i + s_j) in float (diagonal_sylvester_*,
// metal_mixed_sqrt.metal:58-74, :913-932) and out_ij = in_ij * s_j
// (scale_columns_*, :21-37), s the segment's roots (vector layout).
hipError_t launch_sylvester_divide(const Segment* segs, uint32_t count, uint32_t max_ld,
                                   const float* in, const float* roots, float* out,
                                   const int* active, hipStream_t stream);
hipError_t launch_scale_columns(const Segment* segs, uint32_t count, uint32_t max_ld,
                                const float* in, const float* roots, float* out,
                                const int* active, hipStream_t stream);
// dst[0, words) = src[0, words), 4-byte words (float or binary64 data).
hipError_t launch_copy_words(const void* src, void* dst, size_t words, hipStream_t stream);
// Elementwise float -> binary64.
hipError_t launch_widen_f64(const float* src, double* dst, size_t n, hipStream_t stream);
// Tiles of the segments `from` (source batch layout) into consecutive tiles of
// `to` (destination batch layout), element type by size: gather (from =
// batch, to = sub-batch) or scatter (the reverse).  Both lists in the same
// order, same orders.
hipError_t launch_move_tiles(const Segment* from, const Segment* to, uint32_t count,
                             uint32_t max_ld, const void* src, void* dst, size_t elem_bytes,
                             hipStream_t stream);
hipError_t launch_move_vectors(const Segment* from, const Segment* to, uint32_t count,
                               uint32_t max_ld, const float* src, float* dst,
                               hipStream_t stream);

}  // namespace msq
