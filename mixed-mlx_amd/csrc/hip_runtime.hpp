// C++ mirror of the reference's Metal runtime and solver entry points, on top
// of the C ABI (include/msq_hip.h).  This is what a C++ caller of the
// reference's solver layer uses in place of MetalRuntime / runMetalCase:
//
//   MetalRuntime                (src/metal_runtime.hpp:69-176)  -> HipRuntime
//   MetalSolverConfiguration    (src/metal_solver.hpp:12-21)    -> HipSolverConfiguration
//   MetalProblem                (src/metal_solver.hpp:23-31)    -> HipProblem
//   CaseResult                  (src/experiment.hpp:64-125)     -> HipCaseResult
//   runMetalCase / runAdaptiveMetalCase / runTimedMetalCase (src/metal_solver.hpp:33-42)
//                                        -> runHipCase / runAdaptiveHipCase / runTimedHipCase
//   RoundingMode, KernelPath, CorrectionMode, NativeStoppingPolicy,
//   NativeDispatchPolicy and their names (src/precision.hpp, src/configuration.cpp)
//
// The FP64 reference root, prepared matrices (controlled and paper test
// sets) and timing summaries are in host_reference.hpp.  Non-zero status from
// the C ABI becomes an exception carrying msq_last_error(), as the
// reference's checks call fail().  Measured outcomes (not converged,
// non-finite) are results, unless the problem enforces them.
#pragma once

#include <cstddef>
#include <cstdint>
#include <limits>
#include <ostream>
#include <string>
#include <vector>

#include "../../include/msq_hip.h"
#include "host_reference.hpp"

namespace msq_hip {

// Precision values of the boundary (src/precision.hpp:9-20) plus the scaled
// FP8 tier of this backend.
enum class Precision : int {
  Fp32 = MSQ_FP32,
  Bf16 = MSQ_BF16,
  Fp8E4m3fn = MSQ_FP8_E4M3FN,
  Fp8E5m2 = MSQ_FP8_E5M2,
  Mxfp8E4m3fn = MSQ_MXFP8_E4M3FN,
  Mxfp8E5m2 = MSQ_MXFP8_E5M2,
  Fp4E2m1 = MSQ_FP4_E2M1,
  Mxfp4E2m1 = MSQ_MXFP4_E2M1,
  Fp8E4m3fnLogS = MSQ_FP8_E4M3FN_LOG_S,
  Fp4E2m1LogS = MSQ_FP4_E2M1_LOG_S,
  Fp8E4m3fnScaled = MSQ_FP8_E4M3FN_SCALED,
};
const char* precisionName(Precision precision);
Precision parsePrecision(const std::string& name);
// FP32 and BF16 demand full accuracy; the other formats are measured
// (src/precision.hpp:34-90, factorPrecisionPolicy src/configuration.cpp:159-170).
bool requiresFullAccuracy(Precision precision);
// Forward-error bound the enforce policy checks (verificationTolerance,
// src/configuration.cpp:98-107): FP32 1e-8, BF16 1e-6, 8-bit 1e-2, 4-bit 0.25.
double verificationTolerance(Precision precision);

enum class RoundingMode : int { Nearest = MSQ_ROUNDING_NEAREST, Stochastic = MSQ_ROUNDING_STOCHASTIC };
enum class KernelPath : int { Legacy = MSQ_KERNEL_PATH_LEGACY, Tiled = MSQ_KERNEL_PATH_TILED };
enum class CorrectionMode : int {
  Fp32Reorthogonalized = MSQ_CORRECTION_FP32_REORTHOGONALIZED,
  NativeQuantized = MSQ_CORRECTION_NATIVE_QUANTIZED,
};
enum class NativeStoppingPolicy : int {
  FixedBudget = MSQ_STOPPING_FIXED_BUDGET,
  Stagnation = MSQ_STOPPING_STAGNATION,
};
enum class NativeDispatchPolicy : int {
  PerStage = MSQ_DISPATCH_PER_STAGE,
  Fused = MSQ_DISPATCH_FUSED,
};
// Names of the reference's tables: nearest|stochastic, legacy|tiled,
// fp32_reorthogonalized|native_quantized, fixed_budget|stagnation,
// per_stage|fused.  The parsers fail on any other name.
const char* roundingModeName(RoundingMode mode);
RoundingMode parseRoundingMode(const std::string& name);
const char* kernelPathName(KernelPath path);
KernelPath parseKernelPath(const std::string& name);
const char* correctionModeName(CorrectionMode mode);
CorrectionMode parseCorrectionMode(const std::string& name);
const char* stoppingPolicyName(NativeStoppingPolicy policy);
NativeStoppingPolicy parseStoppingPolicy(const std::string& name);
const char* dispatchPolicyName(NativeDispatchPolicy policy);
NativeDispatchPolicy parseDispatchPolicy(const std::string& name);

// Device memory owned by one runtime (makeSharedBuffer / releaseBuffers,
// src/metal_runtime.cpp:1426-1432; explicit copies: no unified memory).
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(msq_runtime* runtime, size_t bytes);
  ~DeviceBuffer();
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& other) noexcept;
  DeviceBuffer& operator=(DeviceBuffer&& other) noexcept;
  void* data() const { return ptr_; }
  size_t bytes() const { return bytes_; }
  void upload(const void* host, size_t bytes);
  void download(void* host, size_t bytes) const;

 private:
  msq_runtime* rt_ = nullptr;
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

// MetalRuntime's role: one device, one stream, the device operations of the
// path.  Not copyable; calls are not reentrant (the reference's runtime is
// single-threaded too).
class HipRuntime {
 public:
  // MetalRuntime(shader, rounding_mode, rounding_seed): every quantization of
  // this runtime rounds with `rounding` (stochastic: seeded by `seed`).
  explicit HipRuntime(int device = 0, RoundingMode rounding = RoundingMode::Nearest,
                      uint32_t seed = 0);
  ~HipRuntime();
  RoundingMode roundingMode() const { return rounding_; }
  uint32_t roundingSeed() const { return seed_; }
  HipRuntime(const HipRuntime&) = delete;
  HipRuntime& operator=(const HipRuntime&) = delete;

  msq_runtime* handle() const { return rt_; }
  DeviceBuffer makeDeviceBuffer(size_t bytes);
  void synchronize();

  // symmetricJacobi (src/metal_runtime.cpp:584-646), batched: device pointers
  // in the batch layout of msq_batch_layout.
  std::vector<msq_jacobi_report> symmetricJacobi(const std::vector<uint32_t>& orders,
                                                 const double* matrices, float* vectors,
                                                 float* diagonal, float* roots,
                                                 const msq_jacobi_options& options);
  // quantizeAndDecode (src/metal_runtime.cpp:672-740).
  void quantizeAndDecode(Precision precision, const float* source, void* values,
                         uint8_t* scales, float* decoded, size_t elements);
  // measureF32Orthogonality (src/metal_runtime.cpp:742-776).
  std::vector<double> measureF32Orthogonality(const std::vector<uint32_t>& orders,
                                              const float* vectors);
  // reconstructTiled / reconstructLowFactorTiled (src/metal_runtime.cpp:429-534).
  void reconstruct(Precision precision, const std::vector<uint32_t>& orders, const float* vectors,
                   const float* roots, float* out);
  // initializeWorkingRoot / computeWorkingResidual (src/metal_runtime.cpp:778-834).
  std::vector<double> computeWorkingResidual(const std::vector<uint32_t>& orders,
                                             const double* matrices, const double* root,
                                             double* residual, float* residual_f32);
  // refineUpdateAndMeasure's correction products (src/metal_runtime.cpp:847-966).
  void correction(Precision precision, const std::vector<uint32_t>& orders, const float* vectors,
                  const float* roots, const float* residual_f32, float* out);
  // solveGpuResidentSpd (src/metal_solver.cpp:92-336), batched.
  // `outputs` (optional) also returns X0, the basis and the roots
  // (msq_solve_outputs; its `roots` member is replaced by `roots`).
  std::vector<msq_solve_report> solveBatched(const std::vector<uint32_t>& orders,
                                             const double* matrices, double* roots,
                                             const std::vector<msq_solve_options>& options,
                                             const msq_solve_outputs* outputs = nullptr);

 private:
  void check(int status, const char* what) const;
  msq_runtime* rt_ = nullptr;
  RoundingMode rounding_ = RoundingMode::Nearest;
  uint32_t seed_ = 0;
};

// MetalSolverConfiguration's fields (src/metal_solver.hpp:12-21), then the
// fields of this backend.  The GPU-resident path of the reference is
// correction_mode Fp32Reorthogonalized with kernel_path Tiled; NativeQuantized
// needs kernel_path Legacy, a basis format other than FP32 equal to the
// spectrum format (no log spectrum), and takes the stopping and dispatch
// policies (src/metal_solver.cpp:359-368, validateCorrectionMode).
struct HipSolverConfiguration {
  Precision basis_precision = Precision::Fp32;
  Precision spectrum_precision = Precision::Fp32;
  KernelPath kernel_path = KernelPath::Tiled;
  CorrectionMode correction_mode = CorrectionMode::Fp32Reorthogonalized;
  NativeStoppingPolicy stopping_policy = NativeStoppingPolicy::FixedBudget;
  size_t stagnation_window = 4;
  double minimum_relative_improvement = 1.0e-3;
  NativeDispatchPolicy dispatch_policy = NativeDispatchPolicy::Fused;
  // Backend fields.  correction_precision: format of the four correction
  // products (FP32: the reference's FP32 corrections).
  Precision correction_precision = Precision::Fp32;
  uint32_t fixed_sweeps = 0;     // 8: the reference's fixed budget
  double off_tolerance = 1e-7;   // off(A) <= off_tolerance ||A||_F
  float correction_rate = 0.0f;  // low-bit tiers: stop Jacobi at this predicted rate
  // First-order forward-error target of the corrections (0: the residual
  // tolerance of the problem alone, as the reference stops).
  double forward_tolerance = 0.0;
  double basis_tolerance = 0.0;   // 0: the tier's basis orthogonality tolerance
  float resolution = 0.0f;        // smallest resolved eigenvalue / largest (0: no test)
  bool refine_unresolved = false; // second Jacobi pass on V^T A V for unresolved spectra
};

// MetalProblem (src/metal_solver.hpp:23-31).
struct HipProblem {
  size_t order = 0;
  double condition = 1.0;
  size_t max_iterations = 12;      // corrections
  double tolerance = 1.0e-12;      // relative FP64 residual target
  bool enforce_policy = false;     // fail() when the result misses its targets
  const PreparedMatrix* prepared_matrix = nullptr;          // else the controlled matrix
  const std::vector<double>* supplied_reference = nullptr;  // FP64 root, order^2
};

struct IterationMeasurement {
  size_t iteration = 0;
  double residual = 0.0;
  double forward_error = 0.0;  // vs the FP64 root: the first and last step, NaN between
};

// PhaseTimes (src/experiment.hpp:54-62).  The device solve is one batched
// call, so its phases are not split: mixed_total_seconds is the call.
struct PhaseTimes {
  double reference_seconds = 0.0;
  double low_eigen_seconds = 0.0;
  double initial_gpu_seconds = 0.0;
  double residual_fp64_seconds = 0.0;
  double correction_gpu_seconds = 0.0;
  double update_fp64_seconds = 0.0;
  double mixed_total_seconds = 0.0;
};

// CaseResult (src/experiment.hpp:64-125), field for field, then the fields of
// this backend.
struct HipCaseResult {
  std::string name;
  std::string paper_test_set = "controlled";
  std::string matrix_family = "conditioned_spd";
  std::string matrix_type = "symmetric";
  std::string solver_backend = "hip";
  uint64_t rng_seed = 0;
  double diagonal_shift = 0.0;
  std::string shift_rule = "not_applicable";
  std::string precision;
  std::string basis_precision;
  std::string correction_mode;
  std::string spectrum_precision;
  std::string precision_policy;
  std::string rounding_mode = "nearest";
  uint32_t rounding_seed = 0;
  std::string execution_model;
  std::string kernel_path;
  std::string status;
  std::string stopping_policy = "fixed_budget";
  std::string dispatch_policy = "not_applicable";
  std::string termination_reason;
  std::string selection_policy = "fixed";
  std::string selected_basis_stage = "not_applicable";
  std::string selected_spectrum_stage = "not_applicable";
  std::string correction_stage_history = "not_applicable";
  size_t precision_attempts = 1;
  size_t order = 0;
  size_t low_factor_basis_bytes = 0;
  size_t fp32_basis_bytes = 0;
  size_t best_iteration = 0;
  size_t stagnation_steps = 0;
  size_t command_buffers_per_correction = 0;
  size_t kernel_dispatches_per_correction = 0;
  size_t precision_escalations = 0;
  size_t low_factor_corrections = 0;
  TimingSummary timing;
  TimingSummary reference_timing;
  std::vector<double> paired_ratios;
  PairedRatioSummary paired_ratio_summary;
  double condition = 0.0;
  bool converged = false;
  bool finite = false;
  double symmetry_error = 0.0;
  double reference_residual = 0.0;
  double frozen_factor_relative_error = 0.0;
  double frozen_factor_error = 0.0;
  double frozen_basis_orthogonality = 0.0;
  double quantized_basis_orthogonality = std::numeric_limits<double>::quiet_NaN();
  double factor_quantization_seconds = 0.0;
  double basis_reorthogonalization_seconds = 0.0;
  double frozen_sylvester_inverse_norm = 0.0;
  double frozen_contraction_bound = 0.0;
  bool certificate_usable = false;
  double best_residual = std::numeric_limits<double>::quiet_NaN();
  double best_forward_error = std::numeric_limits<double>::quiet_NaN();
  double final_residual = std::numeric_limits<double>::quiet_NaN();
  double final_forward_error = std::numeric_limits<double>::quiet_NaN();
  std::vector<IterationMeasurement> history;
  PhaseTimes times;
  // Backend fields.
  std::string correction_precision;
  size_t sweeps = 0;           // Jacobi sweeps (both passes when the second ran)
  size_t corrections = 0;
  size_t max_corrections = 0;
  double off_norm = 0.0;
  bool unresolved = false;
  bool refined = false;
  std::vector<double> forward_estimate_history;  // first-order forward bound per step
};

HipCaseResult runHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                         const HipProblem& problem);
// FP32 basis and spectrum on the controlled problem of the same order and
// condition (the prepared matrix and supplied root are dropped), selection
// policy latency_first; with enforce, the final forward error must be
// <= 1e-9 (src/metal_solver.cpp:976-1001).
HipCaseResult runAdaptiveHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                                 const HipProblem& problem);
// The reference's timing protocol (src/metal_solver.cpp:1003-1091): the FP64
// reference root once for accuracy, then warm-ups and evaluations each
// pairing one FP64 reference solve with one device solve, the reference first
// on even indices and second on odd ones; the result is the median-time
// evaluation with the mean phase times, both timing summaries and the paired
// ratios device / reference.
HipCaseResult runTimedHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                              const HipProblem& problem, size_t warmups, size_t evaluations);

// key=value lines, as the reference's printResult (src/reporting.cpp:22).
void printResult(const HipCaseResult& result);

// CSV output with the reference's schema (writeSummaryHeader / writeCaseCsv /
// writeIterationsHeader, src/reporting.cpp:124-166, :168-305, :305-328), so
// HIP results drop into the reference's analysis pipeline: the same columns
// in the same order, filled from the case result.  The per-phase times of the
// device solve are not split (one batched call): their columns are 0 and
// mixed_total_seconds holds the call.  iterations.csv has one row per
// residual step (the initial root, then each correction).
void writeSummaryHeader(std::ostream& output);
void writeIterationsHeader(std::ostream& output);
void writeCaseCsv(const HipCaseResult& result, std::ostream& summary, std::ostream& iterations);

}  // namespace msq_hip
