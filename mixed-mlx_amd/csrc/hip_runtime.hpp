// C++ mirror of the reference's Metal runtime and solver entry points, on top
// of the C ABI (include/msq_hip.h).  This is what a C++ caller of the
// reference's solver layer uses in place of MetalRuntime / runMetalCase:
//
//   MetalRuntime                (src/metal_runtime.hpp:69-176)  -> HipRuntime
//   MetalSolverConfiguration    (src/metal_solver.hpp:12-21)    -> HipSolverConfiguration
//   MetalProblem                (src/metal_solver.hpp:23-31)    -> HipProblem
//   CaseResult (subset)         (src/experiment.hpp:64-125)     -> HipCaseResult
//   runMetalCase / runTimedMetalCase (src/metal_solver.hpp:33-42)
//                                                               -> runHipCase / runTimedHipCase
//   makeSpdMatrix               (src/numerics.cpp:141-195)      -> makeSpdMatrix
//   fail()                      (src/core.cpp:7-9)              -> fail(): throws std::runtime_error
//
// Non-zero status from the C ABI becomes an exception carrying
// msq_last_error(), as the reference's checks call fail().  Measured outcomes
// (not converged, non-finite) are results, unless the problem enforces them.
#pragma once

#include <cstddef>
#include <cstdint>
#include <ostream>
#include <string>
#include <vector>

#include "../../include/msq_hip.h"

namespace msq_hip {

[[noreturn]] void fail(const std::string& message);

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
  explicit HipRuntime(int device = 0);
  ~HipRuntime();
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
  std::vector<msq_solve_report> solveBatched(const std::vector<uint32_t>& orders,
                                             const double* matrices, double* roots,
                                             const std::vector<msq_solve_options>& options);

 private:
  void check(int status, const char* what) const;
  msq_runtime* rt_ = nullptr;
};

struct HipSolverConfiguration {
  Precision basis_precision = Precision::Fp32;
  Precision correction_precision = Precision::Fp32;
  uint32_t fixed_sweeps = 0;     // 8: the reference's fixed budget
  double off_tolerance = 1e-7;   // off(A) <= off_tolerance ||A||_F
  float correction_rate = 0.0f;  // low-bit tiers: stop Jacobi at this predicted rate
  double forward_tolerance = 0.5e-6;
};

struct HipProblem {
  size_t order = 0;
  double condition = 1.0;
  size_t max_iterations = 12;      // corrections
  double tolerance = 0.0;          // relative FP64 residual target (0: forward estimate only)
  bool enforce_policy = false;     // fail() when the result misses its targets
  const std::vector<double>* prepared_matrix = nullptr;     // column-major, order^2
  const std::vector<double>* supplied_reference = nullptr;  // FP64 root, order^2
};

struct HipCaseResult {
  std::string name;
  std::string solver_backend = "hip";
  std::string basis_precision;
  std::string correction_precision;
  std::string status;
  std::string termination_reason;
  size_t order = 0;
  double condition = 0.0;
  bool converged = false;
  bool finite = false;
  size_t sweeps = 0;
  size_t corrections = 0;
  double off_norm = 0.0;
  double symmetry_error = 0.0;
  double final_residual = 0.0;
  double final_forward_estimate = 0.0;
  double final_forward_error = -1.0;  // vs the reference root (-1: none)
  double basis_orthogonality = 0.0;   // ||V_t^T V_t - I||_F / sqrt(n), low-bit basis
  size_t precision_escalations = 0;   // basis tier rejected for FP32
  size_t max_corrections = 0;
  size_t warmups = 0;
  std::vector<double> residual_history;          // relative FP64 residual per step
  std::vector<double> forward_estimate_history;  // first-order forward error bound per step
  std::vector<double> seconds;        // per evaluation (runTimedHipCase)
};

// The reference's controlled input: a deterministic sin/cos basis
// orthonormalised by Gram-Schmidt, eigenvalues cond^(j / (n - 1)).
std::vector<double> makeSpdMatrix(size_t order, double condition);
// Its square root from the same construction (binary64).
std::vector<double> makeSpdRoot(size_t order, double condition);

HipCaseResult runHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                         const HipProblem& problem);
HipCaseResult runTimedHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                              const HipProblem& problem, size_t warmups, size_t evaluations);

// key=value lines, as the reference's printResult (src/reporting.cpp:22).
void printResult(const HipCaseResult& result);

// CSV output with the reference's schema (writeSummaryHeader / writeCaseCsv /
// writeIterationsHeader, src/reporting.cpp:124-166, :168-305, :305-328), so
// HIP results drop into the reference's analysis pipeline: the same columns
// in the same order.  Columns this backend does not measure (the reference
// timing of the FP64 solve, the frozen-factor certificate, per-phase times)
// are written as 0, the string columns with the values the reference writes
// for a controlled SPD case.  iterations.csv has one row per residual step
// (the initial root, then each correction), with the step's first-order
// forward error bound in `forward_error`.
void writeSummaryHeader(std::ostream& output);
void writeIterationsHeader(std::ostream& output);
void writeCaseCsv(const HipCaseResult& result, std::ostream& summary, std::ostream& iterations);

}  // namespace msq_hip
