// C++ mirror of the Metal runtime and solver entry points over the C ABI.
// See hip_runtime.hpp for the correspondence with the reference.
#include "hip_runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <utility>
#include <vector>

namespace msq_hip {

namespace {

struct NamedPrecision {
  Precision precision;
  const char* name;
};
// Names of the reference's precision table (src/configuration.cpp).
const NamedPrecision kNames[] = {
    {Precision::Fp32, "fp32"},
    {Precision::Bf16, "bf16"},
    {Precision::Fp8E4m3fn, "fp8_e4m3fn"},
    {Precision::Fp8E5m2, "fp8_e5m2"},
    {Precision::Mxfp8E4m3fn, "mxfp8_e4m3fn"},
    {Precision::Mxfp8E5m2, "mxfp8_e5m2"},
    {Precision::Fp4E2m1, "fp4_e2m1"},
    {Precision::Mxfp4E2m1, "mxfp4_e2m1"},
    {Precision::Fp8E4m3fnLogS, "fp8_e4m3fn_log_s"},
    {Precision::Fp4E2m1LogS, "fp4_e2m1_log_s"},
    {Precision::Fp8E4m3fnScaled, "fp8_e4m3fn_scaled"},
};

std::string last_error() {
  const char* e = msq_last_error();
  return e ? std::string(e) : std::string("unknown error");
}

struct Batch {
  std::vector<uint32_t> lds;
  std::vector<uint64_t> offsets, vector_offsets;
  uint64_t total = 0, vector_total = 0;
};

Batch layout(const std::vector<uint32_t>& orders) {
  Batch b;
  b.lds.resize(orders.size());
  b.offsets.resize(orders.size());
  b.vector_offsets.resize(orders.size());
  if (msq_batch_layout(static_cast<uint32_t>(orders.size()), orders.data(), b.lds.data(),
                       b.offsets.data(), b.vector_offsets.data(), &b.total,
                       &b.vector_total) != MSQ_OK)
    fail("batch layout: " + last_error());
  return b;
}

const char* status_name(uint32_t s) {
  switch (s) {
    case MSQ_SOLVE_CONVERGED: return "converged";
    case MSQ_SOLVE_BUDGET_EXHAUSTED: return "budget_exhausted";
    case MSQ_SOLVE_DIVERGED: return "diverged";
    case MSQ_SOLVE_NONFINITE: return "nonfinite";
    case MSQ_SOLVE_INVALID_SPECTRUM: return "invalid_spectrum";
    case MSQ_SOLVE_STAGNATED: return "stagnated";
    case MSQ_SOLVE_UNRESOLVED: return "unresolved";
    default: return "running";
  }
}

}  // namespace

const char* precisionName(Precision precision) {
  for (const auto& p : kNames)
    if (p.precision == precision) return p.name;
  return "unknown";
}

Precision parsePrecision(const std::string& name) {
  for (const auto& p : kNames)
    if (name == p.name) return p.precision;
  fail("unknown precision '" + name + "'");
}

bool requiresFullAccuracy(Precision precision) {
  return precision == Precision::Fp32 || precision == Precision::Bf16;
}

double verificationTolerance(Precision precision) {
  if (precision == Precision::Fp32) return 1.0e-8;
  if (precision == Precision::Bf16) return 1.0e-6;
  const int p = static_cast<int>(precision);
  const bool four_bit = p == MSQ_FP4_E2M1 || p == MSQ_MXFP4_E2M1 || p == MSQ_FP4_E2M1_LOG_S;
  return four_bit ? 2.5e-1 : 1.0e-2;
}

namespace {

template <typename E>
struct EnumName {
  E value;
  const char* name;
};

template <typename E, size_t N>
const char* enum_name(const EnumName<E> (&table)[N], E value) {
  for (const auto& e : table)
    if (e.value == value) return e.name;
  fail("unknown enumeration value");
}

template <typename E, size_t N>
E enum_parse(const EnumName<E> (&table)[N], const std::string& name, const char* what) {
  for (const auto& e : table)
    if (name == e.name) return e.value;
  fail(std::string("unsupported ") + what + ": " + name);
}

const EnumName<RoundingMode> kRounding[] = {{RoundingMode::Nearest, "nearest"},
                                            {RoundingMode::Stochastic, "stochastic"}};
const EnumName<KernelPath> kKernelPaths[] = {{KernelPath::Legacy, "legacy"},
                                             {KernelPath::Tiled, "tiled"}};
const EnumName<CorrectionMode> kCorrectionModes[] = {
    {CorrectionMode::Fp32Reorthogonalized, "fp32_reorthogonalized"},
    {CorrectionMode::NativeQuantized, "native_quantized"}};
const EnumName<NativeStoppingPolicy> kStopping[] = {
    {NativeStoppingPolicy::FixedBudget, "fixed_budget"},
    {NativeStoppingPolicy::Stagnation, "stagnation"}};
const EnumName<NativeDispatchPolicy> kDispatch[] = {{NativeDispatchPolicy::PerStage, "per_stage"},
                                                    {NativeDispatchPolicy::Fused, "fused"}};

}  // namespace

const char* roundingModeName(RoundingMode m) { return enum_name(kRounding, m); }
RoundingMode parseRoundingMode(const std::string& n) {
  return enum_parse(kRounding, n, "rounding mode");
}
const char* kernelPathName(KernelPath k) { return enum_name(kKernelPaths, k); }
KernelPath parseKernelPath(const std::string& n) {
  return enum_parse(kKernelPaths, n, "kernel path");
}
const char* correctionModeName(CorrectionMode m) { return enum_name(kCorrectionModes, m); }
CorrectionMode parseCorrectionMode(const std::string& n) {
  return enum_parse(kCorrectionModes, n, "correction mode");
}
const char* stoppingPolicyName(NativeStoppingPolicy p) { return enum_name(kStopping, p); }
NativeStoppingPolicy parseStoppingPolicy(const std::string& n) {
  // The reference's command line spells the fixed budget "fixed".
  if (n == "fixed") return NativeStoppingPolicy::FixedBudget;
  return enum_parse(kStopping, n, "native stopping policy");
}
const char* dispatchPolicyName(NativeDispatchPolicy p) { return enum_name(kDispatch, p); }
NativeDispatchPolicy parseDispatchPolicy(const std::string& n) {
  return enum_parse(kDispatch, n, "native dispatch policy");
}

// ------------------------------------------------------------- DeviceBuffer
DeviceBuffer::DeviceBuffer(msq_runtime* runtime, size_t bytes) : rt_(runtime), bytes_(bytes) {
  if (msq_alloc(rt_, bytes, &ptr_) != MSQ_OK) fail("device allocation: " + last_error());
}

DeviceBuffer::~DeviceBuffer() {
  if (ptr_) msq_free(rt_, ptr_);
}

DeviceBuffer::DeviceBuffer(DeviceBuffer&& o) noexcept
    : rt_(o.rt_), ptr_(std::exchange(o.ptr_, nullptr)), bytes_(std::exchange(o.bytes_, 0)) {}

DeviceBuffer& DeviceBuffer::operator=(DeviceBuffer&& o) noexcept {
  if (this != &o) {
    if (ptr_) msq_free(rt_, ptr_);
    rt_ = o.rt_;
    ptr_ = std::exchange(o.ptr_, nullptr);
    bytes_ = std::exchange(o.bytes_, 0);
  }
  return *this;
}

void DeviceBuffer::upload(const void* host, size_t bytes) {
  if (bytes > bytes_) fail("upload larger than the buffer");
  if (msq_copy_to_device(rt_, ptr_, host, bytes) != MSQ_OK) fail("upload: " + last_error());
}

void DeviceBuffer::download(void* host, size_t bytes) const {
  if (bytes > bytes_) fail("download larger than the buffer");
  if (msq_copy_to_host(rt_, host, ptr_, bytes) != MSQ_OK) fail("download: " + last_error());
}

// --------------------------------------------------------------- HipRuntime
HipRuntime::HipRuntime(int device, RoundingMode rounding, uint32_t seed)
    : rounding_(rounding), seed_(seed) {
  if (msq_runtime_create(device, &rt_) != MSQ_OK) fail("HIP runtime: " + last_error());
  if (msq_runtime_set_rounding(rt_, static_cast<int>(rounding), seed) != MSQ_OK) {
    const std::string e = last_error();
    msq_runtime_destroy(rt_);
    rt_ = nullptr;
    fail("HIP runtime rounding: " + e);
  }
}

HipRuntime::~HipRuntime() {
  if (rt_) msq_runtime_destroy(rt_);
}

void HipRuntime::check(int status, const char* what) const {
  if (status != MSQ_OK) fail(std::string(what) + ": " + last_error());
}

DeviceBuffer HipRuntime::makeDeviceBuffer(size_t bytes) { return DeviceBuffer(rt_, bytes); }

void HipRuntime::synchronize() { check(msq_synchronize(rt_), "synchronize"); }

std::vector<msq_jacobi_report> HipRuntime::symmetricJacobi(const std::vector<uint32_t>& orders,
                                                           const double* matrices, float* vectors,
                                                           float* diagonal, float* roots,
                                                           const msq_jacobi_options& options) {
  std::vector<msq_jacobi_report> reports(orders.size());
  check(msq_jacobi_batched(rt_, static_cast<uint32_t>(orders.size()), orders.data(), matrices,
                           vectors, diagonal, roots, &options, 1, reports.data()),
        "symmetricJacobi");
  return reports;
}

void HipRuntime::quantizeAndDecode(Precision precision, const float* source, void* values,
                                   uint8_t* scales, float* decoded, size_t elements) {
  check(msq_quantize_and_decode(rt_, static_cast<int>(precision), source, values, scales, decoded,
                                elements),
        "quantizeAndDecode");
}

std::vector<double> HipRuntime::measureF32Orthogonality(const std::vector<uint32_t>& orders,
                                                        const float* vectors) {
  std::vector<double> out(orders.size());
  check(msq_orthogonality(rt_, static_cast<uint32_t>(orders.size()), orders.data(), vectors,
                          out.data()),
        "measureF32Orthogonality");
  return out;
}

void HipRuntime::reconstruct(Precision precision, const std::vector<uint32_t>& orders,
                             const float* vectors, const float* roots, float* out) {
  check(msq_reconstruct(rt_, static_cast<int>(precision), static_cast<uint32_t>(orders.size()),
                        orders.data(), vectors, roots, out),
        "reconstruct");
}

std::vector<double> HipRuntime::computeWorkingResidual(const std::vector<uint32_t>& orders,
                                                       const double* matrices, const double* root,
                                                       double* residual, float* residual_f32) {
  std::vector<double> rel(orders.size());
  check(msq_residual_fp64(rt_, static_cast<uint32_t>(orders.size()), orders.data(), matrices, root,
                          residual, residual_f32, rel.data()),
        "computeWorkingResidual");
  return rel;
}

void HipRuntime::correction(Precision precision, const std::vector<uint32_t>& orders,
                            const float* vectors, const float* roots, const float* residual_f32,
                            float* out) {
  check(msq_correction(rt_, static_cast<int>(precision), static_cast<uint32_t>(orders.size()),
                       orders.data(), vectors, roots, residual_f32, out),
        "correction");
}

std::vector<msq_solve_report> HipRuntime::solveBatched(
    const std::vector<uint32_t>& orders, const double* matrices, double* roots,
    const std::vector<msq_solve_options>& options, const msq_solve_outputs* outputs) {
  if (options.size() != 1 && options.size() != orders.size())
    fail("solveBatched: one options entry, or one per matrix");
  msq_solve_outputs out{};
  if (outputs) out = *outputs;
  out.roots = roots;
  std::vector<msq_solve_report> reports(orders.size());
  check(msq_solve_batched_ex(rt_, static_cast<uint32_t>(orders.size()), orders.data(), matrices,
                             &out, options.data(), static_cast<uint32_t>(options.size()),
                             reports.data()),
        "solveBatched");
  return reports;
}

// -------------------------------------------------------------- the solver
namespace {

// The reference's refinement budget (src/metal_solver.cpp:373-377): the
// requested corrections for full-accuracy formats or a single correction,
// otherwise at least 64 (this backend's history holds 64).
size_t refinement_iterations(const HipSolverConfiguration& c, size_t max_iterations) {
  const bool full = c.basis_precision == c.spectrum_precision &&
                    requiresFullAccuracy(c.basis_precision);
  const size_t k = full || max_iterations == 1 ? max_iterations
                                                : std::max<size_t>(max_iterations, 64);
  return std::min<size_t>(k, MSQ_SOLVE_HISTORY - 1);
}

msq_solve_options solve_options(const HipSolverConfiguration& c, const HipProblem& p) {
  if (c.stagnation_window == 0) fail("--native-stagnation-window must be a positive integer.");
  if (!std::isfinite(c.minimum_relative_improvement) || c.minimum_relative_improvement < 0.0)
    fail("--native-min-relative-improvement must be finite and nonnegative.");
  msq_solve_options o;
  msq_solve_options_init(&o);
  o.basis_precision = static_cast<int32_t>(c.basis_precision);
  o.spectrum_precision = static_cast<int32_t>(c.spectrum_precision);
  o.correction_precision = static_cast<int32_t>(c.correction_precision);
  o.kernel_path = static_cast<int32_t>(c.kernel_path);
  o.correction_mode = static_cast<int32_t>(c.correction_mode);
  o.stopping_policy = static_cast<int32_t>(c.stopping_policy);
  o.dispatch_policy = static_cast<int32_t>(c.dispatch_policy);
  o.stagnation_window = static_cast<uint32_t>(std::min<size_t>(c.stagnation_window, UINT32_MAX));
  o.minimum_relative_improvement = c.minimum_relative_improvement;
  o.jacobi.fixed_sweeps = c.fixed_sweeps;
  o.jacobi.max_sweeps = 30;
  o.jacobi.off_tol = static_cast<float>(c.off_tolerance);
  o.jacobi.off_tol_lambda = 0.0f;
  o.jacobi.stagnation = 0.5f;
  o.jacobi.floor_tol = 1e-5f;
  o.jacobi.correction_rate = c.correction_rate;
  o.min_corrections = 0;
  o.max_corrections = static_cast<uint32_t>(refinement_iterations(c, p.max_iterations));
  o.residual_tolerance = p.tolerance;
  o.forward_tolerance = c.forward_tolerance;
  o.basis_tolerance = c.basis_tolerance;
  o.resolution = c.resolution;
  o.refine_unresolved = c.refine_unresolved ? 1u : 0u;
  o.certificate = 1;
  return o;
}

// factorPrecisionLabel / executionModel (src/configuration.cpp:137-157).
std::string precision_label(const HipSolverConfiguration& c) {
  return c.basis_precision == c.spectrum_precision
             ? std::string(precisionName(c.basis_precision))
             : std::string(precisionName(c.basis_precision)) + "__s_" +
                   precisionName(c.spectrum_precision);
}

std::string condition_label(double condition) {
  std::ostringstream o;
  o << condition;
  std::string s = o.str();
  for (char& ch : s)
    if (ch == '.' || ch == '+') ch = 'p';
  return s;
}

std::vector<double> unpack(const std::vector<double>& tile, size_t n, size_t ld) {
  std::vector<double> out(n * n);
  for (size_t j = 0; j < n; ++j)
    for (size_t i = 0; i < n; ++i) out[i + j * n] = tile[i + j * ld];
  return out;
}

// One device solve of the prepared matrix; fills the device-side fields of
// the result (timings: the call, synchronised).
void solve_on_device(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                     const HipProblem& problem, const std::vector<double>& matrix,
                     const std::vector<double>& reference_root, HipCaseResult& r) {
  const size_t n = problem.order;
  const std::vector<uint32_t> orders = {static_cast<uint32_t>(n)};
  const Batch b = layout(orders);
  const size_t ld = b.lds[0];
  std::vector<double> tile(ld * ld, 0.0);
  for (size_t j = 0; j < n; ++j)
    for (size_t i = 0; i < n; ++i) tile[i + j * ld] = matrix[i + j * n];
  DeviceBuffer a = runtime.makeDeviceBuffer(tile.size() * sizeof(double));
  DeviceBuffer x = runtime.makeDeviceBuffer(tile.size() * sizeof(double));
  DeviceBuffer x0 = runtime.makeDeviceBuffer(tile.size() * sizeof(double));
  a.upload(tile.data(), tile.size() * sizeof(double));
  const std::vector<msq_solve_options> options = {solve_options(configuration, problem)};
  msq_solve_outputs outputs{};
  outputs.initial_roots = static_cast<double*>(x0.data());

  const auto t0 = std::chrono::steady_clock::now();
  const std::vector<msq_solve_report> reports =
      runtime.solveBatched(orders, static_cast<const double*>(a.data()),
                           static_cast<double*>(x.data()), options, &outputs);
  runtime.synchronize();
  r.times.mixed_total_seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  const msq_solve_report& rep = reports[0];

  x.download(tile.data(), tile.size() * sizeof(double));
  const std::vector<double> X = unpack(tile, n, ld);
  x0.download(tile.data(), tile.size() * sizeof(double));
  const std::vector<double> X0 = unpack(tile, n, ld);

  r.sweeps = rep.sweeps;
  r.corrections = rep.corrections;
  r.off_norm = rep.off_norm;
  r.unresolved = rep.unresolved != 0;
  r.refined = rep.refined != 0;
  r.precision_escalations = rep.precision_escalations;
  r.selected_basis_stage =
      configuration.basis_precision == Precision::Fp32
          ? std::string("fp32")
          : std::string(precisionName(configuration.basis_precision)) +
                (rep.basis_rejected ? "_rejected_to_fp32" : "");
  r.selected_spectrum_stage =
      configuration.spectrum_precision == Precision::Fp32
          ? std::string("fp32")
          : std::string(precisionName(configuration.spectrum_precision)) +
                (rep.spectrum_rejected ? "_rejected_to_fp32" : "");
  if (configuration.basis_precision != Precision::Fp32)
    r.quantized_basis_orthogonality = rep.basis_orthogonality;
  const bool native = configuration.correction_mode == CorrectionMode::NativeQuantized;
  const std::string stage = native ? std::string(precisionName(configuration.basis_precision)) +
                                         "_native"
                                   : std::string(precisionName(configuration.correction_precision)) +
                                         "_frozen";
  std::string stages;
  for (size_t k = 0; k < rep.corrections; ++k) stages += (k ? ";" : "") + stage;
  r.correction_stage_history = stages.empty() ? "not_applicable" : stages;
  r.low_factor_corrections =
      native || configuration.correction_precision != Precision::Fp32 ? rep.corrections : 0;

  // Frozen-factor certificate (src/metal_solver.cpp:511-524).
  r.frozen_basis_orthogonality = rep.frozen_basis_orthogonality;
  r.frozen_sylvester_inverse_norm = rep.frozen_sylvester_inverse_norm;
  r.certificate_usable = rep.certificate_usable != 0;
  r.frozen_factor_error = frobeniusDifference(X0, reference_root);
  r.frozen_factor_relative_error = relativeError(X0, reference_root);
  r.frozen_contraction_bound = 3.0 * r.frozen_sylvester_inverse_norm * r.frozen_factor_error;

  // History: relative FP64 residual per step; the forward error against the
  // FP64 root at the first and the last step (src/metal_solver.cpp:525-548).
  const size_t steps = std::min<size_t>(rep.corrections + 1, MSQ_SOLVE_HISTORY);
  const double initial_error = relativeError(X0, reference_root);
  const double final_error = relativeError(X, reference_root);
  r.history.clear();
  r.forward_estimate_history.clear();
  for (size_t k = 0; k < steps; ++k) {
    double fe = std::numeric_limits<double>::quiet_NaN();
    if (k == 0) fe = initial_error;
    if (k + 1 == steps) fe = final_error;
    r.history.push_back({k, rep.residuals[k], fe});
    r.forward_estimate_history.push_back(rep.forward_estimates[k]);
  }
  r.best_iteration = rep.best_iteration;
  r.best_residual = rep.best_residual;
  r.stagnation_steps = rep.stagnation_steps;
  r.final_residual = rep.final_residual;
  r.final_forward_error = final_error;
  // With the stagnation stop the best iterate is returned; its forward error
  // is the final one.
  r.best_forward_error = r.best_iteration + 1 == steps || rep.status == MSQ_SOLVE_STAGNATED
                             ? final_error
                             : std::numeric_limits<double>::quiet_NaN();
  r.finite = allFinite(X) && std::isfinite(r.final_residual) && std::isfinite(final_error);
  r.converged = r.finite && rep.status == MSQ_SOLVE_CONVERGED;
  r.symmetry_error = relativeSymmetryError(X, n);
  r.status = r.finite ? status_name(rep.status) : "nonfinite";
  r.termination_reason = r.status;
}

HipCaseResult run_case(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                       const HipProblem& problem, bool check_enforce) {
  const size_t n = problem.order;
  if (n < 2) fail("The matrix order must be at least two.");
  if (n > 65535) fail("The matrix order exceeds the batch layout's range.");
  if (configuration.correction_mode == CorrectionMode::NativeQuantized &&
      runtime.roundingMode() == RoundingMode::Stochastic)
    fail("Stochastic rounding supports frozen-factor quantization with FP32 correction. It "
         "does not support native quantized correction.");
  HipCaseResult r;
  r.rounding_mode = roundingModeName(runtime.roundingMode());
  r.rounding_seed = runtime.roundingSeed();
  PreparedMatrix controlled;
  const PreparedMatrix* prepared = problem.prepared_matrix;
  if (prepared) {
    if (prepared->order != n || prepared->values.size() != n * n ||
        prepared->matrix_type != "symmetric" || !allFinite(prepared->values))
      fail("HIP requires a finite prepared symmetric matrix of the selected order.");
  } else {
    controlled = prepareControlledMatrix(n, problem.condition);
    prepared = &controlled;
  }
  r.paper_test_set = prepared->paper_test_set;
  r.matrix_family = prepared->matrix_family;
  r.matrix_type = prepared->matrix_type;
  r.rng_seed = prepared->rng_seed;
  r.diagonal_shift = prepared->diagonal_shift;
  r.shift_rule = prepared->shift_rule;
  r.condition = prepared->condition;
  r.order = n;

  const bool native = configuration.correction_mode == CorrectionMode::NativeQuantized;
  r.precision = precision_label(configuration);
  r.basis_precision = precisionName(configuration.basis_precision);
  r.spectrum_precision = precisionName(configuration.spectrum_precision);
  r.correction_precision = precisionName(configuration.correction_precision);
  r.correction_mode = correctionModeName(configuration.correction_mode);
  r.kernel_path = kernelPathName(configuration.kernel_path);
  const bool full = configuration.basis_precision == configuration.spectrum_precision &&
                    requiresFullAccuracy(configuration.basis_precision);
  r.precision_policy = full ? "required_full_accuracy" : "measured_controlled";
  r.stopping_policy = stoppingPolicyName(native ? configuration.stopping_policy
                                                : NativeStoppingPolicy::FixedBudget);
  if (native) {
    r.execution_model = "native_quantized_correction";
    r.dispatch_policy = dispatchPolicyName(configuration.dispatch_policy);
  } else {
    r.execution_model = "gpu_resident_batched_jacobi_" +
                        std::string(precisionName(configuration.correction_precision)) +
                        "_products_fp64_residual";
    r.dispatch_policy = "gpu_resident";
  }
  r.command_buffers_per_correction = 1;
  r.kernel_dispatches_per_correction = native ? 10 : 6;
  r.name = r.paper_test_set == "controlled"
               ? "spd_n" + std::to_string(n) + "_k" + condition_label(r.condition) + "_" +
                     r.precision + "_" + r.correction_mode
               : "paper_" + r.paper_test_set + "_hip_" + r.precision;
  if (runtime.roundingMode() == RoundingMode::Stochastic &&
      (configuration.basis_precision != Precision::Fp32 ||
       configuration.spectrum_precision != Precision::Fp32))
    r.name += "_stochastic";
  const size_t n2 = n * n;
  const int basis = static_cast<int>(configuration.basis_precision);
  r.low_factor_basis_bytes = msq_storage_bytes(basis, n2) + msq_scale_bytes(basis, n2);
  r.fp32_basis_bytes = 4 * n2;
  r.max_corrections = refinement_iterations(configuration, problem.max_iterations);

  std::vector<double> computed_root;
  const std::vector<double>* reference_root = problem.supplied_reference;
  if (!reference_root) {
    r.times.reference_seconds = timeReferenceSolve(*prepared, &computed_root);
    reference_root = &computed_root;
  }
  if (reference_root->size() != n2 || !allFinite(*reference_root))
    fail("The supplied FP64 reference root is invalid.");
  r.reference_residual = relativeResidual(prepared->values, *reference_root, n);

  const double reference_seconds = r.times.reference_seconds;
  solve_on_device(runtime, configuration, problem, prepared->values, *reference_root, r);
  r.times.reference_seconds = reference_seconds;
  const double t = r.times.mixed_total_seconds;
  r.timing = TimingSummary{0, 1, t, t, t, t, 0.0};
  r.reference_timing =
      TimingSummary{0, 1, reference_seconds, reference_seconds, reference_seconds,
                    reference_seconds, 0.0};
  if (check_enforce && problem.enforce_policy && full) {
    if (!r.finite) fail("The HIP solver produced a nonfinite result.");
    if (!r.converged) {
      std::ostringstream m;
      m.precision(17);
      m << "The HIP solver did not reach its target (" << r.status
        << "). Residual: " << r.final_residual;
      fail(m.str());
    }
    if (r.final_forward_error > verificationTolerance(configuration.spectrum_precision))
      fail("The HIP result does not match the FP64 reference.");
    if (r.symmetry_error > 1.0e-12) fail("The HIP result lost symmetric structure.");
  }
  return r;
}

}  // namespace

HipCaseResult runHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                         const HipProblem& problem) {
  return run_case(runtime, configuration, problem, true);
}

HipCaseResult runAdaptiveHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                                 const HipProblem& problem) {
  if (configuration.correction_mode == CorrectionMode::NativeQuantized)
    fail("--adaptive does not support native quantized correction.");
  HipSolverConfiguration fp32 = configuration;
  fp32.basis_precision = Precision::Fp32;
  fp32.spectrum_precision = Precision::Fp32;
  fp32.correction_precision = Precision::Fp32;
  HipProblem controlled = problem;
  controlled.enforce_policy = false;
  controlled.prepared_matrix = nullptr;
  controlled.supplied_reference = nullptr;
  HipCaseResult selected = runHipCase(runtime, fp32, controlled);
  selected.selection_policy = "latency_first";
  selected.precision_attempts = 1;
  selected.execution_model += "_adaptive";
  if (problem.enforce_policy &&
      (!selected.finite || !selected.converged || selected.history.back().forward_error > 1.0e-9 ||
       selected.symmetry_error > 1.0e-12))
    fail("The adaptive solver did not meet the full-accuracy contract.");
  return selected;
}

HipCaseResult runTimedHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                              const HipProblem& problem, size_t warmups, size_t evaluations) {
  if (evaluations == 0) fail("The evaluation count must be nonzero.");
  PreparedMatrix controlled;
  const PreparedMatrix* prepared = problem.prepared_matrix;
  if (!prepared) {
    controlled = prepareControlledMatrix(problem.order, problem.condition);
    prepared = &controlled;
  }
  if (prepared->order != problem.order || prepared->matrix_type != "symmetric")
    fail("The timed HIP case requires a prepared symmetric matrix.");
  std::vector<double> accuracy_reference;
  timeReferenceSolve(*prepared, &accuracy_reference);
  HipProblem measured = problem;
  measured.prepared_matrix = prepared;
  measured.supplied_reference = &accuracy_reference;
  auto run_mixed = [&]() { return runHipCase(runtime, configuration, measured); };
  for (size_t w = 0; w < warmups; ++w) {
    if (w % 2 == 0) {
      timeReferenceSolve(*prepared);
      run_mixed();
    } else {
      run_mixed();
      timeReferenceSolve(*prepared);
    }
  }
  std::vector<HipCaseResult> results;
  std::vector<double> mixed_samples, reference_samples;
  results.reserve(evaluations);
  for (size_t e = 0; e < evaluations; ++e) {
    if (e % 2 == 0) {
      reference_samples.push_back(timeReferenceSolve(*prepared));
      results.push_back(run_mixed());
    } else {
      results.push_back(run_mixed());
      reference_samples.push_back(timeReferenceSolve(*prepared));
    }
    mixed_samples.push_back(results.back().times.mixed_total_seconds);
  }
  PhaseTimes mean;
  for (const HipCaseResult& s : results) {
    mean.low_eigen_seconds += s.times.low_eigen_seconds;
    mean.initial_gpu_seconds += s.times.initial_gpu_seconds;
    mean.residual_fp64_seconds += s.times.residual_fp64_seconds;
    mean.correction_gpu_seconds += s.times.correction_gpu_seconds;
    mean.update_fp64_seconds += s.times.update_fp64_seconds;
    mean.mixed_total_seconds += s.times.mixed_total_seconds;
  }
  const double count = static_cast<double>(results.size());
  mean.low_eigen_seconds /= count;
  mean.initial_gpu_seconds /= count;
  mean.residual_fp64_seconds /= count;
  mean.correction_gpu_seconds /= count;
  mean.update_fp64_seconds /= count;
  mean.mixed_total_seconds /= count;
  const TimingSummary mixed_timing = summarizeTimingSamples(mixed_samples, warmups);
  const TimingSummary reference_timing = summarizeTimingSamples(reference_samples, warmups);
  std::vector<double> ratios;
  const PairedRatioSummary paired =
      summarizePairedRatios(mixed_samples, reference_samples, evaluations, ratios);
  std::stable_sort(results.begin(), results.end(),
                   [](const HipCaseResult& l, const HipCaseResult& r) {
                     return l.times.mixed_total_seconds < r.times.mixed_total_seconds;
                   });
  HipCaseResult result = std::move(results[results.size() / 2]);
  mean.reference_seconds = reference_timing.mean_seconds;
  result.times = mean;
  result.timing = mixed_timing;
  result.reference_timing = reference_timing;
  result.paired_ratios = std::move(ratios);
  result.paired_ratio_summary = paired;
  return result;
}

namespace {

using Field = std::pair<const char*, std::string>;

std::string num(double v) {
  std::ostringstream o;
  o.precision(17);
  o << v;
  return o.str();
}
std::string num(size_t v) { return std::to_string(v); }
std::string num(uint64_t v, int) { return std::to_string(v); }
std::string flag(bool v) { return v ? "1" : "0"; }

// updatesAndOthersSeconds (src/reporting.cpp:10-15).
double updates_and_others(const HipCaseResult& r) {
  const double accounted = r.times.low_eigen_seconds + r.times.initial_gpu_seconds +
                           r.times.residual_fp64_seconds + r.times.correction_gpu_seconds;
  return std::max(0.0, r.times.mixed_total_seconds - accounted);
}

double storage_ratio(const HipCaseResult& r) {
  return r.fp32_basis_bytes > 0 ? static_cast<double>(r.low_factor_basis_bytes) /
                                      static_cast<double>(r.fp32_basis_bytes)
                                : 0.0;
}

double mixed_to_reference(const HipCaseResult& r) {
  return r.reference_timing.mean_seconds > 0.0
             ? r.timing.mean_seconds / r.reference_timing.mean_seconds
             : 0.0;
}

void append(std::vector<Field>& out, std::vector<Field> more) {
  out.insert(out.end(), std::make_move_iterator(more.begin()),
             std::make_move_iterator(more.end()));
}

// paper_test_set .. kernel_path (both files, src/reporting.cpp:124-131).
std::vector<Field> case_fields(const HipCaseResult& r) {
  return {{"paper_test_set", r.paper_test_set},
          {"matrix_family", r.matrix_family},
          {"matrix_type", r.matrix_type},
          {"solver_backend", r.solver_backend},
          {"rng_seed", num(r.rng_seed, 0)},
          {"rounding_mode", r.rounding_mode},
          {"rounding_seed", num(static_cast<size_t>(r.rounding_seed))},
          {"diagonal_shift", num(r.diagonal_shift)},
          {"shift_rule", r.shift_rule},
          {"case", r.name},
          {"precision", r.precision},
          {"basis_precision", r.basis_precision},
          {"spectrum_precision", r.spectrum_precision},
          {"correction_mode", r.correction_mode},
          {"precision_policy", r.precision_policy},
          {"stopping_policy", r.stopping_policy},
          {"dispatch_policy", r.dispatch_policy},
          {"termination_reason", r.termination_reason},
          {"selection_policy", r.selection_policy},
          {"precision_attempts", num(r.precision_attempts)},
          {"execution_model", r.execution_model},
          {"kernel_path", r.kernel_path}};
}

// selected_basis_stage .. basis_reorthogonalization_seconds (both files).
std::vector<Field> basis_fields(const HipCaseResult& r) {
  return {{"selected_basis_stage", r.selected_basis_stage},
          {"selected_spectrum_stage", r.selected_spectrum_stage},
          {"correction_stage_history", r.correction_stage_history},
          {"precision_escalations", num(r.precision_escalations)},
          {"low_factor_corrections", num(r.low_factor_corrections)},
          {"low_factor_basis_bytes", num(r.low_factor_basis_bytes)},
          {"fp32_basis_bytes", num(r.fp32_basis_bytes)},
          {"basis_storage_ratio", num(storage_ratio(r))},
          {"quantized_basis_orthogonality", num(r.quantized_basis_orthogonality)},
          {"factor_quantization_seconds", num(r.factor_quantization_seconds)},
          {"basis_reorthogonalization_seconds", num(r.basis_reorthogonalization_seconds)}};
}

std::vector<Field> reference_timing_fields(const HipCaseResult& r) {
  return {{"reference_timing_min_seconds", num(r.reference_timing.minimum_seconds)},
          {"reference_timing_max_seconds", num(r.reference_timing.maximum_seconds)},
          {"reference_timing_median_seconds", num(r.reference_timing.median_seconds)},
          {"reference_timing_mean_seconds", num(r.reference_timing.mean_seconds)},
          {"reference_timing_variance_seconds2", num(r.reference_timing.variance_seconds2)}};
}

std::vector<Field> paired_fields(const HipCaseResult& r) {
  return {{"paired_ratio_mean", num(r.paired_ratio_summary.mean)},
          {"paired_ratio_variance", num(r.paired_ratio_summary.variance)},
          {"paired_ratio_min", num(r.paired_ratio_summary.minimum)},
          {"paired_ratio_max", num(r.paired_ratio_summary.maximum)},
          {"paired_ratio_median", num(r.paired_ratio_summary.median)}};
}

// factorization_seconds .. mixed_total_seconds.
std::vector<Field> phase_fields(const HipCaseResult& r) {
  return {{"factorization_seconds", num(r.times.low_eigen_seconds)},
          {"initial_root_seconds", num(r.times.initial_gpu_seconds)},
          {"single_seconds", num(r.times.low_eigen_seconds + r.times.initial_gpu_seconds)},
          {"residual_seconds", num(r.times.residual_fp64_seconds)},
          {"correction_system_seconds", num(r.times.correction_gpu_seconds)},
          {"updates_and_others_seconds", num(updates_and_others(r))},
          {"mixed_total_seconds", num(r.times.mixed_total_seconds)}};
}

std::vector<Field> summary_fields(const HipCaseResult& r) {
  const double total = r.times.mixed_total_seconds;
  auto fraction = [total](double seconds) { return num(total > 0.0 ? seconds / total : 0.0); };
  const IterationMeasurement initial = r.history.empty() ? IterationMeasurement{} : r.history.front();
  const IterationMeasurement last = r.history.empty() ? IterationMeasurement{} : r.history.back();
  std::vector<Field> f = case_fields(r);
  append(f, {{"command_buffers_per_correction", num(r.command_buffers_per_correction)},
             {"kernel_dispatches_per_correction", num(r.kernel_dispatches_per_correction)}});
  append(f, basis_fields(r));
  append(f, {{"warmups", num(r.timing.warmups)},
             {"evaluations", num(r.timing.evaluations)},
             {"timing_evaluations", num(r.timing.evaluations)},
             {"timing_min_seconds", num(r.timing.minimum_seconds)},
             {"timing_max_seconds", num(r.timing.maximum_seconds)},
             {"timing_median_seconds", num(r.timing.median_seconds)},
             {"timing_mean_seconds", num(r.timing.mean_seconds)},
             {"timing_variance_seconds2", num(r.timing.variance_seconds2)}});
  append(f, reference_timing_fields(r));
  append(f, {{"order", num(r.order)},
             {"condition", num(r.condition)},
             {"iterations", num(last.iteration)},
             {"best_iteration", num(r.best_iteration)},
             {"best_residual", num(r.best_residual)},
             {"best_forward_error", num(r.best_forward_error)},
             {"stagnation_steps", num(r.stagnation_steps)},
             {"converged", flag(r.converged)},
             {"finite", flag(r.finite)},
             {"status", r.status},
             {"initial_residual", num(initial.residual)},
             {"final_residual", num(r.final_residual)},
             {"initial_forward_error", num(initial.forward_error)},
             {"final_forward_error", num(r.final_forward_error)},
             {"symmetry_error", num(r.symmetry_error)},
             {"reference_residual", num(r.reference_residual)},
             {"frozen_factor_error", num(r.frozen_factor_error)},
             {"frozen_factor_relative_error", num(r.frozen_factor_relative_error)},
             {"frozen_basis_orthogonality", num(r.frozen_basis_orthogonality)},
             {"frozen_sylvester_inverse_norm", num(r.frozen_sylvester_inverse_norm)},
             {"frozen_contraction_bound", num(r.frozen_contraction_bound)},
             {"certificate_usable", flag(r.certificate_usable)},
             {"reference_seconds", num(r.times.reference_seconds)},
             {"low_eigen_seconds", num(r.times.low_eigen_seconds)},
             {"initial_gpu_seconds", num(r.times.initial_gpu_seconds)},
             {"residual_fp64_seconds", num(r.times.residual_fp64_seconds)},
             {"correction_gpu_seconds", num(r.times.correction_gpu_seconds)},
             {"update_fp64_seconds", num(r.times.update_fp64_seconds)}});
  append(f, phase_fields(r));
  append(f, {{"low_eigen_fraction", fraction(r.times.low_eigen_seconds)},
             {"initial_gpu_fraction", fraction(r.times.initial_gpu_seconds)},
             {"residual_fraction", fraction(r.times.residual_fp64_seconds)},
             {"correction_gpu_fraction", fraction(r.times.correction_gpu_seconds)},
             {"update_fraction", fraction(r.times.update_fp64_seconds)}});
  append(f, paired_fields(r));
  append(f, {{"mixed_to_reference", num(mixed_to_reference(r))}});
  return f;
}

std::vector<Field> iteration_fields(const HipCaseResult& r, const IterationMeasurement& m) {
  std::vector<Field> f = case_fields(r);
  append(f, basis_fields(r));
  append(f, {{"order", num(r.order)},
             {"condition", num(r.condition)},
             {"warmups", num(r.timing.warmups)},
             {"evaluations", num(r.timing.evaluations)},
             {"timing_evaluations", num(r.timing.evaluations)}});
  append(f, reference_timing_fields(r));
  append(f, paired_fields(r));
  append(f, phase_fields(r));
  append(f, {{"iteration", num(m.iteration)},
             {"residual", num(m.residual)},
             {"forward_error", num(m.forward_error)}});
  return f;
}

void write_names(std::ostream& out, const std::vector<Field>& f) {
  for (size_t i = 0; i < f.size(); ++i) out << (i ? "," : "") << f[i].first;
  out << '\n';
}
void write_values(std::ostream& out, const std::vector<Field>& f) {
  for (size_t i = 0; i < f.size(); ++i) out << (i ? "," : "") << f[i].second;
  out << '\n';
}

}  // namespace

// printResult (src/reporting.cpp:22-122): space-separated key=value tokens,
// one history line per step.  The backend's own fields follow on a line of
// their own.
void printResult(const HipCaseResult& r) {
  std::ostringstream o;
  o.precision(6);
  o << "paper_test_set=" << r.paper_test_set << " matrix_family=" << r.matrix_family
    << " matrix_type=" << r.matrix_type << " solver_backend=" << r.solver_backend
    << " order=" << r.order << " condition=" << r.condition << " rng_seed=" << r.rng_seed
    << " rounding_mode=" << r.rounding_mode << " rounding_seed=" << r.rounding_seed
    << " diagonal_shift=" << num(r.diagonal_shift) << " shift_rule=" << r.shift_rule << '\n';
  for (const IterationMeasurement& m : r.history)
    o << "case=" << r.name << " selection_policy=" << r.selection_policy
      << " precision_attempts=" << r.precision_attempts << " correction_mode=" << r.correction_mode
      << " precision_policy=" << r.precision_policy << " execution_model=" << r.execution_model
      << " kernel_path=" << r.kernel_path << " dispatch_policy=" << r.dispatch_policy
      << " iteration=" << m.iteration << " residual=" << m.residual
      << " forward_error=" << m.forward_error << '\n';
  o << "reference_residual=" << r.reference_residual << '\n';
  o << "frozen_factor_error=" << r.frozen_factor_error
    << " frozen_factor_relative_error=" << r.frozen_factor_relative_error
    << " frozen_basis_orthogonality=" << r.frozen_basis_orthogonality
    << " frozen_sylvester_inverse_norm=" << r.frozen_sylvester_inverse_norm
    << " frozen_contraction_bound=" << r.frozen_contraction_bound
    << " certificate_usable=" << (r.certificate_usable ? "true" : "false") << '\n';
  o << "selected_basis_stage=" << r.selected_basis_stage
    << " selected_spectrum_stage=" << r.selected_spectrum_stage
    << " correction_stage_history=" << r.correction_stage_history
    << " precision_escalations=" << r.precision_escalations
    << " low_factor_basis_bytes=" << r.low_factor_basis_bytes
    << " fp32_basis_bytes=" << r.fp32_basis_bytes << " basis_storage_ratio=" << storage_ratio(r)
    << " low_factor_corrections=" << r.low_factor_corrections
    << " quantized_basis_orthogonality=" << r.quantized_basis_orthogonality << '\n';
  o << "best_iteration=" << r.best_iteration << " best_residual=" << r.best_residual
    << " best_forward_error=" << r.best_forward_error
    << " stagnation_steps=" << r.stagnation_steps << '\n';
  o << "final_residual=" << r.final_residual << '\n';
  o << "final_forward_error=" << r.final_forward_error << '\n';
  o << "symmetry_error=" << r.symmetry_error << '\n';
  o << "mixed_total_seconds=" << r.times.mixed_total_seconds << '\n';
  o << "factorization_seconds=" << r.times.low_eigen_seconds
    << " initial_root_seconds=" << r.times.initial_gpu_seconds
    << " single_seconds=" << r.times.low_eigen_seconds + r.times.initial_gpu_seconds
    << " residual_seconds=" << r.times.residual_fp64_seconds
    << " correction_system_seconds=" << r.times.correction_gpu_seconds
    << " updates_and_others_seconds=" << updates_and_others(r)
    << " factor_quantization_seconds=" << r.factor_quantization_seconds
    << " basis_reorthogonalization_seconds=" << r.basis_reorthogonalization_seconds << '\n';
  o << "timing_warmups=" << r.timing.warmups << " timing_evaluations=" << r.timing.evaluations
    << " timing_min_seconds=" << r.timing.minimum_seconds
    << " timing_max_seconds=" << r.timing.maximum_seconds
    << " timing_median_seconds=" << r.timing.median_seconds
    << " timing_mean_seconds=" << r.timing.mean_seconds
    << " timing_variance_seconds2=" << r.timing.variance_seconds2 << '\n';
  o << "reference_timing_warmups=" << r.reference_timing.warmups
    << " reference_timing_evaluations=" << r.reference_timing.evaluations
    << " reference_timing_min_seconds=" << r.reference_timing.minimum_seconds
    << " reference_timing_max_seconds=" << r.reference_timing.maximum_seconds
    << " reference_timing_median_seconds=" << r.reference_timing.median_seconds
    << " reference_timing_mean_seconds=" << r.reference_timing.mean_seconds
    << " reference_timing_variance_seconds2=" << r.reference_timing.variance_seconds2
    << " mixed_to_reference=" << mixed_to_reference(r) << '\n';
  if (!r.paired_ratios.empty())
    o << "paired_ratio_mean=" << r.paired_ratio_summary.mean
      << " paired_ratio_variance=" << r.paired_ratio_summary.variance
      << " paired_ratio_min=" << r.paired_ratio_summary.minimum
      << " paired_ratio_max=" << r.paired_ratio_summary.maximum
      << " paired_ratio_median=" << r.paired_ratio_summary.median << '\n';
  o << "converged=" << (r.converged ? "true" : "false")
    << " finite=" << (r.finite ? "true" : "false") << " stopping_policy=" << r.stopping_policy
    << " termination_reason=" << r.termination_reason << '\n';
  o << "command_buffers_per_correction=" << r.command_buffers_per_correction
    << " kernel_dispatches_per_correction=" << r.kernel_dispatches_per_correction << '\n';
  o << "status=" << r.status << " correction_precision=" << r.correction_precision
    << " sweeps=" << r.sweeps << " corrections=" << r.corrections << " off_norm=" << r.off_norm
    << " unresolved=" << (r.unresolved ? "true" : "false")
    << " refined=" << (r.refined ? "true" : "false") << " final_forward_estimate="
    << (r.forward_estimate_history.empty() ? 0.0 : r.forward_estimate_history.back()) << '\n';
  std::fputs(o.str().c_str(), stdout);
}

namespace {
HipCaseResult header_case() {
  HipCaseResult r;
  r.history = {IterationMeasurement{}};
  return r;
}
}  // namespace

void writeSummaryHeader(std::ostream& output) { write_names(output, summary_fields(header_case())); }

void writeIterationsHeader(std::ostream& output) {
  write_names(output, iteration_fields(header_case(), IterationMeasurement{}));
}

void writeCaseCsv(const HipCaseResult& result, std::ostream& summary, std::ostream& iterations) {
  if (result.history.empty()) fail("writeCaseCsv: the case has no history");
  write_values(summary, summary_fields(result));
  for (const IterationMeasurement& m : result.history)
    write_values(iterations, iteration_fields(result, m));
}

}  // namespace msq_hip
