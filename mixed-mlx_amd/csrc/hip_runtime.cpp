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

void fail(const std::string& message) { throw std::runtime_error(message); }

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

// Column-major n x n product C = A B^T in binary64 (host; inputs only).
std::vector<double> gemm_nt(const std::vector<double>& a, const std::vector<double>& b,
                            size_t n) {
  std::vector<double> c(n * n, 0.0);
  for (size_t k = 0; k < n; ++k)
    for (size_t j = 0; j < n; ++j) {
      const double bjk = b[j + k * n];
      const double* ak = a.data() + k * n;
      double* cj = c.data() + j * n;
      for (size_t i = 0; i < n; ++i) cj[i] += ak[i] * bjk;
    }
  return c;
}

// The controlled basis: sin/cos columns, orthonormalised in order.
std::vector<double> controlled_basis(size_t n) {
  std::vector<double> q(n * n);
  for (size_t c = 0; c < n; ++c) {
    double* col = q.data() + c * n;
    for (size_t r = 0; r < n; ++r)
      col[r] = std::sin(0.173 * static_cast<double>((r + 1) * (c + 2))) +
               std::cos(0.117 * static_cast<double>((r + 3) * (c + 1)));
    for (size_t p = 0; p < c; ++p) {
      const double* prior = q.data() + p * n;
      double dot = 0.0;
      for (size_t r = 0; r < n; ++r) dot += prior[r] * col[r];
      for (size_t r = 0; r < n; ++r) col[r] -= dot * prior[r];
    }
    double norm = 0.0;
    for (size_t r = 0; r < n; ++r) norm += col[r] * col[r];
    norm = std::sqrt(norm);
    if (norm <= std::numeric_limits<double>::epsilon()) fail("the controlled basis is singular");
    for (size_t r = 0; r < n; ++r) col[r] /= norm;
  }
  return q;
}

std::vector<double> controlled_product(size_t n, double condition, double power) {
  if (!std::isfinite(condition) || condition < 1.0)
    fail("the SPD condition target must be finite and at least one");
  const std::vector<double> q = controlled_basis(n);
  std::vector<double> scaled = q;
  const double log_cond = std::log(condition);
  for (size_t c = 0; c < n; ++c) {
    const double fraction = static_cast<double>(c) / static_cast<double>(std::max<size_t>(1, n - 1));
    const double s = std::exp(power * log_cond * fraction);
    for (size_t r = 0; r < n; ++r) scaled[r + c * n] *= s;
  }
  return gemm_nt(scaled, q, n);
}

const char* status_name(uint32_t s) {
  switch (s) {
    case MSQ_SOLVE_CONVERGED: return "converged";
    case MSQ_SOLVE_BUDGET_EXHAUSTED: return "budget_exhausted";
    case MSQ_SOLVE_DIVERGED: return "diverged";
    case MSQ_SOLVE_NONFINITE: return "nonfinite";
    case MSQ_SOLVE_INVALID_SPECTRUM: return "invalid_spectrum";
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
HipRuntime::HipRuntime(int device) {
  if (msq_runtime_create(device, &rt_) != MSQ_OK) fail("HIP runtime: " + last_error());
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
    const std::vector<msq_solve_options>& options) {
  if (options.size() != 1 && options.size() != orders.size())
    fail("solveBatched: one options entry, or one per matrix");
  std::vector<msq_solve_report> reports(orders.size());
  check(msq_solve_batched(rt_, static_cast<uint32_t>(orders.size()), orders.data(), matrices,
                          roots, options.data(), static_cast<uint32_t>(options.size()),
                          reports.data()),
        "solveBatched");
  return reports;
}

// -------------------------------------------------------------- the solver
std::vector<double> makeSpdMatrix(size_t order, double condition) {
  return controlled_product(order, condition, 1.0);
}

std::vector<double> makeSpdRoot(size_t order, double condition) {
  return controlled_product(order, condition, 0.5);
}

namespace {

msq_solve_options solve_options(const HipSolverConfiguration& c, const HipProblem& p) {
  msq_solve_options o;
  msq_solve_options_init(&o);
  o.basis_precision = static_cast<int32_t>(c.basis_precision);
  o.correction_precision = static_cast<int32_t>(c.correction_precision);
  o.jacobi.fixed_sweeps = c.fixed_sweeps;
  o.jacobi.max_sweeps = 30;
  o.jacobi.off_tol = static_cast<float>(c.off_tolerance);
  o.jacobi.off_tol_lambda = 0.0f;
  o.jacobi.stagnation = 0.5f;
  o.jacobi.floor_tol = 1e-5f;
  o.jacobi.correction_rate = c.correction_rate;
  o.min_corrections = 0;
  o.max_corrections = static_cast<uint32_t>(
      std::min<size_t>(p.max_iterations, MSQ_SOLVE_HISTORY - 1));
  o.residual_tolerance = p.tolerance;
  o.forward_tolerance = c.forward_tolerance;
  return o;
}

struct CaseBuffers {
  std::vector<uint32_t> orders;
  Batch batch;
  DeviceBuffer a, x;
};

double frobenius(const std::vector<double>& v) {
  double s = 0.0;
  for (double x : v) s += x * x;
  return std::sqrt(s);
}

}  // namespace

HipCaseResult runTimedHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                              const HipProblem& problem, size_t warmups, size_t evaluations) {
  const size_t n = problem.order;
  if (n < 1 || n > 65535) fail("order out of range");
  const std::vector<double> generated =
      problem.prepared_matrix ? std::vector<double>() : makeSpdMatrix(n, problem.condition);
  const std::vector<double>& A = problem.prepared_matrix ? *problem.prepared_matrix : generated;
  if (A.size() != n * n) fail("prepared matrix is not order x order");
  std::vector<double> reference;
  if (problem.supplied_reference) {
    reference = *problem.supplied_reference;
  } else if (!problem.prepared_matrix) {
    reference = makeSpdRoot(n, problem.condition);
  }

  CaseBuffers b;
  b.orders = {static_cast<uint32_t>(n)};
  b.batch = layout(b.orders);
  const size_t ld = b.batch.lds[0];
  std::vector<double> tile(ld * ld, 0.0);
  for (size_t j = 0; j < n; ++j)
    for (size_t i = 0; i < n; ++i) tile[i + j * ld] = A[i + j * n];
  b.a = runtime.makeDeviceBuffer(tile.size() * sizeof(double));
  b.x = runtime.makeDeviceBuffer(tile.size() * sizeof(double));
  b.a.upload(tile.data(), tile.size() * sizeof(double));
  const std::vector<msq_solve_options> options = {solve_options(configuration, problem)};

  HipCaseResult r;
  std::vector<msq_solve_report> reports;
  for (size_t w = 0; w < warmups; ++w)
    reports = runtime.solveBatched(b.orders, static_cast<const double*>(b.a.data()),
                                   static_cast<double*>(b.x.data()), options);
  for (size_t e = 0; e < std::max<size_t>(evaluations, 1); ++e) {
    const auto t0 = std::chrono::steady_clock::now();
    reports = runtime.solveBatched(b.orders, static_cast<const double*>(b.a.data()),
                                   static_cast<double*>(b.x.data()), options);
    runtime.synchronize();
    r.seconds.push_back(
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count());
  }
  const msq_solve_report& rep = reports[0];
  b.x.download(tile.data(), tile.size() * sizeof(double));
  std::vector<double> X(n * n);
  for (size_t j = 0; j < n; ++j)
    for (size_t i = 0; i < n; ++i) X[i + j * n] = tile[i + j * ld];

  r.name = "hip_spd_sqrt";
  r.warmups = warmups;
  r.basis_precision = precisionName(configuration.basis_precision);
  r.correction_precision = precisionName(configuration.correction_precision);
  r.order = n;
  r.condition = problem.condition;
  r.status = status_name(rep.status);
  r.converged = rep.status == MSQ_SOLVE_CONVERGED;
  r.sweeps = rep.sweeps;
  r.corrections = rep.corrections;
  r.off_norm = rep.off_norm;
  r.final_residual = rep.final_residual;
  r.final_forward_estimate = rep.final_forward_estimate;
  for (uint32_t k = 0; k <= rep.corrections && k < MSQ_SOLVE_HISTORY; ++k) {
    r.residual_history.push_back(rep.residuals[k]);
    r.forward_estimate_history.push_back(rep.forward_estimates[k]);
  }
  r.basis_orthogonality = rep.basis_orthogonality;
  r.precision_escalations = rep.precision_escalations;
  r.max_corrections = problem.max_iterations;
  r.finite = std::all_of(X.begin(), X.end(), [](double v) { return std::isfinite(v); });
  double asym = 0.0;
  for (size_t j = 0; j < n; ++j)
    for (size_t i = 0; i < n; ++i) {
      const double d = X[i + j * n] - X[j + i * n];
      asym += d * d;
    }
  r.symmetry_error = std::sqrt(asym) / std::max(frobenius(X), std::numeric_limits<double>::min());
  if (!reference.empty()) {
    std::vector<double> d(n * n);
    for (size_t k = 0; k < n * n; ++k) d[k] = X[k] - reference[k];
    r.final_forward_error = frobenius(d) / frobenius(reference);
  }
  r.termination_reason = r.converged ? "target_met" : r.status;
  if (problem.enforce_policy) {
    if (!r.converged || !r.finite) fail("HIP solve did not meet its target: " + r.status);
    if (r.symmetry_error > 1e-12) fail("HIP solve: root not symmetric");
  }
  return r;
}

HipCaseResult runHipCase(HipRuntime& runtime, const HipSolverConfiguration& configuration,
                         const HipProblem& problem) {
  return runTimedHipCase(runtime, configuration, problem, 0, 1);
}

void printResult(const HipCaseResult& r) {
  std::printf("name=%s\nsolver_backend=%s\nbasis_precision=%s\ncorrection_precision=%s\n",
              r.name.c_str(), r.solver_backend.c_str(), r.basis_precision.c_str(),
              r.correction_precision.c_str());
  std::printf("order=%zu\ncondition=%.6g\nstatus=%s\ntermination_reason=%s\n", r.order,
              r.condition, r.status.c_str(), r.termination_reason.c_str());
  std::printf("converged=%d\nfinite=%d\nsweeps=%zu\ncorrections=%zu\n", r.converged ? 1 : 0,
              r.finite ? 1 : 0, r.sweeps, r.corrections);
  std::printf("off_norm=%.6e\nfinal_residual=%.6e\nfinal_forward_estimate=%.6e\n", r.off_norm,
              r.final_residual, r.final_forward_estimate);
  std::printf("final_forward_error=%.6e\nsymmetry_error=%.6e\n", r.final_forward_error,
              r.symmetry_error);
  std::printf("residual_history=");
  for (size_t k = 0; k < r.residual_history.size(); ++k)
    std::printf(k ? ",%.6e" : "%.6e", r.residual_history[k]);
  std::printf("\n");
  if (!r.seconds.empty()) {
    std::vector<double> s = r.seconds;
    std::sort(s.begin(), s.end());
    std::printf("evaluations=%zu\nmedian_seconds=%.6e\nmin_seconds=%.6e\n", s.size(),
                s[s.size() / 2], s.front());
  }
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

struct TimingSummary {
  double min = 0.0, max = 0.0, median = 0.0, mean = 0.0, variance = 0.0;
};
TimingSummary summarize(std::vector<double> s) {
  TimingSummary t;
  if (s.empty()) return t;
  std::sort(s.begin(), s.end());
  t.min = s.front();
  t.max = s.back();
  t.median = s.size() % 2 ? s[s.size() / 2] : 0.5 * (s[s.size() / 2 - 1] + s[s.size() / 2]);
  for (double x : s) t.mean += x;
  t.mean /= static_cast<double>(s.size());
  for (double x : s) t.variance += (x - t.mean) * (x - t.mean);
  t.variance /= static_cast<double>(s.size());
  return t;
}

// The columns shared by both files, up to kernel_path (src/reporting.cpp:
// 124-131 and :307-311).
std::vector<Field> case_fields(const HipCaseResult& r) {
  return {
      {"paper_test_set", "controlled"},
      {"matrix_family", "conditioned_spd"},
      {"matrix_type", "symmetric"},
      {"solver_backend", r.solver_backend},
      {"rng_seed", "0"},
      {"rounding_mode", "nearest"},
      {"rounding_seed", "0"},
      {"diagonal_shift", "0"},
      {"shift_rule", "not_applicable"},
      {"case", r.name},
      {"precision", r.basis_precision},
      {"basis_precision", r.basis_precision},
      {"spectrum_precision", "fp32"},
      {"correction_mode", r.correction_precision + "_frozen_factor"},
      {"precision_policy", "forward_error_target"},
      {"stopping_policy", "forward_estimate_or_" + num(r.max_corrections) + "_corrections"},
      {"dispatch_policy", "not_applicable"},
      {"termination_reason", r.termination_reason},
      {"selection_policy", "fixed"},
      {"precision_attempts", "1"},
      {"execution_model", "gpu_resident_batched"},
      {"kernel_path", "hip_mfma"},
  };
}

// basis stage .. basis_reorthogonalization_seconds (both files).
std::vector<Field> basis_fields(const HipCaseResult& r) {
  const bool low_basis = r.basis_precision != "fp32";
  const size_t n2 = r.order * r.order;
  const Precision basis = parsePrecision(r.basis_precision);
  const size_t low_bytes = msq_storage_bytes(static_cast<int>(basis), n2) +
                           msq_scale_bytes(static_cast<int>(basis), n2);
  const size_t fp32_bytes = 4 * n2;
  return {
      {"selected_basis_stage",
       !low_basis ? "fp32"
                  : (r.precision_escalations ? r.basis_precision + "_rejected_to_fp32"
                                             : r.basis_precision)},
      {"selected_spectrum_stage", "fp32"},
      {"correction_stage_history", r.correction_precision},
      {"precision_escalations", num(r.precision_escalations)},
      {"low_factor_corrections", num(r.correction_precision != "fp32" ? r.corrections : 0)},
      {"low_factor_basis_bytes", num(low_bytes)},
      {"fp32_basis_bytes", num(fp32_bytes)},
      {"basis_storage_ratio",
       num(fp32_bytes ? static_cast<double>(low_bytes) / static_cast<double>(fp32_bytes) : 0.0)},
      {"quantized_basis_orthogonality",
       low_basis ? num(r.basis_orthogonality) : std::string("nan")},
      {"factor_quantization_seconds", "0"},
      {"basis_reorthogonalization_seconds", "0"},
  };
}

std::vector<Field> reference_timing_fields() {
  return {{"reference_timing_min_seconds", "0"},
          {"reference_timing_max_seconds", "0"},
          {"reference_timing_median_seconds", "0"},
          {"reference_timing_mean_seconds", "0"},
          {"reference_timing_variance_seconds2", "0"}};
}
std::vector<Field> paired_fields() {
  return {{"paired_ratio_mean", "0"},
          {"paired_ratio_variance", "0"},
          {"paired_ratio_min", "0"},
          {"paired_ratio_max", "0"},
          {"paired_ratio_median", "0"}};
}
// Phase times (factorization .. mixed_total): only the total is measured.
std::vector<Field> phase_fields(const TimingSummary& t) {
  return {{"factorization_seconds", "0"},
          {"initial_root_seconds", "0"},
          {"single_seconds", "0"},
          {"residual_seconds", "0"},
          {"correction_system_seconds", "0"},
          {"updates_and_others_seconds", "0"},
          {"mixed_total_seconds", num(t.median)}};
}

void append(std::vector<Field>& out, std::vector<Field> more) {
  out.insert(out.end(), std::make_move_iterator(more.begin()),
             std::make_move_iterator(more.end()));
}

std::vector<Field> summary_fields(const HipCaseResult& r) {
  const TimingSummary t = summarize(r.seconds);
  const size_t evaluations = r.seconds.size();
  std::vector<Field> f = case_fields(r);
  append(f, {{"command_buffers_per_correction", "0"}, {"kernel_dispatches_per_correction", "0"}});
  append(f, basis_fields(r));
  append(f, {{"warmups", num(r.warmups)},
             {"evaluations", num(evaluations)},
             {"timing_evaluations", num(evaluations)},
             {"timing_min_seconds", num(t.min)},
             {"timing_max_seconds", num(t.max)},
             {"timing_median_seconds", num(t.median)},
             {"timing_mean_seconds", num(t.mean)},
             {"timing_variance_seconds2", num(t.variance)}});
  append(f, reference_timing_fields());
  size_t best = 0;
  for (size_t k = 1; k < r.residual_history.size(); ++k)
    if (r.residual_history[k] < r.residual_history[best]) best = k;
  const double initial_residual = r.residual_history.empty() ? 0.0 : r.residual_history.front();
  const double initial_forward =
      r.forward_estimate_history.empty() ? 0.0 : r.forward_estimate_history.front();
  append(f, {{"order", num(r.order)},
             {"condition", num(r.condition)},
             {"iterations", num(r.corrections)},
             {"best_iteration", num(best)},
             {"best_residual", num(r.residual_history.empty() ? 0.0 : r.residual_history[best])},
             {"best_forward_error", num(r.final_forward_error)},
             {"stagnation_steps", "0"},
             {"converged", r.converged ? "1" : "0"},
             {"finite", r.finite ? "1" : "0"},
             {"status", r.status},
             {"initial_residual", num(initial_residual)},
             {"final_residual", num(r.final_residual)},
             {"initial_forward_error", num(initial_forward)},
             {"final_forward_error", num(r.final_forward_error)},
             {"symmetry_error", num(r.symmetry_error)},
             {"reference_residual", "0"},
             {"frozen_factor_error", "0"},
             {"frozen_factor_relative_error", "0"},
             {"frozen_basis_orthogonality", "0"},
             {"frozen_sylvester_inverse_norm", "0"},
             {"frozen_contraction_bound", "0"},
             {"certificate_usable", "0"},
             {"reference_seconds", "0"},
             {"low_eigen_seconds", "0"},
             {"initial_gpu_seconds", "0"},
             {"residual_fp64_seconds", "0"},
             {"correction_gpu_seconds", "0"},
             {"update_fp64_seconds", "0"}});
  append(f, phase_fields(t));
  append(f, {{"low_eigen_fraction", "0"},
             {"initial_gpu_fraction", "0"},
             {"residual_fraction", "0"},
             {"correction_gpu_fraction", "0"},
             {"update_fraction", "0"}});
  append(f, paired_fields());
  append(f, {{"mixed_to_reference", "0"}});
  return f;
}

std::vector<Field> iteration_fields(const HipCaseResult& r, size_t k) {
  const TimingSummary t = summarize(r.seconds);
  const size_t evaluations = r.seconds.size();
  std::vector<Field> f = case_fields(r);
  append(f, basis_fields(r));
  append(f, {{"order", num(r.order)},
             {"condition", num(r.condition)},
             {"warmups", num(r.warmups)},
             {"evaluations", num(evaluations)},
             {"timing_evaluations", num(evaluations)}});
  append(f, reference_timing_fields());
  append(f, paired_fields());
  append(f, phase_fields(t));
  append(f, {{"iteration", num(k)},
             {"residual", num(k < r.residual_history.size() ? r.residual_history[k] : 0.0)},
             {"forward_error",
              num(k < r.forward_estimate_history.size() ? r.forward_estimate_history[k] : 0.0)}});
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

HipCaseResult header_case() {
  HipCaseResult r;
  r.basis_precision = r.correction_precision = "fp32";
  r.residual_history = {0.0};
  r.forward_estimate_history = {0.0};
  return r;
}

}  // namespace

void writeSummaryHeader(std::ostream& output) { write_names(output, summary_fields(header_case())); }

void writeIterationsHeader(std::ostream& output) {
  write_names(output, iteration_fields(header_case(), 0));
}

void writeCaseCsv(const HipCaseResult& result, std::ostream& summary, std::ostream& iterations) {
  write_values(summary, summary_fields(result));
  for (size_t k = 0; k < result.residual_history.size(); ++k)
    write_values(iterations, iteration_fields(result, k));
}

}  // namespace msq_hip
