// Prints the summary and iterations CSV of the C++ mirror (hip_runtime.cpp)
// for one hand-filled case result, for the CPU test that compares the columns
// with the reference's schema.  Output: the summary file, a line "--", then
// the iterations file.
#include <cmath>
#include <iostream>
#include <limits>
#include <sstream>

#include "hip_runtime.hpp"

int main() {
  msq_hip::HipCaseResult r;
  r.name = "spd_n64_k100_bf16_fp32_reorthogonalized";
  r.precision = r.basis_precision = r.spectrum_precision = "bf16";
  r.correction_mode = "fp32_reorthogonalized";
  r.status = r.termination_reason = "converged";
  r.selected_basis_stage = "bf16";
  r.order = 64;
  r.condition = 100.0;
  r.converged = r.finite = true;
  r.certificate_usable = true;
  r.frozen_sylvester_inverse_norm = 0.5;
  r.frozen_factor_error = 1e-4;
  r.frozen_contraction_bound = 1.5e-4;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  r.history = {{0, 1e-3, 2e-3}, {1, 1e-6, nan}, {2, 1e-9, 2e-9}};
  r.best_iteration = 2;
  r.best_residual = 1e-9;
  r.final_residual = 1e-9;
  r.final_forward_error = 2e-9;
  r.timing = {1, 3, 0.002, 0.004, 0.003, 0.003, 1e-6};
  r.reference_timing = {1, 3, 0.004, 0.008, 0.006, 0.006, 4e-6};
  r.paired_ratios = {0.5, 0.5, 0.5};
  r.paired_ratio_summary = {0.5, 0.0, 0.5, 0.5, 0.5};
  r.times.mixed_total_seconds = 0.003;
  r.times.reference_seconds = 0.006;
  std::ostringstream iterations;
  msq_hip::writeSummaryHeader(std::cout);
  msq_hip::writeCaseCsv(r, std::cout, iterations);
  std::cout << "--\n";
  msq_hip::writeIterationsHeader(std::cout);
  std::cout << iterations.str();
  std::cout.flush();
  return 0;
}
