// Prints the summary and iterations CSV of the C++ mirror (hip_runtime.cpp)
// for one hand-filled case result, for the CPU test that compares the columns
// with the reference's schema.  Output: the summary file, a line "--", then
// the iterations file.
#include <iostream>
#include <sstream>

#include "hip_runtime.hpp"

int main() {
  msq_hip::HipCaseResult r;
  r.name = "n64_c100";
  r.basis_precision = "bf16";
  r.correction_precision = "bf16";
  r.status = "converged";
  r.termination_reason = "forward_error_target";
  r.order = 64;
  r.condition = 100.0;
  r.converged = r.finite = true;
  r.corrections = 2;
  r.max_corrections = 4;
  r.warmups = 1;
  r.final_residual = 1e-9;
  r.final_forward_error = 2e-9;
  r.basis_orthogonality = 3e-3;
  r.residual_history = {1e-3, 1e-6, 1e-9};
  r.forward_estimate_history = {2e-3, 2e-6, 2e-9};
  r.seconds = {0.004, 0.002, 0.003};
  std::ostringstream iterations;
  msq_hip::writeSummaryHeader(std::cout);
  msq_hip::writeCaseCsv(r, std::cout, iterations);
  std::cout << "--\n";
  msq_hip::writeIterationsHeader(std::cout);
  std::cout << iterations.str();
  std::cout.flush();
  return 0;
}
