// Command-line driver of the HIP backend, shaped like the reference's
// application (src/application.cpp:79-215: --solver-backend, --precision,
// --basis-precision, --orders, --conditions, --warmups, --evaluations) and
// printing key=value results like its reporting (src/reporting.cpp:22).
// With --output-dir it also writes summary.csv and iterations.csv in the
// reference's column schema (src/benchmark.cpp:105-134).
//
//   msq_cli --solver-backend hip --orders 64,1024 --conditions 100,1e4 \
//           --basis-precision bf16 --correction-precision bf16 --evaluations 3
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <filesystem>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "hip_runtime.hpp"

namespace {

void usage() {
  std::printf(
      "usage: msq_cli [--solver-backend hip] [--orders N[,N...]] [--conditions C[,C...]]\n"
      "               [--precision P] [--basis-precision P] [--correction-precision P]\n"
      "               [--jacobi-sweeps S] [--off-tolerance T] [--correction-rate R]\n"
      "               [--max-iterations K] [--tolerance T] [--warmups W] [--evaluations E]\n"
      "               [--enforce] [--device D] [--output-dir DIR]\n"
      "precisions: fp32 bf16 fp8_e4m3fn_scaled mxfp4_e2m1 (products); S = 8 is the\n"
      "reference's fixed budget, 0 stops at off(A) <= T ||A||_F.\n");
}

template <typename T>
std::vector<T> parse_list(const std::string& text, const char* what) {
  std::vector<T> out;
  std::stringstream ss(text);
  std::string item;
  while (std::getline(ss, item, ',')) {
    std::stringstream is(item);
    T v{};
    if (!(is >> v)) msq_hip::fail(std::string(what) + " expects comma-separated numbers");
    out.push_back(v);
  }
  if (out.empty()) msq_hip::fail(std::string(what) + " is empty");
  return out;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    std::vector<size_t> orders = {64};
    std::vector<double> conditions = {100.0};
    msq_hip::HipSolverConfiguration config;
    msq_hip::HipProblem problem;
    size_t warmups = 0, evaluations = 1;
    int device = 0;
    std::string output_dir;
    for (int i = 1; i < argc; ++i) {
      const std::string a = argv[i];
      auto take = [&](const char* msg) -> std::string {
        if (i + 1 >= argc) msq_hip::fail(msg);
        return argv[++i];
      };
      if (a == "--help" || a == "-h") {
        usage();
        return EXIT_SUCCESS;
      } else if (a == "--solver-backend") {
        const std::string b = take("--solver-backend requires hip");
        if (b != "hip") msq_hip::fail("this driver runs the hip backend only");
      } else if (a == "--orders") {
        orders = parse_list<size_t>(take("--orders requires integers"), "--orders");
      } else if (a == "--conditions") {
        conditions = parse_list<double>(take("--conditions requires reals"), "--conditions");
      } else if (a == "--precision") {
        config.basis_precision = config.correction_precision =
            msq_hip::parsePrecision(take("--precision requires a precision"));
      } else if (a == "--basis-precision") {
        config.basis_precision = msq_hip::parsePrecision(take("--basis-precision requires a precision"));
      } else if (a == "--correction-precision") {
        config.correction_precision =
            msq_hip::parsePrecision(take("--correction-precision requires a precision"));
      } else if (a == "--jacobi-sweeps") {
        config.fixed_sweeps = static_cast<uint32_t>(std::stoul(take("--jacobi-sweeps requires S")));
      } else if (a == "--off-tolerance") {
        config.off_tolerance = std::stod(take("--off-tolerance requires T"));
      } else if (a == "--correction-rate") {
        config.correction_rate = std::stof(take("--correction-rate requires R"));
      } else if (a == "--max-iterations") {
        problem.max_iterations = std::stoul(take("--max-iterations requires K"));
      } else if (a == "--tolerance") {
        problem.tolerance = std::stod(take("--tolerance requires T"));
      } else if (a == "--warmups") {
        warmups = std::stoul(take("--warmups requires W"));
      } else if (a == "--evaluations") {
        evaluations = std::stoul(take("--evaluations requires E"));
      } else if (a == "--enforce") {
        problem.enforce_policy = true;
      } else if (a == "--device") {
        device = std::stoi(take("--device requires D"));
      } else if (a == "--output-dir") {
        output_dir = take("--output-dir requires a directory");
      } else {
        msq_hip::fail("unknown argument " + a);
      }
    }
    std::ofstream summary, iterations;
    std::filesystem::path summary_path, iterations_path;
    if (!output_dir.empty()) {
      std::filesystem::create_directories(output_dir);
      summary_path = std::filesystem::path(output_dir) / "summary.csv";
      iterations_path = std::filesystem::path(output_dir) / "iterations.csv";
      summary.open(summary_path);
      iterations.open(iterations_path);
      if (!summary || !iterations) msq_hip::fail("cannot write CSV files in " + output_dir);
      msq_hip::writeSummaryHeader(summary);
      msq_hip::writeIterationsHeader(iterations);
    }
    msq_hip::HipRuntime runtime(device);
    for (size_t n : orders)
      for (double c : conditions) {
        problem.order = n;
        problem.condition = c;
        const msq_hip::HipCaseResult r =
            msq_hip::runTimedHipCase(runtime, config, problem, warmups, evaluations);
        msq_hip::printResult(r);
        if (summary.is_open()) msq_hip::writeCaseCsv(r, summary, iterations);
        std::printf("\n");
        std::fflush(stdout);
      }
    if (summary.is_open()) {
      summary.close();
      iterations.close();
      if (!summary || !iterations) msq_hip::fail("writing the CSV files failed");
      std::printf("summary_csv=%s\niterations_csv=%s\n", summary_path.c_str(),
                  iterations_path.c_str());
    }
    return EXIT_SUCCESS;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return EXIT_FAILURE;
  }
}
