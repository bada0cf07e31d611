// Command-line driver of the HIP backend, shaped like the reference's
// application for its GPU backend (runApplication, src/application.cpp:47-345;
// runMetal / runSweep, src/metal_backend.cpp:18-119): the same options, a
// single case (controlled matrix of the given order at condition 4, or a
// paper test set; enforced; optionally adaptive) or a timed sweep writing
// summary.csv and iterations.csv in the reference's schema.
//
//   msq_cli --solver-backend hip 256
//   msq_cli --paper-test-set VII --precision fp32
//   msq_cli --sweep out --orders 64,256 --conditions 1e2,1e4 --warmups 2 --evaluations 5
//
// Options of this backend only: --correction-precision, --jacobi-sweeps,
// --off-tolerance, --correction-rate, --forward-tolerance, --basis-tolerance,
// --resolution, --refine-unresolved, --max-iterations, --tolerance,
// --condition, --no-enforce, --device.
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

using msq_hip::fail;

void usage() {
  std::printf(
      "usage: msq_cli [options] [matrix order | sweep output directory]\n"
      "\n"
      "  --solver-backend hip\n"
      "  --precision <fp32|bf16|fp8_e4m3fn|fp8_e5m2|mxfp8_e4m3fn|mxfp8_e5m2|fp4_e2m1|\n"
      "               mxfp4_e2m1|fp8_e4m3fn_log_s|fp4_e2m1_log_s>  (basis fp32 or bf16,\n"
      "               spectrum the format)\n"
      "  --basis-precision <format>  --spectrum-precision <format, log variants too>\n"
      "  --correction-mode fp32_reorthogonalized|native_quantized\n"
      "  --native-dispatch fused|per_stage  --native-stopping fixed|stagnation\n"
      "  --native-stagnation-window <n>  --native-min-relative-improvement <x>\n"
      "  --kernel-path tiled|legacy\n"
      "  --rounding nearest|stochastic  --rounding-seed <uint32>\n"
      "  --orders <a,b,c> --conditions <a,b,c> --sweep [output directory]\n"
      "  --warmups <n> --evaluations <n>\n"
      "  --paper-test-set <V|VI|VII> --paper-order <n>\n"
      "  --adaptive\n"
      "  --help\n"
      "backend options:\n"
      "  --correction-precision <fp32|bf16|fp8_e4m3fn_scaled|mxfp4_e2m1>  correction products\n"
      "  --jacobi-sweeps <S> (8: the reference's fixed budget; 0: off-norm stop)\n"
      "  --off-tolerance <t> --correction-rate <r> --forward-tolerance <t>\n"
      "  --basis-tolerance <t> --resolution <r> --refine-unresolved\n"
      "  --max-iterations <k> --tolerance <t> --condition <c> --no-enforce --device <d>\n"
      "  --output-dir <dir> (same as --sweep <dir>)\n");
}

template <typename T>
std::vector<T> parse_list(const std::string& text, const char* what) {
  std::vector<T> out;
  std::stringstream ss(text);
  std::string item;
  while (std::getline(ss, item, ',')) {
    std::stringstream is(item);
    T v{};
    if (!(is >> v) || !is.eof()) fail(std::string(what) + " expects comma-separated numbers");
    out.push_back(v);
  }
  if (out.empty()) fail(std::string(what) + " is empty");
  return out;
}

template <typename T>
T parse_one(const std::string& text, const char* what) {
  return parse_list<T>(text, what).at(0);
}

size_t parse_positive(const std::string& text, const char* what) {
  if (text.empty() || text[0] == '-') fail(std::string(what) + " requires a positive integer.");
  const size_t v = parse_one<size_t>(text, what);
  if (v == 0) fail(std::string(what) + " requires a positive integer.");
  return v;
}

// basisPrecision (src/configuration.cpp:120-123).
msq_hip::Precision basis_for(msq_hip::Precision p) {
  return p == msq_hip::Precision::Fp32 ? msq_hip::Precision::Fp32 : msq_hip::Precision::Bf16;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    msq_hip::HipSolverConfiguration config;
    std::vector<size_t> orders = {16, 32, 64, 96, 128, 192, 256};
    std::vector<double> conditions = {4.0, 1.0e2, 1.0e4, 1.0e6};
    size_t warmups = 50, evaluations = 100;
    bool sweep = false, adaptive = false, enforce = true, seed_selected = false;
    msq_hip::RoundingMode rounding = msq_hip::RoundingMode::Nearest;
    uint32_t rounding_seed = 0;
    msq_hip::PaperTestSet paper = msq_hip::PaperTestSet::None;
    size_t paper_order = 0;
    double condition = 4.0;
    long max_iterations = -1;
    double tolerance = 1.0e-12;
    int device = 0;
    std::vector<std::string> positional;
    for (int i = 1; i < argc; ++i) {
      const std::string a = argv[i];
      auto take = [&](const char* msg) -> std::string {
        if (i + 1 >= argc) fail(msg);
        return argv[++i];
      };
      if (a == "--help" || a == "-h") {
        usage();
        return EXIT_SUCCESS;
      } else if (a == "--solver-backend") {
        const std::string b = take("--solver-backend requires hip.");
        if (b != "hip") fail("this driver runs the hip backend only (got '" + b + "')");
      } else if (a == "--paper-test-set") {
        paper = msq_hip::paperTestSetForCliName(
            take("--paper-test-set requires I, II, III, IV, V, VI, or VII."));
      } else if (a == "--paper-order") {
        paper_order = parse_positive(take("--paper-order requires a positive integer."),
                                     "--paper-order");
      } else if (a == "--precision") {
        const msq_hip::Precision p = msq_hip::parsePrecision(take("--precision requires a format."));
        config.basis_precision = basis_for(p);
        config.spectrum_precision = p;
      } else if (a == "--basis-precision") {
        config.basis_precision = msq_hip::parsePrecision(take("--basis-precision requires a format."));
      } else if (a == "--spectrum-precision") {
        config.spectrum_precision =
            msq_hip::parsePrecision(take("--spectrum-precision requires a format."));
      } else if (a == "--correction-precision") {
        config.correction_precision =
            msq_hip::parsePrecision(take("--correction-precision requires a format."));
      } else if (a == "--rounding") {
        rounding = msq_hip::parseRoundingMode(take("--rounding requires nearest or stochastic."));
      } else if (a == "--rounding-seed") {
        const unsigned long long v = parse_one<unsigned long long>(
            take("--rounding-seed requires an unsigned 32-bit integer."), "--rounding-seed");
        if (v > 0xffffffffull) fail("--rounding-seed requires an unsigned 32-bit integer.");
        rounding_seed = static_cast<uint32_t>(v);
        seed_selected = true;
      } else if (a == "--kernel-path") {
        config.kernel_path = msq_hip::parseKernelPath(take("--kernel-path requires tiled or legacy."));
      } else if (a == "--correction-mode") {
        config.correction_mode = msq_hip::parseCorrectionMode(
            take("--correction-mode requires fp32_reorthogonalized or native_quantized."));
      } else if (a == "--native-dispatch") {
        config.dispatch_policy =
            msq_hip::parseDispatchPolicy(take("--native-dispatch requires fused or per_stage."));
      } else if (a == "--native-stopping") {
        config.stopping_policy =
            msq_hip::parseStoppingPolicy(take("--native-stopping requires fixed or stagnation."));
      } else if (a == "--native-stagnation-window") {
        config.stagnation_window = parse_positive(
            take("--native-stagnation-window requires a positive integer."),
            "--native-stagnation-window");
      } else if (a == "--native-min-relative-improvement") {
        config.minimum_relative_improvement =
            parse_one<double>(take("--native-min-relative-improvement requires a number."),
                              "--native-min-relative-improvement");
      } else if (a == "--orders") {
        orders = parse_list<size_t>(take("--orders requires integers"), "--orders");
      } else if (a == "--conditions") {
        conditions = parse_list<double>(take("--conditions requires reals"), "--conditions");
      } else if (a == "--warmups") {
        warmups = parse_one<size_t>(take("--warmups requires a count"), "--warmups");
      } else if (a == "--evaluations") {
        evaluations = parse_positive(take("--evaluations requires a positive integer"),
                                     "--evaluations");
      } else if (a == "--sweep") {
        sweep = true;
      } else if (a == "--output-dir") {
        sweep = true;
        positional.push_back(take("--output-dir requires a directory"));
      } else if (a == "--adaptive") {
        adaptive = true;
      } else if (a == "--jacobi-sweeps") {
        config.fixed_sweeps = parse_one<uint32_t>(take("--jacobi-sweeps requires S"), "--jacobi-sweeps");
      } else if (a == "--off-tolerance") {
        config.off_tolerance = parse_one<double>(take("--off-tolerance requires T"), "--off-tolerance");
      } else if (a == "--correction-rate") {
        config.correction_rate = parse_one<float>(take("--correction-rate requires R"), "--correction-rate");
      } else if (a == "--forward-tolerance") {
        config.forward_tolerance =
            parse_one<double>(take("--forward-tolerance requires T"), "--forward-tolerance");
      } else if (a == "--basis-tolerance") {
        config.basis_tolerance =
            parse_one<double>(take("--basis-tolerance requires T"), "--basis-tolerance");
      } else if (a == "--resolution") {
        config.resolution = parse_one<float>(take("--resolution requires R"), "--resolution");
      } else if (a == "--refine-unresolved") {
        config.refine_unresolved = true;
      } else if (a == "--max-iterations") {
        max_iterations =
            static_cast<long>(parse_one<size_t>(take("--max-iterations requires K"), "--max-iterations"));
      } else if (a == "--tolerance") {
        tolerance = parse_one<double>(take("--tolerance requires T"), "--tolerance");
      } else if (a == "--condition") {
        condition = parse_one<double>(take("--condition requires C"), "--condition");
      } else if (a == "--no-enforce") {
        enforce = false;
      } else if (a == "--enforce") {
        enforce = true;
      } else if (a == "--device") {
        device = parse_one<int>(take("--device requires D"), "--device");
      } else if (!a.empty() && a[0] == '-') {
        fail("unknown argument " + a);
      } else {
        positional.push_back(a);
      }
    }
    const bool has_paper = paper != msq_hip::PaperTestSet::None;
    if (has_paper && !msq_hip::paperTestSetIsSymmetric(paper))
      fail("The HIP backend does not support nonsymmetric paper test sets.");
    if (has_paper && (config.basis_precision != msq_hip::Precision::Fp32 ||
                      config.spectrum_precision != msq_hip::Precision::Fp32 ||
                      config.correction_mode != msq_hip::CorrectionMode::Fp32Reorthogonalized))
      fail("Paper test sets require the FP32 mixed correction mode.");
    if (seed_selected && rounding != msq_hip::RoundingMode::Stochastic)
      fail("--rounding-seed requires --rounding stochastic.");
    if (rounding == msq_hip::RoundingMode::Stochastic &&
        config.correction_mode == msq_hip::CorrectionMode::NativeQuantized)
      fail("Stochastic rounding supports frozen-factor quantization with FP32 correction. It "
           "does not support native quantized correction.");
    if (adaptive && config.correction_mode == msq_hip::CorrectionMode::NativeQuantized)
      fail("--adaptive does not support native quantized correction.");
    if (adaptive && sweep) fail("--adaptive does not support sweep mode.");
    if (positional.size() > 1)
      fail(sweep ? "Sweep mode accepts at most one output directory."
                 : "Expected at most one matrix order.");

    msq_hip::PreparedMatrix paper_matrix;
    const msq_hip::PreparedMatrix* paper_pointer = nullptr;
    if (has_paper) {
      paper_matrix = msq_hip::preparePaperMatrix(
          paper, paper_order ? paper_order : msq_hip::paperTestSetOrder(paper));
      paper_pointer = &paper_matrix;
    }
    msq_hip::HipRuntime runtime(device, rounding, rounding_seed);

    if (!sweep) {
      msq_hip::HipProblem problem;
      problem.order = paper_pointer ? paper_pointer->order
                                    : (positional.empty() ? 32 : parse_positive(positional[0],
                                                                                "matrix order"));
      problem.condition = paper_pointer ? paper_pointer->condition : condition;
      problem.max_iterations =
          max_iterations >= 0 ? static_cast<size_t>(max_iterations) : (paper_pointer ? 12 : 8);
      problem.tolerance = tolerance;
      problem.enforce_policy = enforce;
      problem.prepared_matrix = paper_pointer;
      const msq_hip::HipCaseResult r = adaptive
                                           ? msq_hip::runAdaptiveHipCase(runtime, config, problem)
                                           : msq_hip::runHipCase(runtime, config, problem);
      msq_hip::printResult(r);
      std::fflush(stdout);
      return EXIT_SUCCESS;
    }

    const std::filesystem::path dir = positional.empty() ? "hip_benchmark" : positional[0];
    std::filesystem::create_directories(dir);
    std::ofstream summary(dir / "summary.csv"), iterations(dir / "iterations.csv");
    if (!summary || !iterations) fail("Cannot create benchmark CSV files.");
    msq_hip::writeSummaryHeader(summary);
    msq_hip::writeIterationsHeader(iterations);
    size_t failures = 0;
    auto run = [&](const msq_hip::HipProblem& problem) {
      const msq_hip::HipCaseResult r =
          msq_hip::runTimedHipCase(runtime, config, problem, warmups, evaluations);
      msq_hip::printResult(r);
      msq_hip::writeCaseCsv(r, summary, iterations);
      if (!r.converged || !r.finite) ++failures;
      std::printf("\n");
      std::fflush(stdout);
    };
    msq_hip::HipProblem problem;
    problem.max_iterations = max_iterations >= 0 ? static_cast<size_t>(max_iterations) : 12;
    problem.tolerance = tolerance;
    problem.enforce_policy = false;
    if (paper_pointer) {
      problem.order = paper_pointer->order;
      problem.condition = paper_pointer->condition;
      problem.prepared_matrix = paper_pointer;
      run(problem);
    } else {
      for (size_t n : orders)
        for (double c : conditions) {
          problem.order = n;
          problem.condition = c;
          run(problem);
        }
    }
    summary.close();
    iterations.close();
    if (!summary || !iterations) fail("writing the CSV files failed");
    std::printf("benchmark_failures=%zu\nsummary_csv=%s\niterations_csv=%s\n", failures,
                (dir / "summary.csv").c_str(), (dir / "iterations.csv").c_str());
    return EXIT_SUCCESS;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return EXIT_FAILURE;
  }
}
