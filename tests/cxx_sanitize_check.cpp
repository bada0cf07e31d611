// Host code of the library and of the C++ mirror under AddressSanitizer and
// UndefinedBehaviorSanitizer, on the CPU (no GPU calls succeed here): the C
// ABI's host-side helpers and argument checks (every entry point with a null
// runtime or null buffers returns an error and a message), and the mirror's
// host functions (paper test sets, the FP64 reference root, timing
// summaries, names, the CSV writers).  Prints "ok" on success; a sanitizer
// finding aborts with its report.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <limits>
#include <sstream>
#include <string>
#include <vector>

#include "hip_runtime.hpp"

#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      return 1;                                                         \
    }                                                                   \
  } while (0)

static int abi_checks() {
  CHECK(msq_abi_version() == MSQ_ABI_VERSION);
  // Byte sizes of every precision, including 0 elements and odd counts.
  for (int p : {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 16})
    for (size_t n : {size_t{0}, size_t{1}, size_t{31}, size_t{33}, size_t{1000003}}) {
      (void)msq_storage_bytes(p, n);
      (void)msq_scale_bytes(p, n);
      (void)msq_spectrum_metadata_bytes(p, n);
    }
  // Batch layout: regular, empty and invalid orders.
  std::vector<uint32_t> orders = {1, 31, 32, 33, 256, 257, 4096};
  std::vector<uint32_t> lds(orders.size());
  std::vector<uint64_t> offs(orders.size()), voffs(orders.size());
  uint64_t total = 0, vtotal = 0;
  CHECK(msq_batch_layout(static_cast<uint32_t>(orders.size()), orders.data(), lds.data(),
                         offs.data(), voffs.data(), &total, &vtotal) == MSQ_OK);
  CHECK(lds[0] == 32 && lds[3] == 64 && offs[1] == 32 * 32);
  CHECK(msq_batch_layout(0, orders.data(), lds.data(), offs.data(), voffs.data(), &total,
                         &vtotal) != MSQ_OK);
  const uint32_t bad[2] = {0, 70000};
  CHECK(msq_batch_layout(2, bad, lds.data(), offs.data(), voffs.data(), &total, &vtotal) !=
        MSQ_OK);
  CHECK(std::strlen(msq_last_error()) > 0);
  // Options defaults.
  msq_solve_options o;
  std::memset(&o, 0xff, sizeof(o));
  msq_solve_options_init(&o);
  CHECK(o.max_corrections == 12 && o.residual_tolerance == 1e-12);
  // Entry points with a null runtime: errors, never a crash.
  float f = 0.0f;
  double d = 0.0;
  uint32_t n = 4;
  msq_runtime* rt = nullptr;
  CHECK(msq_runtime_create(0, nullptr) != MSQ_OK);
  CHECK(msq_runtime_destroy(nullptr) == MSQ_OK);
  CHECK(msq_synchronize(rt) != MSQ_OK);
  CHECK(msq_runtime_set_rounding(rt, 1, 7) != MSQ_OK);
  CHECK(msq_quantize(rt, 1, &f, &f, nullptr, 1) != MSQ_OK);
  CHECK(msq_decode(rt, 1, &f, nullptr, &f, 1) != MSQ_OK);
  CHECK(msq_jacobi_batched(rt, 1, &n, &d, &f, &f, &f, nullptr, 0, nullptr) != MSQ_OK);
  CHECK(msq_jacobi_batched_noalloc(rt, 1, &n, &d, &f, &f, &f, nullptr, 0, nullptr) != MSQ_OK);
  CHECK(msq_solve_batched(rt, 1, &n, &d, &d, &o, 1, nullptr) != MSQ_OK);
  msq_solve_outputs out{};
  CHECK(msq_solve_batched_ex(rt, 1, &n, &d, &out, &o, 1, nullptr) != MSQ_OK);
  CHECK(msq_mlp_forward_bf16(rt, nullptr, &f, 1, &f) != MSQ_OK);
  CHECK(msq_matrix_features(rt, 1, &n, &d, &f) != MSQ_OK);
  CHECK(std::strlen(msq_last_error()) > 0);
  return 0;
}

static int mirror_checks() {
  using namespace msq_hip;
  // Names and parsers.
  for (const char* p : {"fp32", "bf16", "mxfp4_e2m1", "fp8_e4m3fn_log_s"})
    CHECK(std::string(precisionName(parsePrecision(p))) == p);
  CHECK(std::string(stoppingPolicyName(parseStoppingPolicy("fixed"))) == "fixed_budget");
  CHECK(std::string(kernelPathName(parseKernelPath("legacy"))) == "legacy");
  bool threw = false;
  try {
    parseRoundingMode("upward");
  } catch (const std::exception&) {
    threw = true;
  }
  CHECK(threw);
  // Paper test sets and the FP64 reference root.
  for (PaperTestSet s : {PaperTestSet::V, PaperTestSet::VI, PaperTestSet::VII}) {
    const PreparedMatrix p = preparePaperMatrix(s, 40);
    CHECK(p.values.size() == 1600 && p.condition >= 1.0);
    std::vector<double> root;
    (void)timeReferenceSolve(p, &root);
    CHECK(relativeResidual(p.values, root, 40) < 1e-10);
    CHECK(relativeSymmetryError(root, 40) == 0.0);
  }
  threw = false;
  try {
    preparePaperMatrix(PaperTestSet::II, 8);
  } catch (const std::exception&) {
    threw = true;
  }
  CHECK(threw);
  const std::vector<double> A = makeSpdMatrix(24, 100.0), X = makeSpdRoot(24, 100.0);
  CHECK(relativeError(symmetricReferenceRoot(A, 24), X) < 1e-12);
  // Timing summaries, one and several samples.
  const TimingSummary one = summarizeTimingSamples({0.5}, 0);
  CHECK(one.median_seconds == 0.5 && one.variance_seconds2 == 0.0);
  std::vector<double> ratios;
  const PairedRatioSummary pr = summarizePairedRatios({1, 2, 3}, {2, 2, 2}, 3, ratios);
  CHECK(pr.median == 1.0 && ratios.size() == 3);
  // CSV writers on a filled result.
  HipCaseResult r;
  r.name = "case";
  r.history = {{0, 1e-3, 1e-3}, {1, 1e-9, std::numeric_limits<double>::quiet_NaN()}};
  std::ostringstream s1, s2;
  writeSummaryHeader(s1);
  writeIterationsHeader(s2);
  writeCaseCsv(r, s1, s2);
  CHECK(s1.str().size() > 100 && s2.str().size() > 100);
  return 0;
}

int main() {
  if (abi_checks() != 0) return 1;
  if (mirror_checks() != 0) return 1;
  std::puts("ok");
  return 0;
}
