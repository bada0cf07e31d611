// Prints a paper test set of the C++ mirror (host_reference.cpp) for the CPU
// test that compares it with the Python generator: order, condition, shift,
// seed, labels, then the column-major values.  With "summary", prints the
// timing and paired-ratio summaries of the samples given instead.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "host_reference.hpp"

int main(int argc, char** argv) {
  if (argc >= 2 && std::string(argv[1]) == "summary") {
    // summary m1,m2,... r1,r2,...
    auto parse = [](const char* s) {
      std::vector<double> v;
      for (const char* p = s; *p;) {
        char* end = nullptr;
        v.push_back(std::strtod(p, &end));
        if (end == p) std::exit(2);  // not a number
        p = end;
        if (*p == ',') ++p;
      }
      return v;
    };
    const std::vector<double> mixed = parse(argv[2]), ref = parse(argv[3]);
    const auto t = msq_hip::summarizeTimingSamples(mixed, 3);
    std::printf("%zu %zu %.17g %.17g %.17g %.17g %.17g\n", t.warmups, t.evaluations,
                t.minimum_seconds, t.maximum_seconds, t.median_seconds, t.mean_seconds,
                t.variance_seconds2);
    std::vector<double> ratios;
    const auto p = msq_hip::summarizePairedRatios(mixed, ref, mixed.size(), ratios);
    std::printf("%.17g %.17g %.17g %.17g %.17g\n", p.mean, p.variance, p.minimum, p.maximum,
                p.median);
    for (double r : ratios) std::printf("%.17g\n", r);
    return 0;
  }
  if (argc != 3) return 2;
  try {
    const auto set = msq_hip::paperTestSetForCliName(argv[1]);
    const auto p = msq_hip::preparePaperMatrix(set, std::strtoul(argv[2], nullptr, 10));
    std::printf("%zu %.17g %.17g %llu\n%s %s %s %s\n", p.order, p.condition, p.diagonal_shift,
                static_cast<unsigned long long>(p.rng_seed), p.paper_test_set.c_str(),
                p.matrix_family.c_str(), p.matrix_type.c_str(), p.shift_rule.c_str());
    for (double v : p.values) std::printf("%.17g\n", v);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
