// Prints makeSpdMatrix / makeSpdRoot of the C++ mirror (host_reference.cpp)
// for the CPU test that compares them with the oracle's restatement.
#include <cstdio>
#include <cstdlib>

#include "host_reference.hpp"

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  const size_t n = std::strtoul(argv[1], nullptr, 10);
  const double cond = std::strtod(argv[2], nullptr);
  const auto A = msq_hip::makeSpdMatrix(n, cond);
  const auto X = msq_hip::makeSpdRoot(n, cond);
  for (double v : A) std::printf("%.17g\n", v);
  for (double v : X) std::printf("%.17g\n", v);
  return 0;
}
