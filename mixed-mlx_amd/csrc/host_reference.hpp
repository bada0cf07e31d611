// Host side of the solver layer: the FP64 reference root, the prepared inputs
// (controlled and paper test sets), the host norms, and the timing summaries
// a caller of the solver entry points uses beside the device solve.
//
//   symmetricEigenDouble / rootFromEigen / symmetricSquare / relativeResidual /
//   relativeError / frobeniusDifference / relativeSymmetryError
//                               (src/numerics.cpp:48-264)
//   symmetricReferenceRoot      (src/accelerate_backend.cpp:594-607)
//   DeterministicGenerator, matrixInfinityNorm, prepareControlledMatrix,
//   preparePaperMatrix          (src/accelerate_backend.cpp:624-801)
//   timeReferenceSolve          (src/accelerate_backend.cpp:1093-1108)
//   PreparedMatrix, TimingSummary, PairedRatioSummary (src/experiment.hpp:17-52)
//   summarizeTimingSamples / summarizePairedRatios (src/benchmark.cpp:17-97)
//   paper test set names and orders (src/configuration.cpp:58-63, :237-271)
//
// Binary64 LAPACK/BLAS (dsyevd, dgemm, dsyrk) come from the host's OpenBLAS,
// as the reference's come from Accelerate.  All matrices are column-major
// order x order.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace msq_hip {

// Errors throw std::runtime_error (the reference's fail(), src/core.cpp:7-9).
[[noreturn]] void fail(const std::string& message);

enum class PaperTestSet { None, I, II, III, IV, V, VI, VII };
const char* paperTestSetName(PaperTestSet test_set);
PaperTestSet paperTestSetForCliName(const std::string& name);
bool paperTestSetIsSymmetric(PaperTestSet test_set);
size_t paperTestSetOrder(PaperTestSet test_set);

struct PreparedMatrix {
  std::vector<double> values;
  size_t order = 0;
  double condition = 0.0;
  std::string paper_test_set = "controlled";
  std::string matrix_family = "conditioned_spd";
  std::string matrix_type = "symmetric";
  uint64_t rng_seed = 0;
  double diagonal_shift = 0.0;
  std::string shift_rule = "not_applicable";
};

// xorshift64* uniforms on [0, 1) with 53 bits, Box-Muller normals (the second
// of each pair kept for the next call); seed 0 is replaced by the golden-ratio
// constant.
class DeterministicGenerator {
 public:
  explicit DeterministicGenerator(uint64_t seed);
  double uniform();
  double normal();

 private:
  uint64_t state_;
  double spare_ = 0.0;
  bool has_spare_ = false;
};

// LAPACK dsyevd ('V', 'U'): eigenvalues ascending, eigenvectors overwrite
// `matrix`.
void symmetricEigenDouble(std::vector<double>& matrix, std::vector<double>& values);
// V diag(sqrt(lambda)) V^T (dgemm N, T).
std::vector<double> rootFromEigen(const std::vector<double>& vectors,
                                  const std::vector<double>& values, size_t order);
// X X^T from dsyrk's upper triangle, mirrored.
std::vector<double> symmetricSquare(const std::vector<double>& matrix, size_t order);
void symmetrizeInPlace(std::vector<double>& matrix, size_t order);
// The binary64 principal root; fails on a nonpositive eigenvalue.
std::vector<double> symmetricReferenceRoot(const std::vector<double>& matrix, size_t order);
// ||A - X^2||_F / ||A||_F.
double relativeResidual(const std::vector<double>& matrix, const std::vector<double>& root,
                        size_t order);
double relativeError(const std::vector<double>& root, const std::vector<double>& reference);
double frobeniusDifference(const std::vector<double>& left, const std::vector<double>& right);
double relativeSymmetryError(const std::vector<double>& matrix, size_t order);
double matrixInfinityNorm(const std::vector<double>& matrix, size_t order);
bool allFinite(const std::vector<double>& values);

// The reference's controlled input: a deterministic sin/cos basis Q
// orthonormalised by Gram-Schmidt, A = Q diag(cond^(j / (n - 1))) Q^T.
std::vector<double> makeSpdMatrix(size_t order, double condition);
// Its square root from the same construction (binary64).
std::vector<double> makeSpdRoot(size_t order, double condition);

PreparedMatrix prepareControlledMatrix(size_t order, double condition);
// The symmetric paper test sets: V (uniform), VI (normal) with the diagonal
// shift that sets the smallest eigenvalue to max(1, 0.01 ||A||_inf), and VII
// (Hilbert + 1e-6 I).  The nonsymmetric sets I-IV need a real Schur root,
// which the SPD solver does not take: they fail.
PreparedMatrix preparePaperMatrix(PaperTestSet test_set, size_t order);

// Seconds of one FP64 reference root of the prepared matrix (symmetric only);
// the root is returned through root_output when given.
double timeReferenceSolve(const PreparedMatrix& prepared,
                          std::vector<double>* root_output = nullptr);

struct TimingSummary {
  size_t warmups = 0;
  size_t evaluations = 1;
  double minimum_seconds = 0.0;
  double maximum_seconds = 0.0;
  double median_seconds = 0.0;
  double mean_seconds = 0.0;
  double variance_seconds2 = 0.0;
};

struct PairedRatioSummary {
  double mean = 0.0;
  double variance = 0.0;
  double minimum = 0.0;
  double maximum = 0.0;
  double median = 0.0;
};

// Sorted samples: min, max, median = sorted[size / 2], mean, variance with
// n - 1 (0 for one sample).  Fails on an empty set or a negative or
// nonfinite sample.
TimingSummary summarizeTimingSamples(std::vector<double> samples, size_t warmups);
// ratio_i = mixed_i / reference_i in evaluation order (returned in `ratios`);
// running (Welford) mean and variance with n - 1, min, max, and the median of
// the sorted ratios (mean of the middle two for an even count).
PairedRatioSummary summarizePairedRatios(const std::vector<double>& mixed_samples,
                                         const std::vector<double>& reference_samples,
                                         size_t expected_count, std::vector<double>& ratios);

}  // namespace msq_hip
