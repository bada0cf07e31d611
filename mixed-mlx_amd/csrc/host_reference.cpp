// Host side of the solver layer (see host_reference.hpp for the
// correspondence with the reference).
#include "host_reference.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <stdexcept>
#include <utility>

// Binary64 LAPACK/BLAS of the host OpenBLAS (scipy's build exports them with
// a scipy_ prefix; LP64 integers).  The trailing size_t arguments are the
// Fortran character lengths, passed for the routines compiled from Fortran.
extern "C" {
void scipy_dsyevd_(const char* jobz, const char* uplo, const int* n, double* a, const int* lda,
                   double* w, double* work, const int* lwork, int* iwork, const int* liwork,
                   int* info, size_t jobz_len, size_t uplo_len);
void scipy_dgemm_(const char* transa, const char* transb, const int* m, const int* n,
                  const int* k, const double* alpha, const double* a, const int* lda,
                  const double* b, const int* ldb, const double* beta, double* c, const int* ldc,
                  size_t transa_len, size_t transb_len);
void scipy_dsyrk_(const char* uplo, const char* trans, const int* n, const int* k,
                  const double* alpha, const double* a, const int* lda, const double* beta,
                  double* c, const int* ldc, size_t uplo_len, size_t trans_len);
}

namespace msq_hip {

void fail(const std::string& message) { throw std::runtime_error(message); }

namespace {

int lapack_order(size_t order) {
  if (order == 0 || order > static_cast<size_t>(std::numeric_limits<int>::max() / 2))
    fail("matrix order out of LAPACK range");
  return static_cast<int>(order);
}

struct PaperName {
  PaperTestSet set;
  const char* name;
};
const PaperName kPaperNames[] = {
    {PaperTestSet::None, "controlled"}, {PaperTestSet::I, "I"},   {PaperTestSet::II, "II"},
    {PaperTestSet::III, "III"},         {PaperTestSet::IV, "IV"}, {PaperTestSet::V, "V"},
    {PaperTestSet::VI, "VI"},           {PaperTestSet::VII, "VII"},
};

// The controlled basis: sin/cos columns orthonormalised in order by classical
// Gram-Schmidt (src/numerics.cpp:141-195).
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

// Q diag(cond^(power j / (n - 1))) Q^T: the columns of Q scaled, then dgemm
// (N, T) with Q, as src/numerics.cpp:174-195 forms it.
std::vector<double> controlled_product(size_t n, double condition, double power) {
  if (!std::isfinite(condition) || condition < 1.0)
    fail("the SPD condition target must be finite and at least one");
  const int dim = lapack_order(n);
  const std::vector<double> q = controlled_basis(n);
  std::vector<double> scaled = q;
  const double log_cond = std::log(condition);
  for (size_t k = 0; k < n; ++k) {
    const double fraction = static_cast<double>(k) / static_cast<double>(std::max<size_t>(1, n - 1));
    const double s = std::exp(power * log_cond * fraction);
    for (size_t i = 0; i < n; ++i) scaled[i + k * n] *= s;
  }
  std::vector<double> c(n * n, 0.0);
  const char no = 'N', tr = 'T';
  const double one = 1.0, zero = 0.0;
  scipy_dgemm_(&no, &tr, &dim, &dim, &dim, &one, scaled.data(), &dim, q.data(), &dim, &zero,
               c.data(), &dim, 1, 1);
  return c;
}

double seconds_since(std::chrono::steady_clock::time_point start) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
}

}  // namespace

// ------------------------------------------------------------ test sets
const char* paperTestSetName(PaperTestSet test_set) {
  for (const auto& p : kPaperNames)
    if (p.set == test_set) return p.name;
  fail("unknown paper test set");
}

PaperTestSet paperTestSetForCliName(const std::string& name) {
  for (const auto& p : kPaperNames)
    if (p.set != PaperTestSet::None && name == p.name) return p.set;
  fail("unsupported paper test set: " + name);
}

bool paperTestSetIsSymmetric(PaperTestSet test_set) {
  return test_set == PaperTestSet::V || test_set == PaperTestSet::VI ||
         test_set == PaperTestSet::VII;
}

size_t paperTestSetOrder(PaperTestSet test_set) {
  switch (test_set) {
    case PaperTestSet::I:
    case PaperTestSet::II:
      return 2048;
    case PaperTestSet::III:
    case PaperTestSet::IV:
    case PaperTestSet::V:
    case PaperTestSet::VI:
      return 4096;
    case PaperTestSet::VII:
      return 1024;
    case PaperTestSet::None:
      break;
  }
  fail("the controlled matrix family has no paper order");
}

// ------------------------------------------------------------- generator
DeterministicGenerator::DeterministicGenerator(uint64_t seed)
    : state_(seed ? seed : UINT64_C(0x9e3779b97f4a7c15)) {}

double DeterministicGenerator::uniform() {
  uint64_t x = state_;
  x ^= x >> 12;
  x ^= x << 25;
  x ^= x >> 27;
  state_ = x;
  return static_cast<double>((x * UINT64_C(2685821657736338717)) >> 11) * 0x1.0p-53;
}

double DeterministicGenerator::normal() {
  if (has_spare_) {
    has_spare_ = false;
    return spare_;
  }
  const double u1 = std::max(uniform(), std::numeric_limits<double>::min());
  const double u2 = uniform();
  const double radius = std::sqrt(-2.0 * std::log(u1));
  const double angle = 6.283185307179586476925286766559 * u2;
  spare_ = radius * std::sin(angle);
  has_spare_ = true;
  return radius * std::cos(angle);
}

// ------------------------------------------------------ binary64 numerics
void symmetricEigenDouble(std::vector<double>& matrix, std::vector<double>& values) {
  const int n = lapack_order(values.size());
  if (matrix.size() != values.size() * values.size()) fail("dsyevd: matrix is not order x order");
  const char job = 'V', uplo = 'U';
  int info = 0, lwork = -1, liwork = -1, iwork_query = 0;
  double work_query = 0.0;
  scipy_dsyevd_(&job, &uplo, &n, matrix.data(), &n, values.data(), &work_query, &lwork,
                &iwork_query, &liwork, &info, 1, 1);
  if (info != 0) fail("dsyevd workspace query failed");
  lwork = static_cast<int>(std::ceil(work_query));
  liwork = iwork_query;
  std::vector<double> work(static_cast<size_t>(std::max(lwork, 1)));
  std::vector<int> iwork(static_cast<size_t>(std::max(liwork, 1)));
  scipy_dsyevd_(&job, &uplo, &n, matrix.data(), &n, values.data(), work.data(), &lwork,
                iwork.data(), &liwork, &info, 1, 1);
  if (info != 0) fail("dsyevd failed (info " + std::to_string(info) + ")");
}

std::vector<double> rootFromEigen(const std::vector<double>& vectors,
                                  const std::vector<double>& values, size_t order) {
  const int n = lapack_order(order);
  std::vector<double> weighted = vectors;
  for (size_t j = 0; j < order; ++j) {
    const double s = std::sqrt(values[j]);
    for (size_t i = 0; i < order; ++i) weighted[i + j * order] *= s;
  }
  std::vector<double> root(order * order, 0.0);
  const char no = 'N', tr = 'T';
  const double one = 1.0, zero = 0.0;
  scipy_dgemm_(&no, &tr, &n, &n, &n, &one, weighted.data(), &n, vectors.data(), &n, &zero,
               root.data(), &n, 1, 1);
  return root;
}

std::vector<double> symmetricSquare(const std::vector<double>& matrix, size_t order) {
  const int n = lapack_order(order);
  std::vector<double> product(order * order, 0.0);
  const char uplo = 'U', no = 'N';
  const double one = 1.0, zero = 0.0;
  scipy_dsyrk_(&uplo, &no, &n, &n, &one, matrix.data(), &n, &zero, product.data(), &n, 1, 1);
  for (size_t j = 0; j < order; ++j)
    for (size_t i = j + 1; i < order; ++i) product[i + j * order] = product[j + i * order];
  return product;
}

void symmetrizeInPlace(std::vector<double>& matrix, size_t order) {
  for (size_t j = 0; j < order; ++j)
    for (size_t i = j + 1; i < order; ++i) {
      const double v = 0.5 * (matrix[i + j * order] + matrix[j + i * order]);
      matrix[i + j * order] = v;
      matrix[j + i * order] = v;
    }
}

std::vector<double> symmetricReferenceRoot(const std::vector<double>& matrix, size_t order) {
  std::vector<double> vectors = matrix;
  std::vector<double> values(order);
  symmetricEigenDouble(vectors, values);
  for (double v : values)
    if (!(v > 0.0) || !std::isfinite(v)) fail("the FP64 reference has a nonpositive eigenvalue");
  std::vector<double> root = rootFromEigen(vectors, values, order);
  symmetrizeInPlace(root, order);
  return root;
}

double relativeResidual(const std::vector<double>& matrix, const std::vector<double>& root,
                        size_t order) {
  const std::vector<double> product = symmetricSquare(root, order);
  double a2 = 0.0, r2 = 0.0;
  for (size_t k = 0; k < matrix.size(); ++k) {
    const double r = matrix[k] - product[k];
    a2 += matrix[k] * matrix[k];
    r2 += r * r;
  }
  return std::sqrt(r2 / a2);
}

double relativeError(const std::vector<double>& root, const std::vector<double>& reference) {
  if (root.size() != reference.size()) fail("relativeError: sizes differ");
  double d2 = 0.0, r2 = 0.0;
  for (size_t k = 0; k < root.size(); ++k) {
    const double d = root[k] - reference[k];
    d2 += d * d;
    r2 += reference[k] * reference[k];
  }
  return std::sqrt(d2 / r2);
}

double frobeniusDifference(const std::vector<double>& left, const std::vector<double>& right) {
  if (left.size() != right.size()) fail("frobeniusDifference: sizes differ");
  double d2 = 0.0;
  for (size_t k = 0; k < left.size(); ++k) {
    const double d = left[k] - right[k];
    d2 += d * d;
  }
  return std::sqrt(d2);
}

double relativeSymmetryError(const std::vector<double>& matrix, size_t order) {
  double m2 = 0.0, d2 = 0.0;
  for (size_t j = 0; j < order; ++j)
    for (size_t i = 0; i < order; ++i) {
      const double v = matrix[i + j * order];
      const double d = v - matrix[j + i * order];
      m2 += v * v;
      d2 += d * d;
    }
  return std::sqrt(d2 / m2);
}

double matrixInfinityNorm(const std::vector<double>& matrix, size_t order) {
  double best = 0.0;
  for (size_t i = 0; i < order; ++i) {
    double row = 0.0;
    for (size_t j = 0; j < order; ++j) row += std::abs(matrix[i + j * order]);
    best = std::max(best, row);
  }
  if (!(best > 0.0) || !std::isfinite(best)) fail("the prepared matrix norm is invalid");
  return best;
}

bool allFinite(const std::vector<double>& values) {
  return std::all_of(values.begin(), values.end(), [](double v) { return std::isfinite(v); });
}

std::vector<double> makeSpdMatrix(size_t order, double condition) {
  return controlled_product(order, condition, 1.0);
}

std::vector<double> makeSpdRoot(size_t order, double condition) {
  return controlled_product(order, condition, 0.5);
}

// -------------------------------------------------------- prepared inputs
PreparedMatrix prepareControlledMatrix(size_t order, double condition) {
  lapack_order(order);
  PreparedMatrix p;
  p.values = makeSpdMatrix(order, condition);
  p.order = order;
  p.condition = condition;
  return p;
}

PreparedMatrix preparePaperMatrix(PaperTestSet test_set, size_t order) {
  if (test_set == PaperTestSet::None) fail("a paper test set is required");
  if (!paperTestSetIsSymmetric(test_set))
    fail("the HIP backend solves symmetric positive definite paper sets only (V, VI, VII)");
  lapack_order(order);
  PreparedMatrix p;
  p.order = order;
  p.paper_test_set = paperTestSetName(test_set);
  p.matrix_type = "symmetric";
  p.values.assign(order * order, 0.0);
  std::vector<double> eigenvalues(order);
  if (test_set == PaperTestSet::VII) {
    p.matrix_family = "hilbert_shifted";
    p.shift_rule = "fixed_hilbert_tau";
    for (size_t j = 0; j < order; ++j) {
      for (size_t i = 0; i < order; ++i)
        p.values[i + j * order] = 1.0 / static_cast<double>(i + j + 1);
      p.values[j + j * order] += 1.0e-6;
    }
    p.diagonal_shift = 1.0e-6;
    std::vector<double> discovery = p.values;
    symmetricEigenDouble(discovery, eigenvalues);
    if (!(eigenvalues.front() > 0.0) || !std::isfinite(eigenvalues.back()))
      fail("paper set VII is not positive definite");
    p.condition = eigenvalues.back() / eigenvalues.front();
    return p;
  }
  const bool uniform = test_set == PaperTestSet::V;
  p.matrix_family = uniform ? "uniform_random" : "normal_random";
  p.shift_rule = "set_min_real_part_to_max_1_or_0.01_inf_norm";
  p.rng_seed = UINT64_C(0x243f6a8885a308d3) ^
               (UINT64_C(0x9e3779b97f4a7c15) * static_cast<uint64_t>(static_cast<int>(test_set)));
  DeterministicGenerator g(p.rng_seed);
  // Upper triangle drawn column by column from the diagonal down, mirrored.
  for (size_t j = 0; j < order; ++j)
    for (size_t i = j; i < order; ++i) {
      const double v = uniform ? g.uniform() : g.normal();
      p.values[i + j * order] = v;
      p.values[j + i * order] = v;
    }
  if (!allFinite(p.values)) fail("the generated paper matrix is nonfinite");
  const double separation = std::max(1.0, 0.01 * matrixInfinityNorm(p.values, order));
  std::vector<double> discovery = p.values;
  symmetricEigenDouble(discovery, eigenvalues);
  const double lo = eigenvalues.front(), hi = eigenvalues.back();
  if (!std::isfinite(lo) || !std::isfinite(hi)) fail("the paper matrix shift discovery is nonfinite");
  const double shift = std::max(0.0, separation - lo);
  p.diagonal_shift = shift;
  for (size_t k = 0; k < order; ++k) p.values[k + k * order] += shift;
  p.condition = (hi + shift) / (lo + shift);
  if (!allFinite(p.values)) fail("the shifted paper matrix is nonfinite");
  return p;
}

double timeReferenceSolve(const PreparedMatrix& prepared, std::vector<double>* root_output) {
  if (prepared.matrix_type != "symmetric") fail("the FP64 reference root needs a symmetric matrix");
  const auto start = std::chrono::steady_clock::now();
  std::vector<double> root = symmetricReferenceRoot(prepared.values, prepared.order);
  const double seconds = seconds_since(start);
  if (!allFinite(root)) fail("the timed FP64 reference root is nonfinite");
  if (root_output) *root_output = std::move(root);
  return seconds;
}

// ------------------------------------------------------- timing summaries
TimingSummary summarizeTimingSamples(std::vector<double> samples, size_t warmups) {
  if (samples.empty()) fail("timing samples must not be empty");
  for (double s : samples)
    if (!(s >= 0.0) || !std::isfinite(s)) fail("a timing sample is invalid");
  std::sort(samples.begin(), samples.end());
  double mean = 0.0;
  for (double s : samples) mean += s;
  mean /= static_cast<double>(samples.size());
  double ss = 0.0;
  for (double s : samples) ss += (s - mean) * (s - mean);
  TimingSummary t;
  t.warmups = warmups;
  t.evaluations = samples.size();
  t.minimum_seconds = samples.front();
  t.maximum_seconds = samples.back();
  t.median_seconds = samples[samples.size() / 2];
  t.mean_seconds = mean;
  t.variance_seconds2 = samples.size() > 1 ? ss / static_cast<double>(samples.size() - 1) : ss;
  return t;
}

PairedRatioSummary summarizePairedRatios(const std::vector<double>& mixed_samples,
                                         const std::vector<double>& reference_samples,
                                         size_t expected_count, std::vector<double>& ratios) {
  if (expected_count == 0 || mixed_samples.size() != expected_count ||
      reference_samples.size() != expected_count)
    fail("paired timing sample count does not match the evaluation count");
  ratios.assign(expected_count, 0.0);
  PairedRatioSummary s;
  s.minimum = std::numeric_limits<double>::infinity();
  s.maximum = -std::numeric_limits<double>::infinity();
  double m2 = 0.0;
  for (size_t k = 0; k < expected_count; ++k) {
    const double mixed = mixed_samples[k], ref = reference_samples[k];
    if (!(mixed > 0.0) || !std::isfinite(mixed) || !(ref > 0.0) || !std::isfinite(ref))
      fail("a paired timing sample is nonpositive or nonfinite");
    const double r = mixed / ref;
    if (!(r > 0.0) || !std::isfinite(r)) fail("a paired timing ratio is nonpositive or nonfinite");
    ratios[k] = r;
    s.minimum = std::min(s.minimum, r);
    s.maximum = std::max(s.maximum, r);
    const double before = r - s.mean;
    s.mean += before / static_cast<double>(k + 1);
    m2 += before * (r - s.mean);
  }
  std::vector<double> sorted = ratios;
  std::sort(sorted.begin(), sorted.end());
  const size_t mid = sorted.size() / 2;
  s.median = sorted.size() % 2 ? sorted[mid] : 0.5 * (sorted[mid - 1] + sorted[mid]);
  s.variance = ratios.size() > 1 ? m2 / static_cast<double>(ratios.size() - 1) : 0.0;
  return s;
}

}  // namespace msq_hip
