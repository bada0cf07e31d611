// CPU baseline (test infrastructure only): a C++ restatement of the
// reference's CPU plain path, accelerate_mixed (1_baseline/spd.cpp:348-430),
// on the host LAPACK/BLAS (scipy's OpenBLAS build, symbols prefixed scipy_),
// run over a batch with one matrix per host thread and single-threaded BLAS.
// Only bench.py's cpu_baseline leg and tests/ load it; it is the baseline the
// GPU path is timed beside, never part of the product.
//
// Per matrix (column-major n x n, binary64 in, binary64 root out):
//   Q, lambda = ssyevd(float(A))   (fails on a nonpositive eigenvalue)
//   s = sqrt(lambda);  X = float(Q diag(s)) Q^T (sgemm N,T) in binary64, symmetrised
//   R = A - X X (dgemm N,N);  res = ||R||_F / ||A||_F accumulated with hypot
//   while res > tolerance and k < max_corrections:
//     W = Q^T float(R); C = W Q; C_ij = 0.5 (C_ij + C_ji) / (s_i + s_j)
//     W = Q C; E = W Q^T; X += E (binary64); symmetrise; R, res again
// (symmetrize / validate_positive / relative_residual, 1_baseline/spd.cpp:182-200,
// :317-333).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <exception>
#include <thread>
#include <vector>

extern "C" {
void scipy_ssyevd_(const char* jobz, const char* uplo, const int* n, float* a, const int* lda,
                   float* w, float* work, const int* lwork, int* iwork, const int* liwork,
                   int* info, size_t, size_t);
void scipy_sgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k,
                  const float* alpha, const float* a, const int* lda, const float* b,
                  const int* ldb, const float* beta, float* c, const int* ldc, size_t, size_t);
void scipy_dgemm_(const char* ta, const char* tb, const int* m, const int* n, const int* k,
                  const double* alpha, const double* a, const int* lda, const double* b,
                  const int* ldb, const double* beta, double* c, const int* ldc, size_t, size_t);
void scipy_openblas_set_num_threads(int);
int scipy_openblas_get_num_threads(void);
char* scipy_openblas_get_config(void);
}

namespace {

void sgemm(char ta, char tb, int n, const float* a, const float* b, float* c) {
  const float one = 1.0f, zero = 0.0f;
  scipy_sgemm_(&ta, &tb, &n, &n, &n, &one, a, &n, b, &n, &zero, c, &n, 1, 1);
}

void dgemm(char ta, char tb, int n, const double* a, const double* b, double* c) {
  const double one = 1.0, zero = 0.0;
  scipy_dgemm_(&ta, &tb, &n, &n, &n, &one, a, &n, b, &n, &zero, c, &n, 1, 1);
}

bool eigen_float(std::vector<float>& a, std::vector<float>& w, int n) {
  const char job = 'V', uplo = 'U';
  int info = 0, lwork = -1, liwork = -1, iq = 0;
  float wq = 0.0f;
  scipy_ssyevd_(&job, &uplo, &n, a.data(), &n, w.data(), &wq, &lwork, &iq, &liwork, &info, 1, 1);
  if (info != 0) return false;
  lwork = static_cast<int>(std::ceil(wq));
  liwork = iq;
  std::vector<float> work(static_cast<size_t>(std::max(lwork, 1)));
  std::vector<int> iwork(static_cast<size_t>(std::max(liwork, 1)));
  scipy_ssyevd_(&job, &uplo, &n, a.data(), &n, w.data(), work.data(), &lwork, iwork.data(),
                &liwork, &info, 1, 1);
  return info == 0;
}

void symmetrize(double* x, int n) {
  for (int j = 0; j < n; ++j)
    for (int i = j + 1; i < n; ++i) {
      const double v = 0.5 * (x[i + static_cast<size_t>(j) * n] + x[j + static_cast<size_t>(i) * n]);
      x[i + static_cast<size_t>(j) * n] = v;
      x[j + static_cast<size_t>(i) * n] = v;
    }
}

// R = A - X X and ||R||_F / ||A||_F with the norms accumulated by hypot.
double residual(const double* a, const double* x, double* r, double* product, size_t count,
                int n) {
  dgemm('N', 'N', n, x, x, product);
  double an = 0.0, rn = 0.0;
  for (size_t k = 0; k < count; ++k) {
    r[k] = a[k] - product[k];
    an = std::hypot(an, a[k]);
    rn = std::hypot(rn, r[k]);
  }
  return rn / an;
}

// One matrix; returns -1 on a failed or indefinite eigensolve, else the
// number of corrections run.
int solve_one(const double* a, double* x, int n, uint32_t max_corrections, double tolerance,
              double* final_residual) {
  const size_t count = static_cast<size_t>(n) * n;
  *final_residual = std::nan("");
  std::vector<float> q(a, a + count), lambda(n);
  if (!eigen_float(q, lambda, n)) return -1;
  for (float& v : lambda) {
    if (!(v > 0.0f) || !std::isfinite(v)) return -1;
    v = std::sqrt(v);
  }
  std::vector<float> scaled = q, root32(count);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < n; ++i) scaled[i + static_cast<size_t>(j) * n] *= lambda[j];
  sgemm('N', 'T', n, scaled.data(), q.data(), root32.data());
  for (size_t k = 0; k < count; ++k) x[k] = static_cast<double>(root32[k]);
  symmetrize(x, n);
  std::vector<double> r(count), product(count);
  double res = residual(a, x, r.data(), product.data(), count, n);
  std::vector<float> r32(count), work(count), coords(count), corr(count);
  uint32_t k = 0;
  for (; res > tolerance && k < max_corrections; ++k) {
    for (size_t e = 0; e < count; ++e) r32[e] = static_cast<float>(r[e]);
    sgemm('T', 'N', n, q.data(), r32.data(), work.data());
    sgemm('N', 'N', n, work.data(), q.data(), coords.data());
    for (int j = 0; j < n; ++j)
      for (int i = 0; i <= j; ++i) {
        const size_t up = i + static_cast<size_t>(j) * n, lo = j + static_cast<size_t>(i) * n;
        float v = 0.5f * (coords[up] + coords[lo]);
        v /= lambda[i] + lambda[j];
        coords[up] = v;
        coords[lo] = v;
      }
    sgemm('N', 'N', n, q.data(), coords.data(), work.data());
    sgemm('N', 'T', n, work.data(), q.data(), corr.data());
    for (size_t e = 0; e < count; ++e) x[e] += static_cast<double>(corr[e]);
    symmetrize(x, n);
    res = residual(a, x, r.data(), product.data(), count, n);
  }
  *final_residual = res;
  return static_cast<int>(k);
}

}  // namespace

extern "C" {

// Solves matrices [0, count) (dense column-major, matrix i at offsets[i] in
// `matrices` and `roots`) with `threads` host threads, each taking the next
// index from a shared counter and running BLAS single-threaded.  corrections[i]
// is -1 where the eigensolve failed.  Returns the wall seconds of the batch,
// or a negative value on a bad argument.
double cpu_accelerate_mixed_batch(uint32_t count, const uint32_t* orders,
                                  const uint64_t* offsets, const double* matrices, double* roots,
                                  uint32_t max_corrections, double tolerance, uint32_t threads,
                                  int32_t* corrections, double* residuals) {
  if (!orders || !offsets || !matrices || !roots || !corrections || !residuals || threads == 0)
    return -1.0;
  const int saved = scipy_openblas_get_num_threads();
  scipy_openblas_set_num_threads(1);
  std::atomic<uint32_t> next{0};
  std::atomic<bool> failed{false};
  const auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    try {
      for (uint32_t i = next.fetch_add(1); i < count; i = next.fetch_add(1))
        corrections[i] = solve_one(matrices + offsets[i], roots + offsets[i],
                                   static_cast<int>(orders[i]), max_corrections, tolerance,
                                   residuals + i);
    } catch (const std::exception&) {
      failed = true;
    }
  };
  std::vector<std::thread> pool;
  for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();
  const double seconds =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  scipy_openblas_set_num_threads(saved);
  return failed ? -2.0 : seconds;
}

// The BLAS build string (library and version) for the report.
const char* cpu_blas_config(void) { return scipy_openblas_get_config(); }

}  // extern "C"
