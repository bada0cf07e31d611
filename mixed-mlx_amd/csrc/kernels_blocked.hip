// Blocked Jacobi for orders above the one-workgroup kernel: the reference's
// parallel cyclic Jacobi (src/metal_solver.cpp:50-79, src/metal_runtime.cpp:
// 584-646) at block granularity, with the rotations applied as MFMA GEMMs.
//
// The padded order N (a multiple of 64) is cut into N / 32 blocks of 32
// indices.  The round-robin circle pairs the blocks (the same closed form as
// the scalar schedule, on block indices); each pair p owns a 64-index set P_p
// (its lower block's 32 indices, then its upper block's).  One block round:
//  1. gather: S_p = A[P_p, P_p] (64 x 64) into a binary64 staging batch;
//  2. the one-workgroup Jacobi kernel diagonalises every S_p (in LDS), its
//     eigenvectors U_p being the pair's 64 x 64 rotation;
//  3. apply: A[P_p, P_q] <- U_p^T A[P_p, P_q] U_q for every p <= q (both
//     triangles written), and V[:, P_p] <- V[:, P_p] U_p, as 64 x 64 x 64
//     products on v_mfma_f32_16x16x4_f32 with FP32 accumulation.
// A and V are float in device memory with pitch N.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace msq {
namespace {

constexpr int kB = kBlockedWidth;  // 32
constexpr int kP = 2 * kB;         // 64: order of a pair's subproblem
constexpr int kPitch = kP + 1;     // LDS pitch (odd: conflict-free columns)

__device__ __forceinline__ int block_participant(int t, int x, int nb) {
  if (x == 0) return 0;
  int r = x - 1 - t % (nb - 1);
  if (r < 0) r += nb - 1;
  return 1 + r;
}

// Global index of local index i (0..63) of pair p in round t.
__device__ __forceinline__ int pair_index(int t, int p, int nb, int i) {
  const int a = block_participant(t, p, nb);
  const int b = block_participant(t, nb - 1 - p, nb);
  const int lo = a < b ? a : b, hi = a < b ? b : a;
  return (i < kB ? lo : hi) * kB + (i & (kB - 1));
}

// C (64 x 64, LDS, column-major pitch kPitch) = op(X) Y with op(X) = X^T
// (TRANS) or X, for X, Y in LDS; C may alias X or Y (a barrier separates the
// last read from the first store).  Four waves, each a 32 x 32 quadrant of
// 2 x 2 tiles of 16 x 16; lane l holds X-operand element (row l % 16,
// k = 4 s + l / 16) and Y-operand element (k = 4 s + l / 16, column l % 16);
// accumulator r of lane l is C[4 (l / 16) + r][l % 16] of its tile.
template <bool TRANS>
__device__ __forceinline__ void lds_gemm64(const float* X, const float* Y, float* C) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m = lane & 15, kq = lane >> 4;
  const int r0 = 32 * (wave >> 1), c0 = 32 * (wave & 1);
  typedef float f4 __attribute__((ext_vector_type(4)));
  f4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
  for (int s = 0; s < kP / 4; ++s) {
    const int k = 4 * s + kq;
    float xa[2], yb[2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
      const int row = r0 + 16 * a + m;
      xa[a] = TRANS ? X[k + kPitch * row] : X[row + kPitch * k];
    }
#pragma unroll
    for (int b = 0; b < 2; ++b) yb[b] = Y[k + kPitch * (c0 + 16 * b + m)];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
        acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x4f32(xa[a], yb[b], acc[a][b], 0, 0, 0);
  }
  __syncthreads();
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        C[(r0 + 16 * a + 4 * kq + r) + kPitch * (c0 + 16 * b + m)] = acc[a][b][r];
}

__global__ __launch_bounds__(256) void blocked_gather_kernel(const float* __restrict__ A, int lda,
                                                             int nb, int t,
                                                             double* __restrict__ S) {
  const int p = blockIdx.x;
  double* Sp = S + static_cast<size_t>(p) * kP * kP;
  for (int e = threadIdx.x; e < kP * kP; e += blockDim.x) {
    const int i = e % kP, j = e / kP;
    Sp[e] = static_cast<double>(
        A[pair_index(t, p, nb, i) + static_cast<size_t>(lda) * pair_index(t, p, nb, j)]);
  }
}

// Block (p, q), p <= q, of the pair grid: A[P_p, P_q] <- U_p^T A[P_p, P_q] U_q,
// and its transpose into A[P_q, P_p].
__global__ __launch_bounds__(256) void blocked_apply_a_kernel(float* __restrict__ A, int lda,
                                                              int nb, int t,
                                                              const float* __restrict__ U) {
  __shared__ float Up[kP * kPitch], Uq[kP * kPitch], Bm[kP * kPitch];
  // blockIdx.x -> (p, q), p <= q, over npairs (npairs + 1) / 2 entries.
  const int np = nb / 2;
  int p = 0, rem = blockIdx.x;
  while (rem >= np - p) {
    rem -= np - p;
    ++p;
  }
  const int q = p + rem;
  const float* up = U + static_cast<size_t>(p) * kP * kP;
  const float* uq = U + static_cast<size_t>(q) * kP * kP;
  for (int e = threadIdx.x; e < kP * kP; e += blockDim.x) {
    const int i = e % kP, j = e / kP;
    Up[i + kPitch * j] = up[e];
    Uq[i + kPitch * j] = uq[e];
    Bm[i + kPitch * j] =
        A[pair_index(t, p, nb, i) + static_cast<size_t>(lda) * pair_index(t, q, nb, j)];
  }
  __syncthreads();
  lds_gemm64<true>(Up, Bm, Up);  // T = U_p^T B, over U_p
  __syncthreads();
  lds_gemm64<false>(Up, Uq, Bm);  // B' = T U_q, over B
  __syncthreads();
  for (int e = threadIdx.x; e < kP * kP; e += blockDim.x) {
    const int i = e % kP, j = e / kP;
    A[pair_index(t, p, nb, i) + static_cast<size_t>(lda) * pair_index(t, q, nb, j)] =
        Bm[i + kPitch * j];
    if (p != q)
      A[pair_index(t, q, nb, i) + static_cast<size_t>(lda) * pair_index(t, p, nb, j)] =
          Bm[j + kPitch * i];
  }
}

// V[rows of chunk c, P_p] <- V[rows, P_p] U_p (64 rows per workgroup).
__global__ __launch_bounds__(256) void blocked_apply_v_kernel(float* __restrict__ V, int ldv,
                                                              int nb, int t,
                                                              const float* __restrict__ U) {
  __shared__ float Up[kP * kPitch], Vc[kP * kPitch];
  const int c = blockIdx.x, p = blockIdx.y;
  const float* up = U + static_cast<size_t>(p) * kP * kP;
  for (int e = threadIdx.x; e < kP * kP; e += blockDim.x) {
    const int i = e % kP, j = e / kP;
    Up[i + kPitch * j] = up[e];
    Vc[i + kPitch * j] = V[(kP * c + i) + static_cast<size_t>(ldv) * pair_index(t, p, nb, j)];
  }
  __syncthreads();
  lds_gemm64<false>(Vc, Up, Vc);
  __syncthreads();
  for (int e = threadIdx.x; e < kP * kP; e += blockDim.x) {
    const int i = e % kP, j = e / kP;
    V[(kP * c + i) + static_cast<size_t>(ldv) * pair_index(t, p, nb, j)] = Vc[i + kPitch * j];
  }
}

// Working copies (pitch N, zero padding): A from the binary64 input (pitch
// ld_in), V = I.
__global__ void blocked_init_kernel(const double* __restrict__ a64, int n, int ld_in,
                                    float* __restrict__ A, float* __restrict__ V, int N) {
  const size_t total = static_cast<size_t>(N) * N;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(e % N), j = static_cast<int>(e / N);
    A[e] = (i < n && j < n) ? static_cast<float>(a64[i + static_cast<size_t>(j) * ld_in]) : 0.0f;
    V[e] = i == j ? 1.0f : 0.0f;
  }
}

// Outputs: V (pitch ld_out, zero outside the order), diagonal and roots
// (ld_out entries).
__global__ void blocked_finish_kernel(const float* __restrict__ A, const float* __restrict__ Vw,
                                      int n, int N, float* __restrict__ V, float* __restrict__ diag,
                                      float* __restrict__ roots, int ld_out) {
  const size_t total = static_cast<size_t>(ld_out) * ld_out;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(e % ld_out), j = static_cast<int>(e / ld_out);
    V[e] = (i < n && j < n) ? Vw[i + static_cast<size_t>(j) * N] : 0.0f;
    if (j == 0) {
      const float d = i < n ? A[i + static_cast<size_t>(i) * N] : 0.0f;
      diag[i] = d;
      roots[i] = i < n ? sqrtf(d) : 0.0f;
    }
  }
}

}  // namespace

uint32_t blocked_padded_order(uint32_t order) { return (order + 63u) & ~63u; }

hipError_t launch_blocked_init(const double* a64, uint32_t n, uint32_t ld_in, float* A, float* V,
                               uint32_t N, hipStream_t stream) {
  const size_t total = static_cast<size_t>(N) * N;
  size_t grid = (total + 255) / 256;
  if (grid > 65536) grid = 65536;
  blocked_init_kernel<<<static_cast<int>(grid), 256, 0, stream>>>(a64, n, ld_in, A, V, N);
  return hipGetLastError();
}

hipError_t launch_blocked_round(float* A, float* V, uint32_t N, uint32_t round, double* staging,
                                float* U, const JacobiJob* jobs, float* pair_diag,
                                float* pair_roots, JacobiStatus* pair_status,
                                hipStream_t stream) {
  const int nb = static_cast<int>(N) / kB;
  const int np = nb / 2;
  if (N % kP != 0 || nb < 2) return hipErrorInvalidValue;
  blocked_gather_kernel<<<np, 256, 0, stream>>>(A, N, nb, round, staging);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_batched_jacobi(jobs, np, kP, staging, U, pair_diag, pair_roots, pair_status, stream);
  if (e != hipSuccess) return e;
  blocked_apply_a_kernel<<<np * (np + 1) / 2, 256, 0, stream>>>(A, N, nb, round, U);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  blocked_apply_v_kernel<<<dim3(N / kP, np), 256, 0, stream>>>(V, N, nb, round, U);
  return hipGetLastError();
}

hipError_t launch_blocked_finish(const float* A, const float* Vw, uint32_t n, uint32_t N, float* V,
                                 float* diag, float* roots, uint32_t ld_out, hipStream_t stream) {
  const size_t total = static_cast<size_t>(ld_out) * ld_out;
  size_t grid = (total + 255) / 256;
  if (grid > 65536) grid = 65536;
  blocked_finish_kernel<<<static_cast<int>(grid), 256, 0, stream>>>(A, Vw, n, N, V, diag, roots,
                                                                    ld_out);
  return hipGetLastError();
}

}  // namespace msq
