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
// A and V are float working copies in device memory with pitch N.  All
// matrices of one padded order advance together: every launch covers the
// active members of the class (a device list of member indices, rebuilt by
// the host after each sweep), so a batch of thousands of large matrices costs
// four launches per block round.
#include <hip/hip_runtime.h>

#include "kernels.hpp"

namespace msq {
namespace {

constexpr int kB = kBlockedWidth;  // 32
constexpr int kP = 2 * kB;         // 64: order of a pair's subproblem
constexpr int kPitch = kP + 1;     // LDS pitch (odd: conflict-free columns)
constexpr int kThreads = 256;      // workgroup of the gather and apply kernels
// A 64 x 64 tile moves as 16-byte vectors of 4 consecutive rows of a column
// (a column of a pair's tile is two runs of 32 rows, so no vector straddles
// them): vector v is rows 4 (v % 16) .. + 3 of column v / 16.
constexpr int kVecs = kP * kP / 4 / kThreads;  // vectors per thread (4)
__device__ __forceinline__ int vec_row(int v) { return 4 * (v % (kP / 4)); }
__device__ __forceinline__ int vec_col(int v) { return v / (kP / 4); }
// Four LDS floats at stride 1 (a column run) or kPitch (a row run), and four
// consecutive ones stored from a vector (the odd pitch rules out 16-byte LDS
// accesses).
__device__ __forceinline__ float4 load4(const float* x, int stride) {
  return make_float4(x[0], x[stride], x[2 * stride], x[3 * stride]);
}
__device__ __forceinline__ void store4(float* x, float4 v) {
  x[0] = v.x;
  x[1] = v.y;
  x[2] = v.z;
  x[3] = v.w;
}

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

__global__ __launch_bounds__(kThreads) void blocked_gather_kernel(
    const uint32_t* __restrict__ active, const float* __restrict__ A, int N, int t,
    double* __restrict__ S) {
  const int nb = N / kB, np = nb / 2;
  const int p = blockIdx.x;
  const float* Am = A + static_cast<size_t>(active[blockIdx.y]) * N * N;
  double* Sp = S + (static_cast<size_t>(blockIdx.y) * np + p) * kP * kP;
  // All of a thread's loads are issued before its first store (a load and
  // store per iteration would wait out one memory latency each).
  float4 v[kVecs];
#pragma unroll
  for (int it = 0; it < kVecs; ++it) {
    const int e = threadIdx.x + kThreads * it, i = vec_row(e), j = vec_col(e);
    v[it] = *reinterpret_cast<const float4*>(
        Am + pair_index(t, p, nb, i) + static_cast<size_t>(N) * pair_index(t, p, nb, j));
  }
#pragma unroll
  for (int it = 0; it < kVecs; ++it) {
    double* d = Sp + 4 * (threadIdx.x + kThreads * it);
    *reinterpret_cast<double2*>(d) = make_double2(v[it].x, v[it].y);
    *reinterpret_cast<double2*>(d + 2) = make_double2(v[it].z, v[it].w);
  }
}

// Block (p, q), p <= q, of the pair grid: A[P_p, P_q] <- U_p^T A[P_p, P_q] U_q,
// and its transpose into A[P_q, P_p].  Two LDS tiles (four workgroups per
// CU): U_p and the block first, T = U_p^T B over U_p, U_q (held in registers
// since the start) over the block, then B' = T U_q over U_q.  Every thread
// issues all its global loads before its first LDS store.
__global__ __launch_bounds__(kThreads) void blocked_apply_a_kernel(
    const uint32_t* __restrict__ active, float* __restrict__ A, int N, int t,
    const float* __restrict__ U) {
  __shared__ float X[kP * kPitch], Y[kP * kPitch];
  const int nb = N / kB, np = nb / 2;
  // blockIdx.x -> (p, q), p <= q, over np (np + 1) / 2 entries.
  int p = 0, rem = blockIdx.x;
  while (rem >= np - p) {
    rem -= np - p;
    ++p;
  }
  const int q = p + rem;
  float* Am = A + static_cast<size_t>(active[blockIdx.y]) * N * N;
  const float* up = U + (static_cast<size_t>(blockIdx.y) * np + p) * kP * kP;
  const float* uq = U + (static_cast<size_t>(blockIdx.y) * np + q) * kP * kP;
  float4 a[kVecs], b[kVecs], c[kVecs];
#pragma unroll
  for (int it = 0; it < kVecs; ++it) {
    const int e = threadIdx.x + kThreads * it, i = vec_row(e), j = vec_col(e);
    a[it] = reinterpret_cast<const float4*>(up)[e];
    b[it] = *reinterpret_cast<const float4*>(
        Am + pair_index(t, p, nb, i) + static_cast<size_t>(N) * pair_index(t, q, nb, j));
    c[it] = reinterpret_cast<const float4*>(uq)[e];
  }
#pragma unroll
  for (int it = 0; it < kVecs; ++it) {
    const int e = threadIdx.x + kThreads * it, o = vec_row(e) + kPitch * vec_col(e);
    store4(X + o, a[it]);
    store4(Y + o, b[it]);
  }
  __syncthreads();
  lds_gemm64<true>(X, Y, X);  // T = U_p^T B, over U_p (B read before the store)
#pragma unroll
  for (int it = 0; it < kVecs; ++it) {
    const int e = threadIdx.x + kThreads * it;
    store4(Y + vec_row(e) + kPitch * vec_col(e), c[it]);
  }
  __syncthreads();
  lds_gemm64<false>(X, Y, Y);  // B' = T U_q, over U_q
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kVecs; ++it) {
    const int e = threadIdx.x + kThreads * it, i = vec_row(e), j = vec_col(e);
    *reinterpret_cast<float4*>(Am + pair_index(t, p, nb, i) +
                               static_cast<size_t>(N) * pair_index(t, q, nb, j)) =
        load4(Y + i + kPitch * j, 1);
    if (p != q)
      *reinterpret_cast<float4*>(Am + pair_index(t, q, nb, i) +
                                 static_cast<size_t>(N) * pair_index(t, p, nb, j)) =
          load4(Y + j + kPitch * i, kPitch);
  }
}

// V[rows of chunk c, P_p] <- V[rows, P_p] U_p (64 rows per workgroup).
__global__ __launch_bounds__(kThreads) void blocked_apply_v_kernel(
    const uint32_t* __restrict__ active, float* __restrict__ V, int N, int t,
    const float* __restrict__ U) {
  __shared__ float Up[kP * kPitch], Vc[kP * kPitch];
  const int nb = N / kB, np = nb / 2;
  const int c = blockIdx.x, p = blockIdx.y;
  float* Vm = V + static_cast<size_t>(active[blockIdx.z]) * N * N;
  const float* up = U + (static_cast<size_t>(blockIdx.z) * np + p) * kP * kP;
  float4 a[kVecs], b[kVecs];
#pragma unroll
  for (int it = 0; it < kVecs; ++it) {
    const int e = threadIdx.x + kThreads * it, i = vec_row(e), j = vec_col(e);
    a[it] = reinterpret_cast<const float4*>(up)[e];
    b[it] = *reinterpret_cast<const float4*>(
        Vm + (kP * c + i) + static_cast<size_t>(N) * pair_index(t, p, nb, j));
  }
#pragma unroll
  for (int it = 0; it < kVecs; ++it) {
    const int e = threadIdx.x + kThreads * it, o = vec_row(e) + kPitch * vec_col(e);
    store4(Up + o, a[it]);
    store4(Vc + o, b[it]);
  }
  __syncthreads();
  lds_gemm64<false>(Vc, Up, Vc);
  __syncthreads();
#pragma unroll
  for (int it = 0; it < kVecs; ++it) {
    const int e = threadIdx.x + kThreads * it, i = vec_row(e), j = vec_col(e);
    *reinterpret_cast<float4*>(Vm + (kP * c + i) +
                               static_cast<size_t>(N) * pair_index(t, p, nb, j)) =
        load4(Vc + i + kPitch * j, 1);
  }
}

// Working copies of member blockIdx.y (pitch N, zero padding): A from the
// binary64 input, V = I.
__global__ void blocked_init_kernel(const BlockedMember* __restrict__ members,
                                    const double* __restrict__ a64, float* __restrict__ A,
                                    float* __restrict__ V, int N) {
  const BlockedMember mb = members[blockIdx.y];
  const int n = static_cast<int>(mb.order), ld = static_cast<int>(mb.ld);
  const double* a = a64 + mb.offset;
  float* Am = A + static_cast<size_t>(blockIdx.y) * N * N;
  float* Vm = V + static_cast<size_t>(blockIdx.y) * N * N;
  const size_t total = static_cast<size_t>(N) * N;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(e % N), j = static_cast<int>(e / N);
    Am[e] = (i < n && j < n) ? static_cast<float>(a[i + static_cast<size_t>(j) * ld]) : 0.0f;
    Vm[e] = i == j ? 1.0f : 0.0f;
  }
}

// Block-wide sums / minimum through LDS (blockDim.x a multiple of 64, at most
// 1024); every thread gets the result.
__device__ double block_sum(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int k = 0; k < nw; ++k) t += red[k];
  return t;
}
__device__ float block_min(float v, float* red) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  float t = INFINITY;
  for (int k = 0; k < nw; ++k) t = fminf(t, red[k]);
  return t;
}

// ||A||_F^2 of member blockIdx.x from its binary64 input.
// Part blockIdx.x of gridDim.x of ||A||_F^2 of member blockIdx.y, from the
// binary64 input, at [member][part].
__global__ __launch_bounds__(512) void blocked_norm_kernel(const BlockedMember* __restrict__ members,
                                                           const double* __restrict__ a64,
                                                           double* __restrict__ norm2) {
  __shared__ double red[16];
  const BlockedMember mb = members[blockIdx.y];
  const int n = static_cast<int>(mb.order), ld = static_cast<int>(mb.ld);
  const double* a = a64 + mb.offset;
  const size_t total = static_cast<size_t>(n) * n;
  double s = 0.0;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const double x = a[(e % n) + (e / n) * ld];
    s += x * x;
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) norm2[static_cast<size_t>(blockIdx.y) * kOffnormParts + blockIdx.x] = s;
}

// Stop-rule measures of active member active[blockIdx.y] (entries 0..n-1 of
// its working copy), part blockIdx.x of gridDim.x: off(A)^2, ||G||_F^2 with
// G_ij = a_ij / (s_i + s_j), s_i = sqrt(a_ii) (float quotient, binary64 sum,
// as in the batched kernel), and min a_ii; written at [member][part].
__global__ __launch_bounds__(512) void blocked_offnorm_kernel(
    const uint32_t* __restrict__ active, const BlockedMember* __restrict__ members,
    const float* __restrict__ A, int N, double* __restrict__ off2, double* __restrict__ g2,
    float* __restrict__ dmin) {
  __shared__ double red[16];
  __shared__ float redf[16];
  const uint32_t mem = active[blockIdx.y];
  const int n = static_cast<int>(members[mem].order);
  const float* Am = A + static_cast<size_t>(mem) * N * N;
  const size_t total = static_cast<size_t>(n) * n;
  double s = 0.0, g = 0.0;
  float d = INFINITY;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(e % n), j = static_cast<int>(e / n);
    const float v = Am[i + static_cast<size_t>(j) * N];
    if (i == j) {
      d = fminf(d, v);
    } else {
      s += static_cast<double>(v) * v;
      const float den = sqrtf(fmaxf(Am[i + static_cast<size_t>(i) * N], 0.0f)) +
                        sqrtf(fmaxf(Am[j + static_cast<size_t>(j) * N], 0.0f));
      const float q = den > 0.0f ? v * __frcp_rn(den) : (v != 0.0f ? INFINITY : 0.0f);
      g += static_cast<double>(q * q);
    }
  }
  s = block_sum(s, red);
  g = block_sum(g, red);
  d = block_min(d, redf);
  if (threadIdx.x == 0) {
    const size_t slot = static_cast<size_t>(mem) * kOffnormParts + blockIdx.x;
    off2[slot] = s;
    g2[slot] = g;
    dmin[slot] = d;
  }
}

// Outputs of member blockIdx.y: V (pitch ld, zero outside the order),
// diagonal and roots (ld entries); bad[member] = 1 if a diagonal entry of
// the order is not positive and finite.
__global__ void blocked_finish_kernel(const BlockedMember* __restrict__ members,
                                      const float* __restrict__ A, const float* __restrict__ Vw,
                                      int N, float* __restrict__ V, float* __restrict__ diag,
                                      float* __restrict__ roots, uint32_t* __restrict__ bad) {
  const BlockedMember mb = members[blockIdx.y];
  const int n = static_cast<int>(mb.order), ld = static_cast<int>(mb.ld);
  const float* Am = A + static_cast<size_t>(blockIdx.y) * N * N;
  const float* Vm = Vw + static_cast<size_t>(blockIdx.y) * N * N;
  float* Vo = V + mb.offset;
  const size_t total = static_cast<size_t>(ld) * ld;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(e % ld), j = static_cast<int>(e / ld);
    Vo[e] = (i < n && j < n) ? Vm[i + static_cast<size_t>(j) * N] : 0.0f;
    if (j == 0) {
      const float d = i < n ? Am[i + static_cast<size_t>(i) * N] : 0.0f;
      diag[mb.s_offset + i] = d;
      roots[mb.s_offset + i] = i < n ? sqrtf(d) : 0.0f;
      if (i < n && !(d > 0.0f && isfinite(d))) atomicOr(bad + blockIdx.y, 1u);
    }
  }
}

__global__ void scatter_status_kernel(const JacobiStatus* __restrict__ src,
                                      const uint32_t* __restrict__ index, uint32_t count,
                                      JacobiStatus* __restrict__ dst) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < count) dst[index[k]] = src[k];
}

// Workgroups per member of an elementwise pass: enough for the whole batch
// to fill the chip, at least 4 per member.
int grid_chunks(size_t elements, uint32_t count) {
  const size_t g = (elements + 255) / 256;
  size_t cap = 16384 / count;
  if (cap < 4) cap = 4;
  return static_cast<int>(g < cap ? g : cap);
}

}  // namespace

uint32_t blocked_padded_order(uint32_t order) { return (order + 63u) & ~63u; }

hipError_t launch_blocked_init(const BlockedMember* members, uint32_t count, const double* a64,
                               float* A, float* V, uint32_t N, double* norm2, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (N % kP != 0 || count > 65535) return hipErrorInvalidValue;
  blocked_init_kernel<<<dim3(grid_chunks(static_cast<size_t>(N) * N, count), count), 256, 0, stream>>>(
      members, a64, A, V, static_cast<int>(N));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  blocked_norm_kernel<<<dim3(blocked_offnorm_parts(count), count), 512, 0, stream>>>(members, a64,
                                                                                 norm2);
  return hipGetLastError();
}

hipError_t launch_blocked_round(const uint32_t* active, uint32_t count, float* A, float* V,
                                uint32_t N, uint32_t round, double* staging, float* U,
                                const JacobiJob* jobs, float* pair_diag, float* pair_roots,
                                JacobiStatus* pair_status, hipStream_t stream) {
  const int nb = static_cast<int>(N) / kB;
  const int np = nb / 2;
  if (N % kP != 0 || nb < 2 || count > 65535) return hipErrorInvalidValue;
  if (count == 0) return hipSuccess;
  const int Ni = static_cast<int>(N), t = static_cast<int>(round);
  blocked_gather_kernel<<<dim3(np, count), kThreads, 0, stream>>>(active, A, Ni, t, staging);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  e = launch_batched_jacobi(jobs, count * np, kP, staging, U, pair_diag, pair_roots, pair_status,
                            stream);
  if (e != hipSuccess) return e;
  const uint32_t blocks = np * (np + 1) / 2;
  blocked_apply_a_kernel<<<dim3(blocks, count), kThreads, 0, stream>>>(active, A, Ni, t, U);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  blocked_apply_v_kernel<<<dim3(N / kP, np, count), kThreads, 0, stream>>>(active, V, Ni, t, U);
  return hipGetLastError();
}

uint32_t blocked_offnorm_parts(uint32_t count) {
  const uint32_t p = count ? (16384u + count - 1) / count : 1u;
  return p < 1u ? 1u : (p > static_cast<uint32_t>(kOffnormParts) ? kOffnormParts : p);
}

hipError_t launch_blocked_offnorm(const uint32_t* active, uint32_t count,
                                  const BlockedMember* members, const float* A, uint32_t N,
                                  double* off2, double* g2, float* dmin, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (count > 65535) return hipErrorInvalidValue;
  blocked_offnorm_kernel<<<dim3(blocked_offnorm_parts(count), count), 512, 0, stream>>>(
      active, members, A, static_cast<int>(N), off2, g2, dmin);
  return hipGetLastError();
}

hipError_t launch_blocked_finish(const BlockedMember* members, uint32_t count, const float* A,
                                 const float* Vw, uint32_t N, float* V, float* diag, float* roots,
                                 uint32_t* bad, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (count > 65535) return hipErrorInvalidValue;
  blocked_finish_kernel<<<dim3(grid_chunks(static_cast<size_t>(N) * N, count), count), 256, 0, stream>>>(
      members, A, Vw, static_cast<int>(N), V, diag, roots, bad);
  return hipGetLastError();
}

hipError_t launch_scatter_status(const JacobiStatus* src, const uint32_t* index, uint32_t count,
                                 JacobiStatus* dst, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  scatter_status_kernel<<<(count + 255) / 256, 256, 0, stream>>>(src, index, count, dst);
  return hipGetLastError();
}

}  // namespace msq
