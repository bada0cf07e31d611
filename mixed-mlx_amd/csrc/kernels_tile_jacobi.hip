// Blocked Jacobi in LDS for orders 17..256: pair subproblems diagonalised by
// scalar rotations, the rest of each rotation applied as FP32 MFMA products.
//
// Reference: symmetricJacobi (src/metal_runtime.cpp:584-646) with the
// rotation and element formulas of jacobi_rotations_f32 /
// jacobi_similarity_vectors_f32 (metal_mixed_sqrt.metal:1331-1411) and the
// circle-method schedule of makeJacobiSchedule (src/metal_solver.cpp:50-79),
// here applied to blocks of indices.
//
// The padded order N (a multiple of 16) is cut into NB = N / 8 blocks of 8
// indices.  A round pairs the blocks on the circle (M = N / 16 disjoint block
// pairs); the pair at slot k, blocks b_k = P_t(k) and b'_k = P_t(NB - 1 - k),
// spans 16 indices (local 0..7 = b_k, 8..15 = b'_k).  One wave per pair:
//  1. subproblem: S = A[pair, pair] is rotated in the wave's registers --
//     in round 0 of a sweep by the circle method over its 16 local indices
//     (15 rounds of 8 disjoint rotations), in the other rounds by the 64
//     pairs across its two blocks (8 rounds) -- the rotations accumulated
//     into the pair's 16 x 16 orthogonal Q_k (Q_k <- Q_k J per rotation);
//     S goes back to LDS and Q_k to its LDS slot;
//  2. after a barrier, every off-diagonal 16 x 16 tile-block A[pair k, pair
//     l], k < l, becomes Q_k^T A[k, l] Q_l (two MFMA products), and every
//     wave updates its 16 rows of V: V[:, pair k] <- V[:, pair k] Q_k.
// A sweep is NB - 1 rounds and rotates every pair of indices exactly once,
// as a scalar sweep does, in another order.  The convergence test, its
// quantities and stop rules are the scalar kernel's (off(A) in binary64
// after every sweep).
//
// Storage.  A: the packed upper triangle of 8 x 8 tiles, tile (I, J), I <= J,
// row-major (diagonal tiles use their upper triangle only), so every entry
// a_ij = a_ji has one LDS word: N = 256 takes 528 tiles, 132 KiB.  Q_k: 16 x
// 16 floats, each column four 16-byte chunks of the rows congruent mod 4
// (q_at), so each lane's four MFMA operands of a product are one read.  V lives in registers
// for the whole solve: lane (j, g) of wave w (j = lane % 16, g = lane / 16)
// holds row 16 w + j of V at the columns of every block that are congruent
// to g mod 4, two per block, by slot (vl[k] = slot k, vh[k] = slot NB-1-k);
// between rounds the values move one slot along the circle (register moves).
//
// MFMA operand layouts (v_mfma_f32_16x16x4_f32): lane l supplies A[l % 16][k]
// and B[k][l % 16] with k = 4 s + l / 16 at step s of four; accumulator r of
// lane l is D[4 (l / 16) + r][l % 16].  With sigma(i) = 4 (i % 4) + i / 4:
//  * V: D = P W with P[i][a] = Q_k[a][sigma(i)] and W[a][j] = V[j][pair col a]
//    leaves W'[sigma(4 g + r)] = W'[4 r + g] in accumulator r of lane group g:
//    the same four registers the B operand came from.
//  * tile-block: Z = T' Q_l with T'[i] = T[sigma(i)] leaves Z[4 r + g][j] in
//    accumulator r; that is the B operand of Q_k^T Z at step r (k = 4 r + g),
//    so the second product reads it straight from the accumulators.
#include <hip/hip_runtime.h>

#include <math.h>

#include "jacobi_math.hpp"
#include "kernels.hpp"

namespace msq {
namespace {

using jmath::Block;
using jmath::participant;
using jmath::rotate_block;
using jmath::rotate_rows_cols;
using jmath::rotation_angles;

typedef float f4 __attribute__((ext_vector_type(4)));
typedef float f2 __attribute__((ext_vector_type(2)));
// LDS pointers stay in the local address space (32-bit ds_* addressing).
#define MSQ_LDS __attribute__((address_space(3)))

#ifndef TILE_MIN_PADDED
#define TILE_MIN_PADDED 32
#endif
#ifndef PAIRS_PER_WAVE
#define PAIRS_PER_WAVE 1
#endif
// Pair subproblems (and 16-row groups of V) per wave.
constexpr int kPairsPerWave = PAIRS_PER_WAVE;
constexpr int kTileThreads = 1024 / kPairsPerWave;
constexpr int kTileWaves = kTileThreads / 64;

__host__ __device__ constexpr int tile_index(int I, int J) { return J * (J + 1) / 2 + I; }
__device__ __forceinline__ int sigma16(int i) { return ((i & 3) << 2) | (i >> 2); }
// Offset of Q entry (a, b) inside a pair's 256 floats: column b holds four
// 16-byte chunks, chunk g the rows g, g + 4, g + 8, g + 12, at position
// g ^ (b / 4) (so the 16-byte reads of 16 different columns by one lane group
// fall in different LDS banks).
__device__ __forceinline__ int q_at(int a, int b) {
  return b * 16 + (((a & 3) ^ (b >> 2)) << 2) + (a >> 2);
}
__device__ __forceinline__ int q_chunk(int b, int g) { return b * 16 + ((g ^ (b >> 2)) << 2); }

template <int NP>
struct TileLds {
  static constexpr int NB = NP / 8;
  static constexpr int M = NP / 16;
  static constexpr int kTiles = NB * (NB + 1) / 2;
  static constexpr size_t kA = 0;
  static constexpr size_t kQ = kA + static_cast<size_t>(kTiles) * 64 * sizeof(float);
  static constexpr size_t kRot = kQ + static_cast<size_t>(M) * 256 * sizeof(float);
  static constexpr size_t kRed = kRot + static_cast<size_t>(2 * kTileWaves) * 8 * sizeof(float2);
  static constexpr size_t kRedf = kRed + 2 * kTileWaves * sizeof(double);
  static constexpr size_t kCtrl = kRedf + 16 * sizeof(float);
  // Off-diagonal tile-blocks (k, l), k < l, row-major: k | l << 8.
  static constexpr size_t kItemTable = kCtrl + 4 * sizeof(int);
  static constexpr size_t kBytes = kItemTable + static_cast<size_t>(M * (M - 1) / 2) * 2;
  static_assert(kQ % 16 == 0 && kRot % 16 == 0 && kRed % 8 == 0, "LDS alignment");
  static_assert(NP * sizeof(float) <= static_cast<size_t>(M) * 256 * sizeof(float),
                "square roots fit the Q area");
};
constexpr size_t tile_lds_bytes() { return TileLds<256>::kBytes; }
static_assert(tile_lds_bytes() + 16 <= 160 * 1024, "tile Jacobi LDS");

// Waits for this wave's LDS accesses (the subproblem's rounds are ordered
// within one wave only) and keeps the compiler from moving memory accesses
// across.
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Sums of two doubles over the workgroup in a fixed order: a butterfly inside
// each wave, then the wave totals in wave order.  Every thread returns the
// sums.
__device__ __forceinline__ void tile_sum2(double& a, double& b, MSQ_LDS double* red) {
  for (int s = 32; s > 0; s >>= 1) {
    a += __shfl_xor(a, s);
    b += __shfl_xor(b, s);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    red[2 * wave] = a;
    red[2 * wave + 1] = b;
  }
  __syncthreads();
  double sa = 0.0, sb = 0.0;
  const int waves = static_cast<int>(blockDim.x) >> 6;
  for (int w = 0; w < waves; ++w) {
    sa += red[2 * w];
    sb += red[2 * w + 1];
  }
  a = sa;
  b = sb;
  __syncthreads();
}

// Tile (I, J), I <= J, of the packed tile index t.
__device__ __forceinline__ void tile_coords(int t, int& I, int& J) {
  int j = static_cast<int>((sqrtf(8.0f * static_cast<float>(t) + 1.0f) - 1.0f) * 0.5f);
  while (j > 0 && (j * (j + 1)) / 2 > t) --j;
  while (((j + 1) * (j + 2)) / 2 <= t) ++j;
  J = j;
  I = t - (j * (j + 1)) / 2;
}

// LDS float offset of a_xy for x in block X (row x8 inside it) and y in block
// Y (column y8), X != Y.
__device__ __forceinline__ int cross_at(int X, int x8, int Y, int y8) {
  return X < Y ? tile_index(X, Y) * 64 + x8 * 8 + y8 : tile_index(Y, X) * 64 + y8 * 8 + x8;
}

// One pair subproblem in registers.  The 16 x 16 block S and the
// accumulated rotations Q are held in slot order: slot y of inner round tau
// holds local index P(y).  Lane (x, g) holds row x of S at the slots c0 = 2g,
// c1 = 15 - 2g, c2 = 2g + 1, c3 = 14 - 2g (s0..s3), and row x (a local
// index, fixed) of Q at the same slots (q0..q3).  Slot k pairs with slot
// 15 - k, so a row's partner is lane 15 - x of the same 16 (DPP row mirror)
// and a column's partner is the lane's own other register.
struct SubState {
  float s0, s1, s2, s3;
  float q0, q1, q2, q3;
};

__device__ __forceinline__ float dpp_mirror(float v) {  // lane 15 - x of the row
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x140, 0xf, 0xf, true));
}
__device__ __forceinline__ float dpp_ror1(float v) {  // lane (x - 1) mod 16
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x121, 0xf, 0xf, true));
}
__device__ __forceinline__ float dpp_ror9(float v) {  // lane (x - 9) mod 16
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x129, 0xf, 0xf, true));
}
__device__ __forceinline__ float dpp_shr1(float v) {  // lane x - 1 (lane 0: 0)
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x111, 0xf, 0xf, true));
}

// One inner round of two subproblems (independent, interleaved for
// instruction-level parallelism).  FULL: the circle over 16 (slot y >= 1 holds
// local index 1 + ((y - 1 - tau) mod 15), slot 0 index 0; the slots then move
// y -> y + 1, 15 -> 1).  Otherwise: slot k < 8 holds local k and slot y >= 8
// holds 8 + ((y - 8 - tau) mod 8), so slot k pairs local k with
// 8 + ((7 - k - tau) mod 8) (the pairs across the two blocks); then the slots
// 8..15 move y -> y + 1, 15 -> 8.  rw[u]: subproblem u's 8 (c, sigma)
// records, sigma the coefficient of the pair's lower slot k (new row k =
// c row k + sigma row 15 - k; the upper slot has -sigma).
template <bool FULL, int P>
__device__ __forceinline__ void sub_round(SubState (&st)[P], int tau, int x, int g,
                                          MSQ_LDS f2* const (&rw)[P]) {
  const int kx = min(x, 15 - x);
  const bool even = (kx & 1) == 0, lo = x < 8;
  // Rotation inputs, held by lane (x, g) with g = min(x, 15 - x) / 2: its
  // diagonal entry and the entry of its row in its partner's column.  Values,
  // not struct fields, go into the selects (a select between two fields
  // becomes a load from a select of their addresses, which keeps the state in
  // private memory).
  // The partner's values come by DPP with every lane active (a DPP read of an
  // inactive lane returns nothing).
  float diag[P], off[P], pdiag[P], poff[P];
#pragma unroll
  for (int u = 0; u < P; ++u) {
    float v0 = st[u].s0, v1 = st[u].s1, v2 = st[u].s2, v3 = st[u].s3;
    asm volatile("" : "+v"(v0), "+v"(v1), "+v"(v2), "+v"(v3));
    diag[u] = lo ? (even ? v0 : v2) : (even ? v1 : v3);
    off[u] = lo ? (even ? v1 : v3) : (even ? v0 : v2);
    pdiag[u] = dpp_mirror(diag[u]);
    poff[u] = dpp_mirror(off[u]);
  }
  if (lo && g == (x >> 1)) {
    int a, b;  // local indices at slots x and 15 - x
    if constexpr (FULL) {
      a = participant(tau, x, 16);
      b = participant(tau, 15 - x, 16);
    } else {
      a = x;
      b = 8 + ((7 - x - tau) & 7);
    }
    // The pair p < q: a_pp, a_qq, and a_pq from row p.
    const bool xp = a < b;
#pragma unroll
    for (int u = 0; u < P; ++u) {
      const float2 r = rotation_angles<1, 0>(xp ? diag[u] : pdiag[u], xp ? pdiag[u] : diag[u],
                                       xp ? off[u] : poff[u]);
      rw[u][x] = f2{r.x, xp ? -r.y : r.y};
    }
  }
  wave_lds_sync();
  f2 rr[P], ra[P], rb[P];
#pragma unroll
  for (int u = 0; u < P; ++u) {
    rr[u] = rw[u][kx];
    ra[u] = rw[u][2 * g];
    rb[u] = rw[u][2 * g + 1];
  }
  wave_lds_sync();  // the next round's records overwrite these
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int u = 0; u < P; ++u) {
    SubState& q = st[u];
    // Rows: new row x = c row x + sigma_x row 15 - x.
    const float sg = lo ? rr[u].y : -rr[u].y;
    const float p0 = dpp_mirror(q.s0), p1 = dpp_mirror(q.s1);
    const float p2 = dpp_mirror(q.s2), p3 = dpp_mirror(q.s3);
    float s0 = rr[u].x * q.s0 + sg * p0;
    float s1 = rr[u].x * q.s1 + sg * p1;
    float s2 = rr[u].x * q.s2 + sg * p2;
    float s3 = rr[u].x * q.s3 + sg * p3;
    // Columns of S and Q: slots 2g / 15 - 2g with ra, 2g + 1 / 14 - 2g with rb.
    q.s0 = ra[u].x * s0 + ra[u].y * s1;
    q.s1 = ra[u].x * s1 - ra[u].y * s0;
    q.s2 = rb[u].x * s2 + rb[u].y * s3;
    q.s3 = rb[u].x * s3 - rb[u].y * s2;
    const float t0 = ra[u].x * q.q0 + ra[u].y * q.q1;
    const float t1 = ra[u].x * q.q1 - ra[u].y * q.q0;
    const float t2 = rb[u].x * q.q2 + rb[u].y * q.q3;
    const float t3 = rb[u].x * q.q3 - rb[u].y * q.q2;
    q.q0 = t0;
    q.q1 = t1;
    q.q2 = t2;
    q.q3 = t3;
    // Slots move for the next round: rows of S (lanes), columns of S and Q
    // (registers, and lane groups g +- 1).
    if constexpr (FULL) {
      auto rows = [x](float v) {
        const float r1 = dpp_ror1(v);
        const float w = dpp_shr1(r1);
        return x == 0 ? v : (x == 1 ? w : r1);
      };
      q.s0 = rows(q.s0);
      q.s1 = rows(q.s1);
      q.s2 = rows(q.s2);
      q.s3 = rows(q.s3);
      // new slot 2g <- 2g - 1 (group g - 1's c2; slot 0 stays), 2g + 1 <- 2g
      // (slot 1 <- 15), 15 - 2g <- 14 - 2g, 14 - 2g <- 13 - 2g (group g + 1's
      // c1; slot 8 <- 7).
      const int dn = lane >= 16 ? lane - 16 : lane, up = lane < 48 ? lane + 16 : lane;
      auto cols = [g, dn, up](float& v0, float& v1, float& v2, float& v3) {
        const float from_dn = __shfl(v2, dn);
        const float from_up = __shfl(v1, up);
        const float n0 = g == 0 ? v0 : from_dn;
        const float n2 = g == 0 ? v1 : v0;
        const float n1 = v3;
        const float n3 = g == 3 ? v2 : from_up;
        v0 = n0;
        v1 = n1;
        v2 = n2;
        v3 = n3;
      };
      cols(q.s0, q.s1, q.s2, q.s3);
      cols(q.q0, q.q1, q.q2, q.q3);
    } else {
      auto rows = [x](float v) {
        const float r1 = dpp_ror1(v);
        const float r9 = dpp_ror9(v);
        return x < 8 ? v : (x == 8 ? r9 : r1);
      };
      q.s0 = rows(q.s0);
      q.s1 = rows(q.s1);
      q.s2 = rows(q.s2);
      q.s3 = rows(q.s3);
      // new slot 15 - 2g <- 14 - 2g (own c3), 14 - 2g <- 13 - 2g (group
      // g + 1's c1; slot 8 <- 15, group 0's c1); slots 0..7 stay.
      const int up = (lane + 16) & 63;
      const float su = __shfl(q.s1, up), qu = __shfl(q.q1, up);
      q.s1 = q.s3;
      q.s3 = su;
      q.q1 = q.q3;
      q.q3 = qu;
    }
  }
}

// V <- V Q_k for pair k on this wave's row groups: D = P W with P[i][a] =
// Q_k[a][sigma(i)] and W[a][j] = V[j][pair column a] leaves W'[4 r + g] in
// accumulator r of lane group g, the registers the B operand came from.
template <int P, int M>
__device__ __forceinline__ void v_update(float (&vl)[P][M][2], float (&vh)[P][M][2],
                                         const MSQ_LDS float* Qs, int k, int qoff) {
  const f4 qa = *reinterpret_cast<const MSQ_LDS f4*>(Qs + k * 256 + qoff);
#pragma unroll
  for (int h2 = 0; h2 < P; ++h2) {
    f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(qa[0], vl[h2][k][0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(qa[1], vl[h2][k][1], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(qa[2], vh[h2][k][0], acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(qa[3], vh[h2][k][1], acc, 0, 0, 0);
    vl[h2][k][0] = acc[0];
    vl[h2][k][1] = acc[1];
    vh[h2][k][0] = acc[2];
    vh[h2][k][1] = acc[3];
  }
}

// Offset of s_ab of a pair subproblem, local indices a, b (0..7: block bp,
// 8..15: block bq; tpp, tqq: the offsets of the two diagonal tiles); s_ab
// and s_ba share one word.
__device__ __forceinline__ int sub_at(int a, int b, int bp, int bq, int tpp, int tqq) {
  const int ia = a & 7, ib = b & 7;
  if ((a >> 3) == (b >> 3)) return (a < 8 ? tpp : tqq) + min(ia, ib) * 8 + max(ia, ib);
  return a < 8 ? cross_at(bp, ia, bq, ib) : cross_at(bp, ib, bq, ia);
}

template <int NP>
__device__ __forceinline__ void tile_jacobi_solve(const JacobiJob& job,
                                                  const double* __restrict__ a64,
                                                  float* __restrict__ vectors,
                                                  float* __restrict__ diagonal,
                                                  float* __restrict__ roots,
                                                  JacobiStatus* __restrict__ status,
                                                  MSQ_LDS unsigned char* __restrict__ smem) {
  using L = TileLds<NP>;
  constexpr int NB = L::NB;
  constexpr int M = L::M;
  constexpr int kTiles = L::kTiles;
  constexpr int kItems = M * (M - 1) / 2;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  // Wave-uniform (scalar registers for everything derived from it).
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int nthreads = static_cast<int>(blockDim.x);
  const int n = static_cast<int>(job.order);
  const int ld = static_cast<int>(job.ld);
  MSQ_LDS float* const At = reinterpret_cast<MSQ_LDS float*>(smem + L::kA);
  MSQ_LDS float* const Qs = reinterpret_cast<MSQ_LDS float*>(smem + L::kQ);
  MSQ_LDS f2* const rot = reinterpret_cast<MSQ_LDS f2*>(smem + L::kRot);
  MSQ_LDS double* const red = reinterpret_cast<MSQ_LDS double*>(smem + L::kRed);
  MSQ_LDS float* const redf = reinterpret_cast<MSQ_LDS float*>(smem + L::kRedf);
  MSQ_LDS int* const ctrl = reinterpret_cast<MSQ_LDS int*>(smem + L::kCtrl);
  MSQ_LDS uint16_t* const items = reinterpret_cast<MSQ_LDS uint16_t*>(smem + L::kItemTable);
  for (int it = tid; it < kItems; it += nthreads) {
    int k = 0, l = it;
    while (l >= M - 1 - k) {
      l -= M - 1 - k;
      ++k;
    }
    items[it] = static_cast<uint16_t>(k | (k + 1 + l) << 8);
  }

  // A into the tiles: element (r, c) of tile (I, J) = a_{8I+r, 8J+c}; rows
  // fastest, so that 8 lanes read 64 consecutive bytes of a column.
  const double* A = a64 + job.a_offset;
  double norm2 = 0.0;
#pragma unroll 1
  for (int e = tid; e < kTiles * 64; e += nthreads) {
    const int t = e >> 6, c = (e >> 3) & 7, r = e & 7;
    int I, J;
    tile_coords(t, I, J);
    const int i = 8 * I + r, j = 8 * J + c;
    float v = 0.0f;
    if (i < n && j < n && (I != J || r <= c)) {
      const double a = A[static_cast<size_t>(i) + static_cast<size_t>(j) * ld];
      v = static_cast<float>(a);
      norm2 += (i == j ? 1.0 : 2.0) * a * a;
    }
    At[t * 64 + r * 8 + c] = v;
  }
  double unused = 0.0;
  tile_sum2(norm2, unused, red);
  const double matrix_norm = sqrt(norm2);

  // V = I in registers.  Lane (j, g) of wave w: rows 16 (2 w + h2) + j,
  // h2 = 0, 1 (row groups below M), columns 8 b + 4 h + g of the block b at
  // each slot (round 0: slot x holds block x).
  const int vj = lane & 15, vg = lane >> 4;
  constexpr int P = kPairsPerWave;
  float vl[P][M][2], vh[P][M][2];
#pragma unroll
  for (int h2 = 0; h2 < P; ++h2) {
    const int vrow = 16 * (P * wave + h2) + vj;
#pragma unroll
    for (int k = 0; k < M; ++k) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        vl[h2][k][h] = vrow == 8 * k + 4 * h + vg ? 1.0f : 0.0f;
        vh[h2][k][h] = vrow == 8 * (NB - 1 - k) + 4 * h + vg ? 1.0f : 0.0f;
      }
    }
  }
  // This wave's pairs (the second one a copy of the first, storing nothing,
  // when M is odd and this is the last wave with pairs).
  const bool has_pairs = P * wave < M;
  const int pairs = min(P, M - P * wave);

  const uint32_t fixed = job.fixed_sweeps;
  const uint32_t cap = fixed ? fixed : job.max_sweeps;
  constexpr int rounds_per_sweep = NB - 1;
  double previous = INFINITY;
  double off = 0.0, rate = INFINITY;
  float min_diag = 0.0f;
  uint32_t sweeps = 0;
  int t = 0;
  JacobiStatus* st = status + job.status_index;

#pragma unroll 1
  for (;;) {
    const int round = t % rounds_per_sweep;
    // ---- 1. subproblems: wave w rotates pairs 2w and 2w + 1 in registers
    // (round 0 of a sweep: all their index pairs; later rounds: the pairs
    // across their two blocks), writes them back and stores Q_k.
    if (has_pairs) {
      const int x = lane & 15, g = lane >> 4;
      const int c0 = 2 * g, c1 = 15 - 2 * g, c2 = 2 * g + 1, c3 = 14 - 2 * g;
      int bp[P], bq[P], tpp[P], tqq[P];
      SubState sub[P];
#pragma unroll
      for (int u = 0; u < P; ++u) {
        const int k = P * wave + (u < pairs ? u : 0);
        bp[u] = participant(round, k, NB);
        bq[u] = participant(round, NB - 1 - k, NB);
        tpp[u] = tile_index(bp[u], bp[u]) * 64;
        tqq[u] = tile_index(bq[u], bq[u]) * 64;
        sub[u] = SubState{At[sub_at(x, c0, bp[u], bq[u], tpp[u], tqq[u])],
                          At[sub_at(x, c1, bp[u], bq[u], tpp[u], tqq[u])],
                          At[sub_at(x, c2, bp[u], bq[u], tpp[u], tqq[u])],
                          At[sub_at(x, c3, bp[u], bq[u], tpp[u], tqq[u])],
                          x == c0 ? 1.0f : 0.0f, x == c1 ? 1.0f : 0.0f,
                          x == c2 ? 1.0f : 0.0f, x == c3 ? 1.0f : 0.0f};
      }
      MSQ_LDS f2* rw[P];
#pragma unroll
      for (int u = 0; u < P; ++u) rw[u] = rot + (P * wave + u) * 8;
      if (round == 0) {
#pragma unroll 1
        for (int tau = 0; tau < 15; ++tau) sub_round<true, P>(sub, tau, x, g, rw);
      } else {
#pragma unroll 1
        for (int tau = 0; tau < 8; ++tau) sub_round<false, P>(sub, tau, x, g, rw);
      }
      // Slot order is local order again (15 circle rounds, 8 half-rotations).
#pragma unroll
      for (int u = 0; u < P; ++u) {
        if (u >= pairs) break;
        const float sv[4] = {sub[u].s0, sub[u].s1, sub[u].s2, sub[u].s3};
        const float qv[4] = {sub[u].q0, sub[u].q1, sub[u].q2, sub[u].q3};
        const int cols[4] = {c0, c1, c2, c3};
        MSQ_LDS float* const Q = Qs + (P * wave + u) * 256;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int y = cols[j];
          const bool own = (x >> 3) == (y >> 3) ? x <= y : x < y;
          if (own) At[sub_at(x, y, bp[u], bq[u], tpp[u], tqq[u])] = sv[j];
          Q[q_at(x, y)] = qv[j];
        }
      }
    }
    __syncthreads();

    // ---- 2. products: the off-diagonal tile-blocks (k, l), k < l (item
    // wave + 16 j in pass j), and this wave's V update, VPJ pairs per pass
    // beside the item (independent MFMA chains in one instruction stream).
    // Per item the four tiles (row half a: block bk[a], column half b: block
    // bl[b]) have wave-uniform offsets and orientations; a lane only selects
    // among them by its halves.
    {
      const int i = lane & 15, g = lane >> 4;
      const int rho = sigma16(i);
      const int hr = rho >> 3;  // row half of this lane's A-operand row
      // In-tile offsets of T[rho][4 s + g] when the stored tile is (row
      // block, column block) (offr) or (column block, row block) (offc).
      int offr[4], offc[4];
#pragma unroll
      for (int s2 = 0; s2 < 4; ++s2) {
        const int y8 = 4 * (s2 & 1) + g;
        offr[s2] = (rho & 7) * 8 + y8;
        offc[s2] = y8 * 8 + (rho & 7);
      }
      const int hs = g >> 1, hc = i >> 3, r8 = 4 * (g & 1), i8 = i & 7;
      const int qoff = q_chunk(rho, g);
      constexpr int J = (kItems + kTileWaves - 1) / kTileWaves;
      constexpr int VPJ = (M + J - 1) / J;
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int it = wave + kTileWaves * j;
        if (it < kItems) {
          const int kl = __builtin_amdgcn_readfirstlane(static_cast<int>(items[it]));
          const int k = kl & 0xff, l = kl >> 8;
          const int bk[2] = {participant(round, k, NB), participant(round, NB - 1 - k, NB)};
          const int bl[2] = {participant(round, l, NB), participant(round, NB - 1 - l, NB)};
          // Tile (row half a, column half b): its offset and whether it is
          // stored as (row block, column block) (X < Y) or transposed.
          auto tile_base = [](int X, int Y) {
            const int lo = min(X, Y), hi = max(X, Y);
            return (hi * (hi + 1) / 2 + lo) * 64;
          };
          const int base0 = tile_base(bk[0], bl[0]), base1 = tile_base(bk[0], bl[1]);
          const int base2 = tile_base(bk[1], bl[0]), base3 = tile_base(bk[1], bl[1]);
          const int upm = (bk[0] < bl[0] ? 1 : 0) | (bk[0] < bl[1] ? 2 : 0) |
                          (bk[1] < bl[0] ? 4 : 0) | (bk[1] < bl[1] ? 8 : 0);
          // This lane's row half hr: its two tiles and their orientations,
          // selected once (explicit selects, pinned: a select between array
          // elements by a lane value becomes a private-memory array access).
          int rb0 = hr ? base2 : base0, rb1 = hr ? base3 : base1;
          int ru = hr ? upm >> 2 : upm;
          asm volatile("" : "+v"(rb0), "+v"(rb1), "+v"(ru));
          // Z = T' Q_l: A operand T[sigma(i)][4 s + g].
          float ta[4];
#pragma unroll
          for (int s2 = 0; s2 < 4; ++s2) {
            const int b0 = (s2 >> 1) ? rb1 : rb0;
            ta[s2] = At[b0 + (((ru >> (s2 >> 1)) & 1) ? offr[s2] : offc[s2])];
          }
          const f4 ql = *reinterpret_cast<const MSQ_LDS f4*>(Qs + l * 256 + q_chunk(i, g));
          const f4 qk = *reinterpret_cast<const MSQ_LDS f4*>(Qs + k * 256 + q_chunk(i, g));
          f4 z = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int s2 = 0; s2 < 4; ++s2)
            z = __builtin_amdgcn_mfma_f32_16x16x4f32(ta[s2], ql[s2], z, 0, 0, 0);
          f4 o = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
          for (int s2 = 0; s2 < 4; ++s2)
            o = __builtin_amdgcn_mfma_f32_16x16x4f32(qk[s2], z[s2], o, 0, 0, 0);
          // o[r] = T_new[4 g + r][j], j = i: row half g / 2, column half i / 8.
          const int t = 2 * hs + hc;
          const int b0 = hs ? (hc ? base3 : base2) : (hc ? base1 : base0);
          if ((upm >> t) & 1) {
            MSQ_LDS float* dst = At + b0 + r8 * 8 + i8;
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[8 * r] = o[r];
          } else {
            *reinterpret_cast<MSQ_LDS f4*>(At + b0 + i8 * 8 + r8) = o;
          }
        }
        if (has_pairs) {
#pragma unroll
          for (int v = 0; v < VPJ; ++v)
            if (j * VPJ + v < M) v_update<P, M>(vl, vh, Qs, j * VPJ + v, qoff);
        }
      }
    }
    __syncthreads();
    // V: every block moves one slot along the circle (slot x -> x + 1,
    // NB - 1 -> 1, 0 stays).
#pragma unroll
    for (int h2 = 0; h2 < P; ++h2) {
      const float l1[2] = {vh[h2][0][0], vh[h2][0][1]};
      const float hl[2] = {vl[h2][M - 1][0], vl[h2][M - 1][1]};
#pragma unroll
      for (int k = M - 1; k >= 2; --k) {
        vl[h2][k][0] = vl[h2][k - 1][0];
        vl[h2][k][1] = vl[h2][k - 1][1];
      }
      vl[h2][1][0] = l1[0];
      vl[h2][1][1] = l1[1];
#pragma unroll
      for (int k = 0; k < M - 1; ++k) {
        vh[h2][k][0] = vh[h2][k + 1][0];
        vh[h2][k][1] = vh[h2][k + 1][1];
      }
      vh[h2][M - 1][0] = hl[0];
      vh[h2][M - 1][1] = hl[1];
    }
    ++t;
    if (t % rounds_per_sweep != 0) continue;

    // ---- Convergence after each sweep.
    sweeps = static_cast<uint32_t>(t / rounds_per_sweep);
    const bool at_limit = sweeps >= cap;
    MSQ_LDS float* sroot = Qs;  // idle until the next round
#pragma unroll 1
    for (int i = tid; i < NP; i += nthreads) {
      const float d = At[tile_index(i >> 3, i >> 3) * 64 + (i & 7) * 9];
      sroot[i] = i < n ? sqrtf(fmaxf(d, 0.0f)) : 0.0f;
    }
    if (tid == 0) redf[0] = INFINITY;
    __syncthreads();
    double off2 = 0.0, g2 = 0.0;
    float dmin = INFINITY;
    // Thread tid takes entry tid % 64 of the tiles tid / 64 + (blockDim / 64) q,
    // the tile coordinates advanced in place.
    int I, J;
    tile_coords(tid >> 6, I, J);
#pragma unroll 1
    for (int e = tid; e < kTiles * 64; e += nthreads) {
      const int r = (e >> 3) & 7, c = e & 7;
      if (e != tid) {
        I += nthreads >> 6;
        while (I > J) {
          I -= J + 1;
          ++J;
        }
      }
      const int i = 8 * I + r, j = 8 * J + c;
      if (i >= n || j >= n || (I == J && r > c)) continue;
      const float v = At[e];
      if (i == j) {
        dmin = fminf(dmin, v);
      } else {
        off2 += 2.0 * static_cast<double>(v) * static_cast<double>(v);
        const float den = sroot[i] + sroot[j];
        const float gg = den > 0.0f ? v * __frcp_rn(den) : (v != 0.0f ? INFINITY : 0.0f);
        g2 += 2.0 * static_cast<double>(gg * gg);
      }
    }
    if (dmin < INFINITY) atomicMin((float*)redf, dmin);
    tile_sum2(off2, g2, red);
    if (tid == 0) {
      const float d = redf[0];
      off = sqrt(off2);
      min_diag = d;
      rate = d > 0.0f ? sqrt(g2) / sqrt(static_cast<double>(d)) : INFINITY;
      if (sweeps >= 1 && sweeps <= kJacobiHistory)
        st->off_history[sweeps - 1] = static_cast<float>(off / matrix_norm);
      int stop = at_limit ? 1 : 0;
      if (!fixed) {
        double tol = static_cast<double>(job.off_tol) * matrix_norm;
        if (job.off_tol_lambda > 0.0f && d > 0.0f) {
          const double tl = static_cast<double>(job.off_tol_lambda) * d;
          tol = tl > tol ? tl : tol;
        }
        stop = stop || off <= tol ||
               (job.correction_rate > 0.0f && rate <= static_cast<double>(job.correction_rate)) ||
               (off > static_cast<double>(job.stagnation) * previous &&
                off <= static_cast<double>(job.floor_tol) * matrix_norm);
      }
      previous = off;
      ctrl[0] = stop;
    }
    __syncthreads();
    const int stop = ctrl[0];
    __syncthreads();
    if (stop) break;
  }

  // ---- Outputs: diagonal and roots (ld entries), V (ld x ld, zero outside
  // the order).
  float* dg = diagonal + job.s_offset;
  float* rt = roots + job.s_offset;
  int bad = 0;
#pragma unroll 1
  for (int i = tid; i < ld; i += nthreads) {
    const float d = i < n ? At[tile_index(i >> 3, i >> 3) * 64 + (i & 7) * 9] : 0.0f;
    dg[i] = d;
    rt[i] = i < n ? sqrtf(d) : 0.0f;
    if (i < n && !(d > 0.0f && isfinite(d))) bad = 1;
  }
  float* Vg = vectors + job.v_offset;
#pragma unroll 1
  for (int e = tid; e < ld * ld; e += nthreads) {
    const int i = e % ld;
    const int x = e / ld;
    if (i >= n || x >= n) Vg[e] = 0.0f;
  }
  {
    const int round = t % rounds_per_sweep;
#pragma unroll
    for (int h2 = 0; h2 < P; ++h2) {
      const int vrow = 16 * (P * wave + h2) + vj;
      if (vrow >= n) continue;
#pragma unroll
      for (int k = 0; k < M; ++k) {
        const int bl = participant(round, k, NB), bh = participant(round, NB - 1 - k, NB);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          const int cl = 8 * bl + 4 * h + vg, ch = 8 * bh + 4 * h + vg;
          if (cl < n) Vg[vrow + static_cast<size_t>(cl) * ld] = vl[h2][k][h];
          if (ch < n) Vg[vrow + static_cast<size_t>(ch) * ld] = vh[h2][k][h];
        }
      }
    }
  }
  if (tid == 0) ctrl[2] = 0;
  __syncthreads();
  if (bad) ctrl[2] = 1;
  __syncthreads();
  bad = ctrl[2];
  if (tid == 0) {
    st->sweeps = sweeps;
    st->flags = bad ? kJacobiInvalidSpectrum : 0u;
    st->off_norm = off;
    st->matrix_norm = matrix_norm;
    st->min_diagonal = min_diag;
    st->correction_rate = static_cast<float>(rate);
  }
}

// One job: a call, not inlined, so that the orders' bodies do not share one
// register allocation.
template <int NP>
__device__ __attribute__((noinline)) void tile_job(
    const JacobiJob* __restrict__ job, const double* __restrict__ a64, float* __restrict__ vectors,
    float* __restrict__ diagonal, float* __restrict__ roots, JacobiStatus* __restrict__ status,
    __attribute__((address_space(3))) unsigned char* lds) {
  tile_jacobi_solve<NP>(*job, a64, vectors, diagonal, roots, status, lds);
}

// Persistent launch: one 1024-thread workgroup per CU takes the next job
// from a device counter until the jobs run out (largest order first); every
// workgroup reads the counter once per job and leaves when it passes the job
// count, so the grid drains.  Padded orders from TILE_MIN_PADDED (32: the
// smaller orders keep the scalar per-class launch) to 256; below 256 some
// waves have no pair and only share the tile-blocks, which still beats the
// scalar kernel's per-class launches at every padded order from 32 (measured
// on MI355X over the batched uneven workload: its Jacobi phase 28.6 ms with
// the persistent family from 144, 25.3 ms from 32).
__global__ __launch_bounds__(kTileThreads) void tile_jacobi_kernel(
    const JacobiJob* __restrict__ jobs, uint32_t count, uint32_t* __restrict__ counter,
    const double* __restrict__ a64, float* __restrict__ vectors, float* __restrict__ diagonal,
    float* __restrict__ roots, JacobiStatus* __restrict__ status) {
  __shared__ __align__(16) unsigned char smem[tile_lds_bytes()];
  __shared__ uint32_t next;
  auto* lds = (__attribute__((address_space(3))) unsigned char*)smem;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
    __syncthreads();
    // The job's fields are read from device memory where they are used, at a
    // wave-uniform address (scalar loads, no private copy).
    const uint32_t i = __builtin_amdgcn_readfirstlane(next);
    __syncthreads();
    if (i >= count) return;
    const JacobiJob* job = jobs + i;
    switch ((job->order + 15u) & ~15u) {
#if TILE_MIN_PADDED <= 32
      case 32: tile_job<32>(job, a64, vectors, diagonal, roots, status, lds); break;
#endif
#if TILE_MIN_PADDED <= 48
      case 48: tile_job<48>(job, a64, vectors, diagonal, roots, status, lds); break;
#endif
#if TILE_MIN_PADDED <= 64
      case 64: tile_job<64>(job, a64, vectors, diagonal, roots, status, lds); break;
#endif
#if TILE_MIN_PADDED <= 80
      case 80: tile_job<80>(job, a64, vectors, diagonal, roots, status, lds); break;
#endif
#if TILE_MIN_PADDED <= 96
      case 96: tile_job<96>(job, a64, vectors, diagonal, roots, status, lds); break;
#endif
#if TILE_MIN_PADDED <= 112
      case 112: tile_job<112>(job, a64, vectors, diagonal, roots, status, lds); break;
#endif
#if TILE_MIN_PADDED <= 128
      case 128: tile_job<128>(job, a64, vectors, diagonal, roots, status, lds); break;
#endif
      case 144: tile_job<144>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 160: tile_job<160>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 176: tile_job<176>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 192: tile_job<192>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 208: tile_job<208>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 224: tile_job<224>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 240: tile_job<240>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 256: tile_job<256>(job, a64, vectors, diagonal, roots, status, lds); break;
      default: break;  // the host launches the orders above tile_jacobi_min_padded() - 16 only
    }
    __syncthreads();
  }
}

}  // namespace

uint32_t tile_jacobi_min_padded() { return TILE_MIN_PADDED; }

hipError_t launch_tile_jacobi(const JacobiJob* jobs, uint32_t count, uint32_t* counter, int grid,
                              const double* a64, float* vectors, float* diagonal, float* roots,
                              JacobiStatus* status, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (!jobs || !counter || grid < 1) return hipErrorInvalidValue;
  const uint32_t blocks = static_cast<uint32_t>(grid) < count ? static_cast<uint32_t>(grid) : count;
  hipError_t e = launch_fill_u32(counter, 1, 0u, stream);
  if (e != hipSuccess) return e;
  tile_jacobi_kernel<<<blocks, kTileThreads, 0, stream>>>(jobs, count, counter, a64, vectors,
                                                          diagonal, roots, status);
  return hipGetLastError();
}

}  // namespace msq
