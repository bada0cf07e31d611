// Parallel cyclic Jacobi eigensolver, round-robin disjoint pairs.
//
// Reference: makeJacobiSchedule (src/metal_solver.cpp:50-79), symmetricJacobi
// (src/metal_runtime.cpp:584-646), jacobi_rotations_f32 and
// jacobi_similarity_vectors_f32 (metal_mixed_sqrt.metal:1331-1411).
//
// The reference dispatches two kernels per round, ping-ponging A and V through
// device memory (16 (n_pad - 1) dispatches per sweep pair).  Here, for orders
// up to 256, batched_jacobi_kernel runs one workgroup per matrix: A lives in
// LDS as its packed upper triangle (n(n+1)/2 floats, 129 KiB at n = 256) and
// is rotated in place -- the blocks {p_k, q_k} x {p_l, q_l} of one round are
// disjoint, so each block is read, rotated on both sides and written back by
// one thread with no second buffer.  V lives in registers up to n = 128 and in
// device memory above.  The whole solve -- all sweeps, the FP64 off-norm test
// after each sweep and the square roots -- is one launch per order class of
// the batch.  Larger orders use the blocked solver (kernels_blocked.hip),
// whose 64 x 64 subproblems run on this kernel.
//
// The schedule is the reference's circle method in closed form: in round t the
// participant at slot x is P_t(0) = 0, P_t(x) = 1 + ((x - 1 - t) mod (N - 1)),
// and slot k pairs with slot N - 1 - k.  N is the order rounded up to 16 (the
// reference rounds to even, N = n + n % 2): the extra indices are zero rows
// and columns, whose pairs have a_pq = 0 and rotate by the identity, so the
// rotations of the matrix's own pairs are the reference's in another order
// (for n a multiple of 16, the same schedule).
//
// Rotation angles: tau = (a_qq - a_pp) / (2 a_pq), t = sign(tau) / (|tau| +
// sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c, with p < q, evaluated in
// binary64 (an equivalent form, rotation_angles below) and rounded once to
// float (the reference evaluates them in float);
// entries are updated with the reference's element formula
//   A'[r][c] = rr.x cc.x A[r][c] + rr.y cc.x A[r'][c] + rr.x cc.y A[r][c'] +
//              rr.y cc.y A[r'][c']
// where index p sees (c, -s) and index q sees (c, s).
#include <hip/hip_runtime.h>

#include <atomic>

#include <math.h>

#include <type_traits>

#include "jacobi_math.hpp"
#include "kernels.hpp"

namespace msq {
namespace {

using jmath::Block;
using jmath::participant;
using jmath::rotate_block;
using jmath::rotate_rows_cols;
using jmath::rotation_angles;

__device__ __forceinline__ int tri(int i, int j) {  // i <= j
  return (j * (j + 1)) / 2 + i;
}
__device__ __forceinline__ int tri_sym(int i, int j) {
  return i <= j ? tri(i, j) : tri(j, i);
}
__device__ __forceinline__ int tri_col(int e) {
  int j = static_cast<int>((sqrtf(8.0f * static_cast<float>(e) + 1.0f) - 1.0f) * 0.5f);
  while (j > 0 && (j * (j + 1)) / 2 > e) --j;
  while (((j + 1) * (j + 2)) / 2 <= e) ++j;
  return j;
}

// Launch shape for a padded order N (a multiple of 16 up to 256).
//
// V rows are split into H parts of S slots on adjacent lanes; a lane holds
// the 2 S positions of its slots ("lo" = k, "hi" = N - 1 - k).
//  * N <= 128: S = 8, H = N / 16.  V stays in registers for the whole solve
//    (one row per H lanes, all rows at once).
//  * N > 128: S = N / 16 (9..16), H = 8.  V lives in device memory between
//    groups of rounds; a group is applied with rows r and r + N / 2 on one
//    lane (4 S V values per lane, N / 2 x 8 lanes).
__host__ __device__ constexpr int slots_for_order(int N) { return N <= 128 ? 8 : N / 16; }
__host__ __device__ constexpr bool v_in_memory(int S) { return S != 8; }
__host__ __device__ constexpr int parts_for_order(int N) { return N <= 128 ? N / 16 : 8; }
// History row stride per part: odd, so the H parts' broadcast reads of one
// slot fall in distinct LDS banks.
__host__ __device__ constexpr int padded_slots(int S) { return S | 1; }
// LDS of the kernel but for the rotation history: packed A, slot records
// (two rounds of m 16-byte records), reduction scratch.
__host__ __device__ constexpr int lds_without_history(int N) {
  return (((N * (N + 1) / 2) + 3) & ~3) * 4 + 2 * (N / 2) * 16 + 32 * 8 + 16 * 4 + 4 * 4;
}
constexpr int kLdsPerCu = 160 * 1024;
// V in device memory is updated in chunks of rounds, each unrolled (the
// values shift through fixed registers): 16 rounds, or 8 at N = 256, the one
// order whose A leaves LDS for a single chunk of 16 but three of 8.
__host__ __device__ constexpr int v_chunk_for_order(int N) { return N == 256 ? 8 : 16; }
// Rounds buffered before V is updated.
//  * Register-resident V: a whole phase cycle, S rounds.
//  * V in device memory: V is loaded and stored once per group, so the group
//    is as long as the LDS left by A holds history for (every such order runs
//    one workgroup per CU, bound by its registers), in whole chunks, and at
//    most a sweep.  Groups end at sweep boundaries, so a group is a multiple
//    of the chunk or, the last of a sweep, one round short of one (a sweep is
//    N - 1 = 15 mod 16 rounds).  N = 256: 24; 240: 32; 224: 48; 208: 80;
//    192: 96; 176: 128; 160: 144; 144: 144.
__host__ __device__ constexpr int group_for_order(int N) {
  const int S = slots_for_order(N);
  if (!v_in_memory(S)) return S;
  const int C = v_chunk_for_order(N);
  const int chunk = C * 8 * padded_slots(S) * 8;  // bytes of a chunk of history
  int l = (kLdsPerCu - 512 - lds_without_history(N)) / chunk;  // 512: alignment slack
  const int sweep_chunks = (N - 1 + C - 1) / C;
  l = l < sweep_chunks ? l : sweep_chunks;
  return C * (l < 1 ? 1 : l);
}
// Lanes of the register-resident V: one row per H = N / 16 lanes, whole waves.
__host__ __device__ constexpr int v_threads_for_order(int N) {
  return ((N + (64 / (N / 16)) - 1) / (64 / (N / 16))) * 64;
}
// Threads per workgroup.  N > 128: 4 N.  N <= 128: the V lanes, or (wide)
// one thread per off-diagonal block where that fits 1024 lanes and exceeds
// the V lanes (N = 32 .. 80), so that a round's block phase is a single
// block per thread; the waves past the V lanes skip the V updates.  A
// launch whose workgroups are all resident at once is bound by the latency
// of a round (the pair phase runs after its thread's block), which the wide
// shape shortens (one 64 x 64 sweep, as in the subproblems of the blocked
// Jacobi: 75 -> 55 us); a launch with more workgroups than that is bound by
// throughput, where the narrow shape's fuller CUs win (measured on MI355X:
// 20,000 mixed orders up to 512 ran 13-19% slower all wide).
__host__ __device__ constexpr int threads_for_order(int N, bool wide = false) {
  if (N > 128) return 4 * N;
  const int v = v_threads_for_order(N);
  const int blocks = ((N / 2) * (N / 2 - 1) / 2 + 63) / 64 * 64;
  return wide && blocks <= 1024 && blocks > v ? blocks : v;
}
__host__ __device__ constexpr bool has_wide_shape(int N) {
  return threads_for_order(N, true) != threads_for_order(N, false);
}

// Off-diagonal 2x2 blocks per thread when each thread takes single blocks:
// the m (m - 1) / 2 blocks over the launch's threads, m = N / 2, the
// next-pair blocks among them (one per thread below m).
__host__ __device__ constexpr int single_blocks_per_thread(int N, bool wide = false) {
  return ((N / 2) * (N / 2 - 1) / 2 + threads_for_order(N, wide) - 1) /
         threads_for_order(N, wide);
}
// Orders whose block phase is bound by LDS throughput rather than by the
// latency of a round take super-blocks (below); smaller orders take single
// blocks, which keep the per-wave critical path shortest.  Measured on
// MI355X (one launch per padded order of a 4096-matrix batch): N = 256 runs
// 5.5% faster with super-blocks, N = 240 4% slower, N <= 128 about 2x slower.
__host__ __device__ constexpr bool use_super_blocks(int N) { return N >= 256; }
// V in device memory: two rows per lane (rows r and r + N / 2; each history
// read serves both), 4 S values per lane, which spill a little at
// N = 192 .. 240 (36 to 84 bytes per lane).  Measured on MI355X (256
// matrices per order, 4 fixed sweeps) with V groups as long as the LDS
// allows, two rows run 2-3% faster than one row per lane in two passes at
// N = 144 .. 176 and 224 .. 240, 0.3-1% at 192, 208 and 256.

// Large orders handle off-diagonal blocks as 2 x 2 super-blocks of slot pairs:
// super-block (K, L), K <= L < m / 2, covers slots {2K, 2K + 1} x {2L, 2L + 1}
// (four blocks for K < L; the single block (2K, 2K + 1) for K = L), so one
// thread reads four slot records for four blocks.  The super-diagonal ones
// (K, K + 1) and the two corner diagonal ones hold the blocks that complete
// the next round's pairs and belong to the pair-phase owners; the others
// ("items": K + 2 <= L, and K = L for 0 < K < m / 2 - 1) are dealt to all
// threads.
__host__ __device__ constexpr int super_items(int N) {
  return (N / 4 - 1) * (N / 4 - 2) / 2 + (N / 4 - 2);
}
// Sums of one or two doubles per thread in a fixed order: the values of
// threads 0 .. nt - 1 in part0 / part1 (nt entries each; threads beyond nt
// take part in the barriers only), then a tree over thread indices (stride
// halving from the power of two below nt).  The result does not depend on
// which waves arrive first, so a matrix's norms and stop decisions are the
// same in every launch and whatever else shares the batch.  Every thread
// returns the sums; part0 / part1 are free again on return.
__device__ __forceinline__ void block_sum2(double& a, double& b, double* part0, double* part1,
                                           int tid, int nt) {
  if (tid < nt) {
    part0[tid] = a;
    part1[tid] = b;
  }
  __syncthreads();
  int s = 1;
  while (2 * s < nt) s *= 2;
  for (; s > 0; s >>= 1) {
    if (tid < s && tid + s < nt) {
      part0[tid] += part0[tid + s];
      part1[tid] += part1[tid + s];
    }
    __syncthreads();
  }
  a = part0[0];
  b = part1[0];
  __syncthreads();
}

__host__ __device__ constexpr int items_per_thread(int N) {
  return super_items(N) / threads_for_order(N) + 1;
}

// Lane shifts inside a V row (8 adjacent lanes, inside one 16-lane DPP row):
// value of lane - 1 (row_shr:1) and of lane + 1 (row_shl:1).  Lanes whose
// source lies outside their DPP row get 0 and never use it.
__device__ __forceinline__ float from_lane_below(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x111, 0xf, 0xf, false));
}
__device__ __forceinline__ float from_lane_above(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x101, 0xf, 0xf, false));
}

// One round of V updates at phase PHI = round mod S (register-resident V).
// Register r holds the lo value of slot (r + PHI) mod S of this part and hi[r]
// the hi value of slot (r - PHI) mod S: values stay in their registers while
// the round-robin moves positions, and only the mapping advances with the
// phase.  After the rotations one lo value leaves for the next part (one hi
// value for the previous part) and the registers they free take the arriving
// values.
template <int S, int PHI>
__device__ __forceinline__ void v_round(float (&lo)[S], float (&hi)[S],
                                        const float2* __restrict__ rotv, int h, int H,
                                        int lane) {
#pragma unroll
  for (int r0 = 0; r0 < S; r0 += 4) {
    asm volatile("" ::: "memory");
    float2 r4[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (r0 + u < S) r4[u] = rotv[h * padded_slots(S) + (r0 + u + PHI) % S];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int rl = r0 + u;
      if (rl < S) {
        const int rh = (rl + 2 * PHI) % S;
        const float x = lo[rl], y = hi[rh];
        lo[rl] = r4[u].x * x + r4[u].y * y;
        hi[rh] = r4[u].x * y - r4[u].y * x;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  constexpr int r_last_lo = ((S - 1 - PHI) % S + S) % S;  // slot S-1's lo
  constexpr int r_first_lo = ((0 - PHI) % S + S) % S;     // slot 0's lo
  constexpr int r_first_hi = PHI % S;                     // slot 0's hi
  const float out_up = lo[r_last_lo];
  const float out_dn = hi[r_first_hi];
  const float pos0 = lo[r_first_lo];
  const float in_up = __shfl(out_up, lane > 0 ? lane - 1 : 0);
  const float in_dn = __shfl(out_dn, lane < 63 ? lane + 1 : 63);
  // Selects, not branches (h differs between lanes): part 0 keeps position 0
  // (slot 0's lo next round) and moves position N - 1 to 1; the last part
  // turns its last lo into its last hi.
  const bool first = h == 0;
  lo[r_last_lo] = first ? pos0 : in_up;
  lo[r_first_lo] = first ? out_dn : lo[r_first_lo];
  hi[r_first_hi] = h == H - 1 ? out_up : in_dn;
  // Round boundary: pin every V value to its register (empty asm, no
  // instruction), so the next round's arithmetic cannot overlap this one's
  // and the live set stays 2 S values plus one round's temporaries.
#pragma unroll
  for (int r = 0; r < S; ++r) asm volatile("" : "+v"(lo[r]), "+v"(hi[r]));
}

// Applies G buffered rounds to V, starting at phase P0; each round's
// (c, s') per slot is one row of `hist` (`row` entries apart).  Straight-line
// code: rounds a group did not reach are identity rotations in `hist`.
template <int S, int G, int P0, int I = 0>
__device__ __forceinline__ void v_group(float (&lo)[S], float (&hi)[S],
                                        const float2* __restrict__ hist, int row, int h, int H,
                                        int lane) {
  if constexpr (I < G) {
    v_round<S, (P0 + I) % S>(lo, hi, hist + I * row, h, H, lane);
    v_group<S, G, P0, I + 1>(lo, hi, hist, row, h, H, lane);
  }
}

// One round on two V rows held slot-aligned (8 parts per row, the same
// part of rows r and r + N / 2): lo[j] / hi[j] are slot j's two positions,
// rotated together, then shifted one position along the round-robin circle
// (straight-line code turns the shifts into register renames; the caller
// stores every register back to its own position).  Each rotation is read
// from LDS once and applied to both rows.
template <int S>
__device__ __forceinline__ void v_round_shift2(float (&lo0)[S], float (&hi0)[S],
                                               float (&lo1)[S], float (&hi1)[S],
                                               const float2* __restrict__ rotv, int h) {
#pragma unroll
  for (int j0 = 0; j0 < S; j0 += 4) {
    float2 r4[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (j0 + u < S) r4[u] = rotv[h * padded_slots(S) + j0 + u];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int j = j0 + u;
      if (j < S) {
        const float x0 = lo0[j], y0 = hi0[j];
        lo0[j] = r4[u].x * x0 + r4[u].y * y0;
        hi0[j] = r4[u].x * y0 - r4[u].y * x0;
        const float x1 = lo1[j], y1 = hi1[j];
        lo1[j] = r4[u].x * x1 + r4[u].y * y1;
        hi1[j] = r4[u].x * y1 - r4[u].y * x1;
      }
    }
  }
  // The shift: lo[S - 1] goes to the next part's lo[0] (the last part's to
  // its own hi[S - 1]), hi[0] to the previous part's hi[S - 1] (part 0's to
  // its own lo[1]); position 0 stays.
  const bool first = h == 0, last = h == 7;
  {
    const float out_up = lo0[S - 1], out_dn = hi0[0], keep0 = lo0[0];
    const float in_up = from_lane_below(out_up);
    const float in_dn = from_lane_above(out_dn);
#pragma unroll
    for (int j = S - 1; j >= 2; --j) lo0[j] = lo0[j - 1];
    lo0[1] = first ? out_dn : lo0[0];
    lo0[0] = first ? keep0 : in_up;
#pragma unroll
    for (int j = 0; j < S - 1; ++j) hi0[j] = hi0[j + 1];
    hi0[S - 1] = last ? out_up : in_dn;
  }
  {
    const float out_up = lo1[S - 1], out_dn = hi1[0], keep0 = lo1[0];
    const float in_up = from_lane_below(out_up);
    const float in_dn = from_lane_above(out_dn);
#pragma unroll
    for (int j = S - 1; j >= 2; --j) lo1[j] = lo1[j - 1];
    lo1[1] = first ? out_dn : lo1[0];
    lo1[0] = first ? keep0 : in_up;
#pragma unroll
    for (int j = 0; j < S - 1; ++j) hi1[j] = hi1[j + 1];
    hi1[S - 1] = last ? out_up : in_dn;
  }
  // Round boundary (empty asm, no instruction): V values pinned to registers
  // and memory accesses kept on their side, so rounds do not overlap.
#pragma unroll
  for (int r = 0; r < S; ++r) {
    asm volatile("" : "+v"(lo0[r]), "+v"(hi0[r]));
    asm volatile("" : "+v"(lo1[r]), "+v"(hi1[r]));
  }
  asm volatile("" ::: "memory");
}

template <int S, int G, int I = 0>
__device__ __forceinline__ void v_group_shift2(float (&lo0)[S], float (&hi0)[S],
                                               float (&lo1)[S], float (&hi1)[S],
                                               const float2* __restrict__ hist, int row, int h) {
  if constexpr (I < G) {
    v_round_shift2<S>(lo0, hi0, lo1, hi1, hist + I * row, h);
    v_group_shift2<S, G, I + 1>(lo0, hi0, lo1, hi1, hist, row, h);
  }
}

// Byte offset of the packed-triangle entry (x, y), x != y, from the
// byte-scaled participants and column bases (cb(x) = x (x + 1) / 2): cb is
// increasing, so the larger participant has the larger base.
__device__ __forceinline__ int packed_offset(int x4, int cb4x, int y4, int cb4y) {
  return max(cb4x, cb4y) + min(x4, y4);
}
__device__ __forceinline__ float& lds_at(float* base, int byte_offset) {
  return *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + byte_offset);
}

// Per slot and round, one 16-byte record: the rotation (c, s) of the pair
// p < q and, for each of p and q, the byte-scaled participant x4 = 4 x in the
// low 10 bits and the byte-scaled column base cb4 = 4 x (x + 1) / 2 above
// them (x <= 255: 10 + 17 bits).  One ds_read_b128 per slot serves a block.
struct __align__(16) SlotRec {
  float c, s;
  uint32_t p, q;  // x4 | cb4 << 10
};
__device__ __forceinline__ uint32_t slot_index(int x) {
  return static_cast<uint32_t>(4 * x) | (static_cast<uint32_t>(2 * x * (x + 1)) << 10);
}
__device__ __forceinline__ int idx4(uint32_t w) { return static_cast<int>(w & 0x3ffu); }
// Record position of slot k.  Super-blocks read slots 2L and 2L + 1 for
// consecutive L on consecutive lanes, so their records are stored even slots
// first, then odd ones: each of those reads is then 16-byte consecutive
// across lanes (a stride of two records would put four lanes of every
// 16-lane group on the same LDS banks).  Single blocks read slot l for
// consecutive l: identity.
template <bool SPLIT, int M>
__device__ __forceinline__ int slot_pos(int k) {
  return SPLIT ? (k & 1) * (M / 2) + (k >> 1) : k;
}
__device__ __forceinline__ int base4(uint32_t w) { return static_cast<int>(w >> 10); }

// Rotates U off-diagonal blocks (k < l) in place.  All loads are issued before
// any store (the blocks are disjoint, which the compiler cannot know); entries
// of `code` equal to ~0 are padding: they load a real block and store nothing.
template <int U>
__device__ __forceinline__ void rotate_blocks(float* __restrict__ Ap,
                                              const SlotRec* __restrict__ rec,
                                              const uint32_t* code) {
  int i00[U], i01[U], i10[U], i11[U];
  Block blk[U];
  float ck[U], sk[U], cl[U], sl[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const uint32_t c = code[u] == ~0u ? (1u << 16) : code[u];
    const int k = static_cast<int>(c & 0xffffu), l = static_cast<int>(c >> 16);
    const SlotRec rk = rec[k], rl = rec[l];
    i00[u] = packed_offset(idx4(rk.p), base4(rk.p), idx4(rl.p), base4(rl.p));
    i01[u] = packed_offset(idx4(rk.p), base4(rk.p), idx4(rl.q), base4(rl.q));
    i10[u] = packed_offset(idx4(rk.q), base4(rk.q), idx4(rl.p), base4(rl.p));
    i11[u] = packed_offset(idx4(rk.q), base4(rk.q), idx4(rl.q), base4(rl.q));
    blk[u].a00 = lds_at(Ap, i00[u]);
    blk[u].a01 = lds_at(Ap, i01[u]);
    blk[u].a10 = lds_at(Ap, i10[u]);
    blk[u].a11 = lds_at(Ap, i11[u]);
    ck[u] = rk.c;
    sk[u] = rk.s;
    cl[u] = rl.c;
    sl[u] = rl.s;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const Block b = blk[u];
    Block o;
    rotate_rows_cols(b, ck[u], sk[u], cl[u], sl[u], o);
    const float o00 = o.a00, o01 = o.a01, o10 = o.a10, o11 = o.a11;
    if (code[u] != ~0u) {
      lds_at(Ap, i00[u]) = o00;
      lds_at(Ap, i01[u]) = o01;
      lds_at(Ap, i10[u]) = o10;
      lds_at(Ap, i11[u]) = o11;
    }
  }
}

// Rotates the blocks of super-block `code` = K | L << 16 in place: all four
// for K < L, only (2K, 2K + 1) for K = L (the other three load and are
// dropped).  All loads are issued before any store (the blocks of a round are
// disjoint, which the compiler cannot know).
// rec: records in the split order of slot_pos (even slots, then odd ones,
// m / 2 = M2 each).
template <int M2>
__device__ __forceinline__ void rotate_super(float* __restrict__ Ap,
                                             const SlotRec* __restrict__ rec, uint32_t code) {
  const int K = static_cast<int>(code & 0xffffu), L = static_cast<int>(code >> 16);
  const SlotRec rk[2] = {rec[K], rec[M2 + K]};
  const SlotRec rl[2] = {rec[L], rec[M2 + L]};
  int i00[4], i01[4], i10[4], i11[4];
  Block blk[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const SlotRec& x = rk[u >> 1];
    const SlotRec& y = rl[u & 1];
    i00[u] = packed_offset(idx4(x.p), base4(x.p), idx4(y.p), base4(y.p));
    i01[u] = packed_offset(idx4(x.p), base4(x.p), idx4(y.q), base4(y.q));
    i10[u] = packed_offset(idx4(x.q), base4(x.q), idx4(y.p), base4(y.p));
    i11[u] = packed_offset(idx4(x.q), base4(x.q), idx4(y.q), base4(y.q));
    blk[u].a00 = lds_at(Ap, i00[u]);
    blk[u].a01 = lds_at(Ap, i01[u]);
    blk[u].a10 = lds_at(Ap, i10[u]);
    blk[u].a11 = lds_at(Ap, i11[u]);
  }
  const bool diag = K == L;
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const SlotRec& x = rk[u >> 1];
    const SlotRec& y = rl[u & 1];
    Block o;
    rotate_rows_cols(blk[u], x.c, x.s, y.c, y.s, o);
    const float o00 = o.a00, o01 = o.a01, o10 = o.a10, o11 = o.a11;
    if (!diag || u == 1) {
      lds_at(Ap, i00[u]) = o00;
      lds_at(Ap, i01[u]) = o01;
      lds_at(Ap, i10[u]) = o10;
      lds_at(Ap, i11[u]) = o11;
    }
  }
}

// Blocks whose rotated entries complete the next round's pairs.  Round t + 1
// pairs, at slot k', the participants that round t holds at positions k' - 1
// and N - 2 - k' (the round-robin moves every position x >= 1 to x + 1,
// N - 1 to 1); as slots of round t these are
//   k' = 0:           lo(0), hi(1)        -> block (0, 1)
//   1 <= k' <= m - 2: lo/hi(k' - 1), hi(k' + 1) -> block (k' - 1, k' + 1)
//   k' = m - 1:       lo(m - 2), lo(m - 1) -> block (m - 2, m - 1)
// (for k' = 1 position 0 is fixed and position N - 1 arrives: hi(0)).  With
// single blocks, thread k' owns next_pair_block(k').  With super-blocks,
// super-block (j, j + 1) holds the blocks of slots 2j + 1 and 2j + 2, and
// the corner super-blocks (0, 0) and (m/2 - 1, m/2 - 1) those of slots 0 and
// m - 1.  Owner thread j < m/2 - 1 takes super-block (j, j + 1); thread
// m/2 - 1 the corner (0, 0); thread m/2 the corner (m/2 - 1, m/2 - 1).
__device__ __forceinline__ uint32_t next_pair_block(int kp, int m) {
  const int k = kp == 0 ? 0 : (kp == m - 1 ? m - 2 : kp - 1);
  const int l = kp == 0 ? 1 : (kp == m - 1 ? m - 1 : kp + 1);
  return static_cast<uint32_t>(k) | (static_cast<uint32_t>(l) << 16);
}
// Item r of the dealt single blocks, the ones no next_pair_block names:
// (k, k + 1) for 0 < k < m - 2, then (k, l) with k + 3 <= l row-major;
// m (m - 3) / 2 items in all.
__device__ __forceinline__ uint32_t item_single(int r, int m) {
  if (r < m - 3) return static_cast<uint32_t>(r + 1) | (static_cast<uint32_t>(r + 2) << 16);
  r -= m - 3;
  int k = 0;
  while (r >= m - 3 - k) {
    r -= m - 3 - k;
    ++k;
  }
  return static_cast<uint32_t>(k) | (static_cast<uint32_t>(k + 3 + r) << 16);
}
__device__ __forceinline__ uint32_t super_code(int K, int L) {
  return static_cast<uint32_t>(K) | (static_cast<uint32_t>(L) << 16);
}
__device__ __forceinline__ uint32_t owner_super(int tid, int m2) {
  if (tid < m2 - 1) return super_code(tid, tid + 1);
  if (tid == m2 - 1) return super_code(0, 0);
  if (tid == m2) return super_code(m2 - 1, m2 - 1);
  return ~0u;
}
// Item r of the dealt super-blocks: (K, L) with K + 2 <= L row-major, then
// the diagonal ones (K, K), 0 < K < m/2 - 1.
__device__ __forceinline__ uint32_t item_super(int r, int m2) {
  const int off = (m2 - 1) * (m2 - 2) / 2;
  if (r >= off) return super_code(1 + r - off, 1 + r - off);
  int K = 0;
  while (r >= m2 - 2 - K) {
    r -= m2 - 2 - K;
    ++K;
  }
  return super_code(K, K + 2 + r);
}

// One workgroup per matrix, orders up to 256.
//
// A: packed upper triangle in LDS, indexed by matrix index, rotated in place.
// A round is two phases.  Pair phase: one thread per slot k computes the
// rotation of its pair from a_pp, a_qq, a_pq and rotates the diagonal block
// {p, q} x {p, q} (no other block touches it).  Block phase: every
// off-diagonal block {p_k, q_k} x {p_l, q_l} (k < l) is read, rotated on both
// sides and written back by one thread; the blocks of a round are disjoint,
// so there is no second buffer.  A thread works on 2 x 2 super-blocks of slot
// pairs, fixed for the whole solve.  Inside a group of rounds the pair phase
// of round t + 1 runs in the block phase of round t: the owner of the
// super-block that completes round t + 1's pairs at slots 2j + 1 and 2j + 2
// (owner_super) rotates it first, then computes those two rotations and
// rotates their diagonal blocks, which no other thread of round t touches.
// Slot data is double-buffered by round parity, so a round costs one
// barrier.
//
// V: slot-aligned rows.  Thread (row i, part h) holds V[i][x] for the
// positions x of slots k in [h S, (h + 1) S): "lo" = position k, "hi" =
// position N - 1 - k.  Slot k pairs positions k and N - 1 - k in every round,
// so each rotation is local to a thread.  Between rounds the round-robin
// moves every position x >= 1 to x + 1 (N - 1 to 1): lo shifts up, hi shifts
// down, and one value crosses to each neighbouring part of the row (adjacent
// lanes).  The rotations of a group of rounds are buffered in LDS and applied
// to V after the group.
// LDS of one solve of padded order NP, carved from one byte array at
// compile-time offsets (addresses fold into the ds instructions' offset
// fields): packed A, slot records (two rounds x m), the rotation history (G
// rounds x 8 parts x SP of (c, s') for V), the fixed-order reduction
// partials (above N = 128 they live in the history, idle at both reduction
// points, after the s_i = sqrt(a_ii) values the convergence test keeps at its
// start; below, where LDS is plentiful, in an array of their own), and
// scratch for the minimum and the stop flag.
template <int NP, bool WIDE>
struct JacobiLds {
  static constexpr int kS = slots_for_order(NP);
  static constexpr int kPacked = NP * (NP + 1) / 2;
  static constexpr int kHistEntries = group_for_order(NP) * 8 * padded_slots(kS);
  static constexpr int kThreads = threads_for_order(NP, WIDE);
  static constexpr bool kPartInHistory = NP > 128;
  static constexpr int kSrootBytes = (4 * NP + 7) / 8 * 8;
  static constexpr size_t kAp = 0;
  static constexpr size_t kRec = kAp + static_cast<size_t>((kPacked + 3) & ~3) * sizeof(float);
  static constexpr size_t kHist = kRec + 2 * static_cast<size_t>(NP / 2) * sizeof(SlotRec);
  static constexpr size_t kPart = kHist + static_cast<size_t>(kHistEntries) * sizeof(float2);
  static constexpr size_t kRedf =
      kPart + (kPartInHistory ? 0 : 2 * static_cast<size_t>(kThreads) * sizeof(double));
  static constexpr size_t kCtrl = kRedf + 16 * sizeof(float);
  static constexpr size_t kBytes = kCtrl + 4 * sizeof(int);
  static_assert(kRec % 16 == 0 && kHist % 16 == 0 && kPart % 8 == 0, "LDS alignment");
  static_assert(!kPartInHistory || sizeof(float2) * static_cast<size_t>(kHistEntries) >=
                                       kSrootBytes + 2 * sizeof(double) * kThreads,
                "room for the reduction partials in the rotation history");
};

// One matrix on one workgroup.  NT = threads_for_order(NP, WIDE) threads do
// the work; a workgroup may have more (the persistent launch runs every order
// on 1024 threads): those take part in the barriers only, every work loop
// strides by NT, and the results do not depend on the workgroup's size.
template <int NP, bool WIDE>
__device__ __forceinline__ void jacobi_solve(const JacobiJob& job, const double* __restrict__ a64,
                                             float* __restrict__ vectors,
                                             float* __restrict__ diagonal,
                                             float* __restrict__ roots,
                                             JacobiStatus* __restrict__ status,
                                             unsigned char* __restrict__ smem, const int nt) {
  using L = JacobiLds<NP, WIDE>;
  constexpr int S = slots_for_order(NP);
  constexpr bool VMEM = v_in_memory(S);
  constexpr int G = group_for_order(NP);
  constexpr int VC = v_chunk_for_order(NP);
  constexpr int SP = padded_slots(S);
  constexpr bool SUPER = use_super_blocks(NP);
  constexpr int IPT = items_per_thread(NP);
  constexpr int BPT = single_blocks_per_thread(NP, WIDE);
  const int tid = threadIdx.x;
  // nt: the working threads, threads_for_order(NP, WIDE), passed in as a
  // run-time value (the work loops' stride: with a compile-time stride the
  // compiler reshapes them and spills in the block phase).
  constexpr int nthreads = threads_for_order(NP, WIDE);
  const bool worker = tid < nt;  // whole waves: nt is a multiple of 64
  const int n = static_cast<int>(job.order);
  const int ld = static_cast<int>(job.ld);
  constexpr int N = NP;  // padded order: zero rows/columns beyond n
  constexpr int m = N / 2;
  constexpr int H = parts_for_order(N);
  // History entries per round: room for 8 parts whatever H is, so every
  // history offset is a compile-time constant from one base address.
  constexpr int hrow = 8 * SP;
  constexpr int rows_per_wave = 64 / H;
  constexpr int packed = N * (N + 1) / 2;
  constexpr int m2 = m / 2;  // slot pairs
  constexpr int items = super_items(NP);

  float* const Ap = reinterpret_cast<float*>(smem + L::kAp);
  SlotRec (*const srec)[m] = reinterpret_cast<SlotRec(*)[m]>(smem + L::kRec);
  float2* const hist = reinterpret_cast<float2*>(smem + L::kHist);
  float* const redf = reinterpret_cast<float*>(smem + L::kRedf);
  int* const ctrl = reinterpret_cast<int*>(smem + L::kCtrl);
  double* const part0 = L::kPartInHistory
                            ? reinterpret_cast<double*>(smem + L::kHist + L::kSrootBytes)
                            : reinterpret_cast<double*>(smem + L::kPart);
  double* const part1 = part0 + nthreads;

  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int h = lane % H;
  // Register-resident V: one row per H lanes; lanes beyond the last whole
  // row of a wave idle in the V work but take part in everything else.
  const int row = wave * rows_per_wave + lane / H;
  const bool vlane = lane / H < rows_per_wave && row < N;

  // This thread's super-blocks, K | L << 16: first its owner super-block
  // (pair-phase owners only), then its dealt items, item r to thread
  // r mod nthreads: the empty entries (~0) gather in the last threads, so
  // whole waves skip them (LDS bandwidth is the block phase's bound).
  //
  // Single blocks (smaller orders): (k | l << 16), BPT per thread, the first
  // of a thread below m the block that completes the next round's pair at
  // slot tid; the other blocks (item_single) fill the remaining entries in
  // order, entry 0 of threads m .. nthreads - 1 first, then entry 1 of every
  // thread, and so on.  Empty entries (~0) load a real block and store
  // nothing, keeping the waves in step.
  constexpr int CODES = SUPER ? IPT + 1 : BPT;
  uint32_t code[CODES];
  if (!worker) {
#pragma unroll
    for (int i = 0; i < CODES; ++i) code[i] = ~0u;
  } else if constexpr (SUPER) {
    code[0] = owner_super(tid, m2);
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int r = tid + i * nthreads;
      code[i + 1] = r < items ? item_super(r, m2) : ~0u;
    }
  } else {
    constexpr int dealt = m * (m - 3) / 2;
    code[0] = tid < m ? next_pair_block(tid, m)
                      : (tid - m < dealt ? item_single(tid - m, m) : ~0u);
#pragma unroll
    for (int i = 1; i < BPT; ++i) {
      // The last entry in reverse thread order: its empty entries fall to the
      // lowest threads, whose waves also run the pair phases, and a wave with
      // every last entry empty skips it.
      const int r = nthreads - m + (i - 1) * nthreads + (i == BPT - 1 ? nthreads - 1 - tid : tid);
      code[i] = r < dealt ? item_single(r, m) : ~0u;
    }
  }

  const double* A = a64 + job.a_offset;
  double norm2 = 0.0;
  #pragma unroll 1
  for (int e = tid; worker && e < packed; e += nt) {
    const int j = tri_col(e);
    const int i = e - (j * (j + 1)) / 2;
    float v = 0.0f;
    if (i < n && j < n) {
      const double a = A[static_cast<size_t>(i) + static_cast<size_t>(j) * ld];
      v = static_cast<float>(a);
      norm2 += (i == j ? 1.0 : 2.0) * a * a;
    }
    Ap[e] = v;
  }
  float* Vg = vectors + job.v_offset;
  // Register-resident V (VMEM false) starts as the identity; the device-memory
  // copy (VMEM true) holds V in position order: column x = position x.
  float lo[VMEM ? 1 : S], hi[VMEM ? 1 : S];
  if constexpr (!VMEM) {
#pragma unroll
    for (int j = 0; j < S; ++j) {
      const int k = h * S + j;
      lo[j] = (vlane && row == k) ? 1.0f : 0.0f;
      hi[j] = (vlane && row == N - 1 - k) ? 1.0f : 0.0f;
    }
  } else {
    #pragma unroll 1
    for (int e = tid; worker && e < N * N; e += nt) {
      const int i = e % N;
      const int x = e / N;
      Vg[i + static_cast<size_t>(x) * ld] = i == x ? 1.0f : 0.0f;
    }
  }
  // Block reductions go through LDS (wave shuffles would keep their
  // lane-address registers live through the whole kernel), in a fixed order.
  double unused = 0.0;
  block_sum2(norm2, unused, part0, part1, tid, nt);
  const double matrix_norm = sqrt(norm2);

  const uint32_t fixed = job.fixed_sweeps;
  const uint32_t cap = fixed ? fixed : job.max_sweeps;
  const int rounds_per_sweep = N - 1;
  const int limit = static_cast<int>(cap) * rounds_per_sweep;
  double previous = INFINITY;
  double off = 0.0, rate = INFINITY;
  float min_diag = 0.0f;
  uint32_t sweeps = 0;
  int t = 0;   // rounds applied to A
  int tv = 0;  // rounds applied to V (register V: whole groups, t rounded up to G)
  JacobiStatus* st = status + job.status_index;

  // One group: up to G rounds on A (rotation parameters buffered), the G
  // buffered rounds on V, then the convergence test if a sweep finished.
  // Returns true when the solve stops.
  // Pair phase of absolute round tt at slot k: rotation of the pair, the
  // rotated diagonal block, slot data into buffer tt & 1, V's (c, s') into
  // history row rv.
  auto pair_phase = [&](int k, int tt, float2* rv) {
    const int round = tt % rounds_per_sweep;
    const int a = participant(round, k, N);
    const int b = participant(round, N - 1 - k, N);
    const int p = a < b ? a : b;
    const int q = a < b ? b : a;
    const int cbp = (p * (p + 1)) >> 1;
    const int cbq = (q * (q + 1)) >> 1;
    const int ipp = cbp + p, ipq = cbq + p, iqq = cbq + q;
    Block blk;
    blk.a00 = Ap[ipp];
    blk.a01 = Ap[ipq];
    blk.a10 = blk.a01;
    blk.a11 = Ap[iqq];
    const float2 r = rotation_angles(blk.a00, blk.a11, blk.a01);
    const float2 rp = make_float2(r.x, -r.y);
    const Block o = rotate_block(blk, rp, r, rp, r);
    Ap[ipp] = o.a00;
    Ap[ipq] = o.a01;
    Ap[iqq] = o.a11;
    SlotRec rec;
    rec.c = r.x;
    rec.s = r.y;
    rec.p = slot_index(p);
    rec.q = slot_index(q);
    srec[tt & 1][slot_pos<SUPER, m>(k)] = rec;
    // position k holds a: it is p (sees (c, -s)) when a < b
    rv[(k / S) * SP + k % S] = make_float2(r.x, a < b ? -r.y : r.y);
  };
  // Pair phases of two slots with every load before any store, so that the
  // two binary64 angle computations overlap (the pairs are disjoint).
  auto pair_phase2 = [&](int k0, int k1, int tt, float2* rv) {
    const int round = tt % rounds_per_sweep;
    const int ks[2] = {k0, k1};
    int a[2], b[2], p[2], q[2], cbp[2], cbq[2];
    Block blk[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      a[u] = participant(round, ks[u], N);
      b[u] = participant(round, N - 1 - ks[u], N);
      p[u] = a[u] < b[u] ? a[u] : b[u];
      q[u] = a[u] < b[u] ? b[u] : a[u];
      cbp[u] = (p[u] * (p[u] + 1)) >> 1;
      cbq[u] = (q[u] * (q[u] + 1)) >> 1;
      blk[u].a00 = Ap[cbp[u] + p[u]];
      blk[u].a01 = Ap[cbq[u] + p[u]];
      blk[u].a10 = blk[u].a01;
      blk[u].a11 = Ap[cbq[u] + q[u]];
    }
    float2 r[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) r[u] = rotation_angles(blk[u].a00, blk[u].a11, blk[u].a01);
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const float2 rp = make_float2(r[u].x, -r[u].y);
      const Block o = rotate_block(blk[u], rp, r[u], rp, r[u]);
      Ap[cbp[u] + p[u]] = o.a00;
      Ap[cbq[u] + p[u]] = o.a01;
      Ap[cbq[u] + q[u]] = o.a11;
      SlotRec rec;
      rec.c = r[u].x;
      rec.s = r[u].y;
      rec.p = slot_index(p[u]);
      rec.q = slot_index(q[u]);
      srec[tt & 1][slot_pos<SUPER, m>(ks[u])] = rec;
      rv[(ks[u] / S) * SP + ks[u] % S] = make_float2(r[u].x, a[u] < b[u] ? -r[u].y : r[u].y);
    }
  };

  auto group = [&]() -> bool {
    const int group_start = t;
    // V in device memory: groups end at sweep boundaries (the stop rules are
    // checked there, and V applies exactly the group's rounds).
    const int sweep_end = (t / rounds_per_sweep + 1) * rounds_per_sweep;
    const int group_end = VMEM ? min(min(t + G, limit), sweep_end) : min(t + G, limit);
    for (int k = tid; k < m; k += nt) pair_phase(k, t, hist);
    __syncthreads();
#pragma unroll 1
    for (; t < group_end; ++t) {
      const SlotRec* cur = srec[t & 1];
      if constexpr (SUPER) {
        // Owner super-blocks first (waves without owners skip them), then the
        // pair phases of round t + 1, then the dealt items.
        if (code[0] != ~0u) rotate_super<m2>(Ap, cur, code[0]);
        if (t + 1 < group_end) {
          float2* rv = hist + (t + 1 - group_start) * hrow;
          if (tid < m2 - 1) {
            pair_phase2(2 * tid + 1, 2 * tid + 2, t + 1, rv);
          } else if (tid == m2 - 1) {
            pair_phase(0, t + 1, rv);
          } else if (tid == m2) {
            pair_phase(m - 1, t + 1, rv);
          }
        }
#pragma unroll
        for (int i = 1; i <= IPT; ++i)
          if (code[i] != ~0u) rotate_super<m2>(Ap, cur, code[i]);
      } else {
        auto blocks = [&](auto entries) {
          constexpr int NE = decltype(entries)::value;
#pragma unroll
          for (int i = 0; i < NE; i += 2) {
            if (i + 1 < NE) {
              rotate_blocks<2>(Ap, cur, code + i);
            } else {
              rotate_blocks<1>(Ap, cur, code + i);
            }
          }
        };
        if (worker) {
          if constexpr (BPT > 1) {
            if (__ballot(code[BPT - 1] != ~0u) == 0) {
              blocks(std::integral_constant<int, BPT - 1>{});
            } else {
              blocks(std::integral_constant<int, BPT>{});
            }
          } else {
            blocks(std::integral_constant<int, BPT>{});
          }
        }
        if (tid < m && t + 1 < group_end)
          pair_phase(tid, t + 1, hist + (t + 1 - group_start) * hrow);
      }
      __syncthreads();
    }
    // Register V: rounds the group did not reach take identity rotations (V's
    // positions still advance; the output maps them through V's own round
    // count).  V in device memory applies the group's rounds only.
    if constexpr (!VMEM) {
      #pragma unroll 1
      for (int e = (group_end - group_start) * hrow + tid; worker && e < G * hrow; e += nt)
        hist[e] = make_float2(1.0f, 0.0f);
    }
    __syncthreads();
    // Rounds of this group for V in device memory: whole chunks, then one
    // chunk less a round at the end of a sweep.
    const int vchunks = (group_end - group_start) / VC;
    const bool vtail = (group_end - group_start) % VC != 0;
    if constexpr (!VMEM) {
      // Every lane of the V waves runs the V code (lanes past the last whole
      // row of a wave compute on zeros that no valid lane reads): no divergent
      // branch around the 2 S live V registers.  Waves past the V lanes skip
      // it (a wave-uniform branch).  G = S: every group starts at phase 0.
      if (worker && wave * rows_per_wave < N) v_group<S, G, 0>(lo, hi, hist, hrow, h, H, lane);
    } else {
      // Each lane loads its positions, applies the group's rounds with the
      // values shifting along the circle in registers, and stores every
      // register back to its own position: no other lane touches them, so no
      // barrier in between.  Element offsets from a value the compiler must
      // treat as new in every group (empty asm): otherwise it hoists all the
      // addresses out of the loop and keeps them live (spilled) through the A
      // update.
      if (worker) {
        // Rows r and r + N / 2 on the same lane (8 lanes per row), so each
        // rotation read from the history serves two rows.
        const int vrow = tid >> 3;
        const int vh = tid & 7;
        uint32_t lo_off = static_cast<uint32_t>(vrow + vh * S * ld);
        uint32_t hi_off = static_cast<uint32_t>(vrow + (N - 1 - vh * S) * ld);
        uint32_t step = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(ld));
        asm volatile("" : "+v"(lo_off), "+v"(hi_off), "+s"(step));
        float vl0[S], vh0[S], vl1[S], vh1[S];
#pragma unroll
        for (int j = 0; j < S; ++j) {
          vl0[j] = Vg[lo_off + j * step];
          vh0[j] = Vg[hi_off - j * step];
          vl1[j] = Vg[lo_off + m + j * step];
          vh1[j] = Vg[hi_off + m - j * step];
        }
#pragma unroll 1
        for (int c = 0; c < vchunks; ++c)
          v_group_shift2<S, VC>(vl0, vh0, vl1, vh1, hist + c * VC * hrow, hrow, vh);
        if (vtail)
          v_group_shift2<S, VC - 1>(vl0, vh0, vl1, vh1, hist + vchunks * VC * hrow, hrow, vh);
        asm volatile("" : "+v"(lo_off), "+v"(hi_off), "+s"(step));
#pragma unroll
        for (int j = 0; j < S; ++j) {
          Vg[lo_off + j * step] = vl0[j];
          Vg[hi_off - j * step] = vh0[j];
          Vg[lo_off + m + j * step] = vl1[j];
          Vg[hi_off + m - j * step] = vh1[j];
        }
      }
    }
    tv += VMEM ? group_end - group_start : G;
    __syncthreads();  // the next group overwrites the history (and V)

    // Convergence, at the first group boundary after each finished sweep.
    const bool at_limit = t >= limit;
    if (!(t / rounds_per_sweep > static_cast<int>(sweeps) || at_limit)) return false;
    sweeps = static_cast<uint32_t>(t / rounds_per_sweep);
    double off2 = 0.0, g2 = 0.0;
    float dmin = INFINITY;
    // s_i = sqrt(a_ii) in the history buffer (idle until the next group).
    float* sroot = reinterpret_cast<float*>(hist);
    #pragma unroll 1
    for (int i = tid; worker && i < N; i += nt)
      sroot[i] = i < n ? sqrtf(fmaxf(Ap[tri(i, i)], 0.0f)) : 0.0f;
    if (tid == 0) redf[0] = INFINITY;
    __syncthreads();
    #pragma unroll 1
    for (int e = tid; worker && e < packed; e += nt) {
      const int j = tri_col(e);
      const int i = e - (j * (j + 1)) / 2;
      if (j >= n) continue;
      const double v = Ap[e];
      if (i == j) {
        dmin = fminf(dmin, Ap[e]);
      } else {
        off2 += 2.0 * v * v;
        // The rate estimate needs no binary64 quotient: float ratio, double sum.
        const float den = sroot[i] + sroot[j];
        const float g = den > 0.0f ? Ap[e] * __frcp_rn(den) : (Ap[e] != 0.0f ? INFINITY : 0.0f);
        g2 += 2.0 * static_cast<double>(g * g);
      }
    }
    // (The minimum is exact in any order.)
    if (dmin < INFINITY) atomicMin(&redf[0], dmin);
    block_sum2(off2, g2, part0, part1, tid, nt);
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
    return stop != 0;
  };

#pragma unroll 1
  while (!group()) {
  }
  static_assert(VMEM || G == S, "register-resident V needs whole phase cycles");

  // Outputs: diagonal and square roots (ld entries), V (ld x ld, zero outside
  // the order).
  float* dg = diagonal + job.s_offset;
  float* rt = roots + job.s_offset;
  int bad = 0;
  #pragma unroll 1
  for (int i = tid; worker && i < ld; i += nt) {
    const float d = i < n ? Ap[tri(i, i)] : 0.0f;
    dg[i] = d;
    rt[i] = i < n ? sqrtf(d) : 0.0f;
    if (i < n && !(d > 0.0f && isfinite(d))) bad = 1;
  }
  const int round = tv % rounds_per_sweep;
  if constexpr (VMEM) {
    // Positions -> matrix indices: position x of round tv holds P_tv(x).
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
      const int vrow = pass * m + (tid >> 3);
      const int vh = tid & 7;
      float vl[S], vhv[S];
      const uint32_t base = static_cast<uint32_t>(vrow);
      if (worker) {
#pragma unroll
        for (int j = 0; j < S; ++j) {
          vl[j] = Vg[base + static_cast<uint32_t>((vh * S + j) * ld)];
          vhv[j] = Vg[base + static_cast<uint32_t>((N - 1 - (vh * S + j)) * ld)];
        }
      }
      __syncthreads();
      if (worker) {
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const int kl = participant(round, vh * S + j, N);
          const int kh = participant(round, N - 1 - (vh * S + j), N);
          Vg[base + static_cast<uint32_t>(kl * ld)] = vl[j];
          Vg[base + static_cast<uint32_t>(kh * ld)] = vhv[j];
        }
      }
      __syncthreads();
    }
  }
  #pragma unroll 1
  for (int e = tid; worker && e < ld * ld; e += nt) {
    const int i = e % ld;
    const int x = e / ld;
    if (i >= n || x >= n) Vg[e] = 0.0f;
  }
  if (!VMEM && worker && vlane && row < n) {
    // G = S, so V's phase is 0: register r holds slot h S + r; the position
    // x of a slot holds matrix index P_t(x) of the round reached.
#pragma unroll
    for (int r = 0; r < S; ++r) {
      const int kl = participant(round, h * S + r, N);
      const int kh = participant(round, N - 1 - (h * S + r), N);
      if (kl < n) Vg[row + static_cast<size_t>(kl) * ld] = lo[r];
      if (kh < n) Vg[row + static_cast<size_t>(kh) * ld] = hi[r];
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

// One workgroup per matrix of one padded order.
template <int NP, bool WIDE>
__global__ __launch_bounds__(1024) void batched_jacobi_kernel(
    const JacobiJob* __restrict__ jobs, const double* __restrict__ a64,
    float* __restrict__ vectors, float* __restrict__ diagonal,
    float* __restrict__ roots, JacobiStatus* __restrict__ status) {
  __shared__ __align__(16) unsigned char smem[JacobiLds<NP, WIDE>::kBytes];
  jacobi_solve<NP, WIDE>(jobs[blockIdx.x], a64, vectors, diagonal, roots, status, smem,
                         static_cast<int>(blockDim.x));
}

// Persistent launch of the orders 129..256 (padded 144..256), whose
// workgroups each take a whole CU (A's packed triangle and the rotation
// history fill the LDS): one 1024-thread workgroup per CU takes the next
// job from a device counter until the jobs run out, so the CUs stay busy
// across order classes instead of waiting for each class's slowest matrix.
// The jobs come largest order first (longest processing time first).  Every
// workgroup reads the counter once per job and leaves when it passes the
// job count, so the grid drains.  A job's results are those of
// batched_jacobi_kernel for it, bit for bit.
constexpr int kPersistentMinPadded = 144;
constexpr size_t persistent_lds_bytes() {
  size_t b = 0;
  b = JacobiLds<144, false>::kBytes > b ? JacobiLds<144, false>::kBytes : b;
  b = JacobiLds<160, false>::kBytes > b ? JacobiLds<160, false>::kBytes : b;
  b = JacobiLds<176, false>::kBytes > b ? JacobiLds<176, false>::kBytes : b;
  b = JacobiLds<192, false>::kBytes > b ? JacobiLds<192, false>::kBytes : b;
  b = JacobiLds<208, false>::kBytes > b ? JacobiLds<208, false>::kBytes : b;
  b = JacobiLds<224, false>::kBytes > b ? JacobiLds<224, false>::kBytes : b;
  b = JacobiLds<240, false>::kBytes > b ? JacobiLds<240, false>::kBytes : b;
  b = JacobiLds<256, false>::kBytes > b ? JacobiLds<256, false>::kBytes : b;
  return b;
}
static_assert(persistent_lds_bytes() + 16 <= 160 * 1024, "persistent Jacobi LDS");

// One job of the persistent launch: a call, not inlined, so that the eight
// orders' bodies do not share one register allocation and stack; the LDS
// pointer keeps its address space across the call.
template <int NP>
__device__ __attribute__((noinline)) void persistent_job(
    const JacobiJob& job, const double* __restrict__ a64, float* __restrict__ vectors,
    float* __restrict__ diagonal, float* __restrict__ roots, JacobiStatus* __restrict__ status,
    __attribute__((address_space(3))) unsigned char* lds) {
  // The 1024-thread workgroup's first threads_for_order(NP) threads work
  // (blockDim.x - (1024 - that): equal, but not a compile-time constant).
  jacobi_solve<NP, false>(job, a64, vectors, diagonal, roots, status, (unsigned char*)lds,
                          static_cast<int>(blockDim.x) - (1024 - threads_for_order(NP, false)));
}

__global__ __launch_bounds__(1024) void persistent_jacobi_kernel(
    const JacobiJob* __restrict__ jobs, uint32_t count, uint32_t* __restrict__ counter,
    const double* __restrict__ a64, float* __restrict__ vectors, float* __restrict__ diagonal,
    float* __restrict__ roots, JacobiStatus* __restrict__ status) {
  __shared__ __align__(16) unsigned char smem[persistent_lds_bytes()];
  __shared__ uint32_t next;
  auto* lds = (__attribute__((address_space(3))) unsigned char*)smem;
  for (;;) {
    if (threadIdx.x == 0) next = atomicAdd(counter, 1u);
    __syncthreads();
    const uint32_t i = next;
    __syncthreads();  // every lane has read `next` before it is reused
    if (i >= count) return;
    const JacobiJob job = jobs[i];
    switch ((job.order + 15u) & ~15u) {
      case 144: persistent_job<144>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 160: persistent_job<160>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 176: persistent_job<176>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 192: persistent_job<192>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 208: persistent_job<208>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 224: persistent_job<224>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 240: persistent_job<240>(job, a64, vectors, diagonal, roots, status, lds); break;
      case 256: persistent_job<256>(job, a64, vectors, diagonal, roots, status, lds); break;
      default: break;  // the host launches orders 129..256 only
    }
    __syncthreads();  // the next job reuses the LDS
  }
}

template <int NP>
hipError_t launch_order(const JacobiJob* jobs, uint32_t count, int threads, bool wide,
                        const double* a64, float* vectors, float* diagonal, float* roots,
                        JacobiStatus* status, hipStream_t stream) {
  if constexpr (has_wide_shape(NP)) {
    if (wide) {
      batched_jacobi_kernel<NP, true><<<count, threads, 0, stream>>>(jobs, a64, vectors, diagonal,
                                                                     roots, status);
      return hipGetLastError();
    }
  }
  batched_jacobi_kernel<NP, false><<<count, threads, 0, stream>>>(jobs, a64, vectors, diagonal,
                                                                  roots, status);
  return hipGetLastError();
}

// CUs of the current device (queried once per device; 256 on MI355X).
uint32_t device_cu_count() {
  constexpr int kMaxDevices = 64;
  static std::atomic<uint32_t> cache[kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return 256u;
  uint32_t cu = cache[dev].load(std::memory_order_relaxed);
  if (cu == 0) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
      return 256u;
    cu = static_cast<uint32_t>(v);
    cache[dev].store(cu, std::memory_order_relaxed);
  }
  return cu;
}

// Wide shape when every workgroup of the launch is resident at once: the
// kernels of these orders use about 100 VGPRs, 20 waves per CU.
bool use_wide_shape(int N, uint32_t count) {
  if (!has_wide_shape(N)) return false;
  const uint32_t waves = static_cast<uint32_t>(threads_for_order(N, true) / 64);
  return count * waves <= device_cu_count() * 20u;
}

}  // namespace

uint32_t batched_jacobi_padded_order(uint32_t order) { return (order + 15u) & ~15u; }

int batched_jacobi_threads(uint32_t order, uint32_t count) {
  const int N = static_cast<int>(batched_jacobi_padded_order(order));
  return threads_for_order(N, use_wide_shape(N, count));
}

size_t batched_jacobi_lds_bytes(uint32_t order) {
  const int N = static_cast<int>(batched_jacobi_padded_order(order));
  const size_t packed = static_cast<size_t>(N) * (N + 1) / 2;
  const int S = slots_for_order(N);
  const size_t hist = static_cast<size_t>(group_for_order(N)) * 8 * padded_slots(S);
  // Reduction partials: in the history above N = 128, an array of their own
  // (two doubles per thread of the widest shape) below.
  const size_t part = N > 128 ? 2 * sizeof(double)
                              : 2 * sizeof(double) * static_cast<size_t>(threads_for_order(N, true));
  return ((packed + 3) & ~size_t{3}) * sizeof(float) + 2 * (N / 2) * sizeof(SlotRec) +
         hist * sizeof(float2) + part + 16 * sizeof(float) + 4 * sizeof(int);
}

hipError_t launch_persistent_jacobi(const JacobiJob* jobs, uint32_t count, uint32_t* counter,
                                    int grid, const double* a64, float* vectors,
                                    float* diagonal, float* roots, JacobiStatus* status,
                                    hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (!jobs || !counter || grid < 1) return hipErrorInvalidValue;
  static_assert(kPersistentMinPadded == ((kPersistentMinOrder + 15u) & ~15u),
                "persistent orders");
  const uint32_t blocks = static_cast<uint32_t>(grid) < count ? static_cast<uint32_t>(grid) : count;
  hipError_t e = launch_fill_u32(counter, 1, 0u, stream);
  if (e != hipSuccess) return e;
  persistent_jacobi_kernel<<<blocks, 1024, 0, stream>>>(jobs, count, counter, a64, vectors,
                                                        diagonal, roots, status);
  return hipGetLastError();
}

hipError_t launch_batched_jacobi(const JacobiJob* jobs, uint32_t count, uint32_t order,
                                 const double* a64, float* vectors, float* diagonal,
                                 float* roots, JacobiStatus* status, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (order < 1 || order > kBatchedJacobiMaxOrder) return hipErrorInvalidValue;
  const int N = static_cast<int>(batched_jacobi_padded_order(order));
  const int S = slots_for_order(N);
  const bool wide = use_wide_shape(N, count);
  const int threads = batched_jacobi_threads(order, count);
  const size_t lds = batched_jacobi_lds_bytes(order);
  const int m = N / 2;
  // The kernel's shape assumptions: whole parts per row, at most 1024 lanes
  // (the count the kernel's item dealing is compiled for), a thread for each
  // of the N / 4 + 1 pair-phase owners, every dealt super-block in some
  // thread's list, LDS within 160 KiB, and for V in device memory N / 2 rows
  // of 8 lanes (two rows per lane in the updates, two passes in the output).
  const int offblocks = m * (m - 1) / 2;
  if (parts_for_order(N) * S != m || threads > 1024 || threads % 64 != 0 || threads < m ||
      threads != threads_for_order(N, wide) || threads < N / 4 + 1 ||
      super_items(N) > items_per_thread(N) * threads ||
      single_blocks_per_thread(N, wide) * threads < offblocks || m < 8 || lds > 160 * 1024 ||
      (v_in_memory(S) && threads != 8 * m))
    return hipErrorInvalidValue;
  switch (N) {
#define MSQ_ORDER(NV) \
    case NV:                                                                                  \
      return launch_order<NV>(jobs, count, threads, wide, a64, vectors, diagonal, roots, status, \
                              stream);
    MSQ_ORDER(16) MSQ_ORDER(32) MSQ_ORDER(48) MSQ_ORDER(64) MSQ_ORDER(80) MSQ_ORDER(96)
    MSQ_ORDER(112) MSQ_ORDER(128) MSQ_ORDER(144) MSQ_ORDER(160) MSQ_ORDER(176) MSQ_ORDER(192)
    MSQ_ORDER(208) MSQ_ORDER(224) MSQ_ORDER(240) MSQ_ORDER(256)
#undef MSQ_ORDER
  }
  return hipErrorInvalidValue;
}


}  // namespace msq
