// Rotation arithmetic shared by the Jacobi kernels: the circle-method
// schedule in closed form, the binary64 rotation angles and the 2 x 2 block
// updates.
//
// Reference: makeJacobiSchedule (src/metal_solver.cpp:50-79), the rotation
// and element formulas of jacobi_rotations_f32 / jacobi_similarity_vectors_f32
// (metal_mixed_sqrt.metal:1331-1411).
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>

namespace msq {
namespace jmath {

// Round-robin (circle method) over np participants: in round t the
// participant at slot x is P_t(0) = 0, P_t(x) = 1 + ((x - 1 - t) mod (np - 1)),
// and slot k pairs with slot np - 1 - k.
__device__ __forceinline__ int participant(int t, int x, int np) {
  if (x == 0) return 0;
  int r = x - 1 - t;
  if (r < 0) r += np - 1;
  return 1 + r;
}

// Rotation (c, s) for the pair p < q from a_pp, a_qq, a_pq, in binary64.
// With d = a_qq - a_pp and e = 2 a_pq (tau = d / e),
//   t = sign(d e) |e| / (|d| + sqrt(d^2 + e^2)) = sign(tau) / (|tau| + sqrt(1 + tau^2)),
// c = 1 / sqrt(1 + t^2), s = t c.  The square roots and the reciprocal come
// from the hardware estimates (v_rsq_f64, v_rcp_f64) refined by two Newton
// steps each: a binary64 result within a few units of the last place, which
// rounds to the same float as the correctly rounded quotients and roots
// except at float rounding ties, and one third of their dependent chain (the
// pair phase is on every round's critical path).  d^2 + e^2 neither
// overflows nor underflows in binary64 for float inputs, and t is in [-1, 1].
template <int STEPS = 2>
__device__ __forceinline__ double rsqrt_refined(double w) {  // w > 0, normal
  double y = __builtin_amdgcn_rsq(w);
#pragma unroll
  for (int i = 0; i < STEPS; ++i) y = y * fma(-0.5 * w * y, y, 1.5);
  return y;
}
// STEPS = 1: one Newton step per estimate, still far below the float
// rounding of the result.  T_STEPS (the blocked eigensolver: 0) refines only
// the estimates that set t: an error there changes the angle, so the
// rotation leaves a_pq at that relative size instead of zero, far below the
// FP32 rounding of the update itself; c = 1 / sqrt(1 + t^2) keeps STEPS, so
// c^2 + s^2 = 1 to binary64 accuracy for whatever t came out.
template <int STEPS = 2, int T_STEPS = STEPS>
__device__ __forceinline__ float2 rotation_angles(float app, float aqq, float apq) {
  if (apq == 0.0f) return make_float2(1.0f, 0.0f);
  const double d = static_cast<double>(aqq) - static_cast<double>(app);
  const double e = 2.0 * static_cast<double>(apq);
  const double w = fma(d, d, e * e);
  const double den = fabs(d) + w * rsqrt_refined<T_STEPS>(w);
  double r = __builtin_amdgcn_rcp(den);
#pragma unroll
  for (int i = 0; i < T_STEPS; ++i) r = r * fma(-den, r, 2.0);
  const double t = copysign(fabs(e) * r, d * e);
  const double c = rsqrt_refined<STEPS>(fma(t, t, 1.0));
  return make_float2(static_cast<float>(c), static_cast<float>(t * c));
}

// Rotated 2x2 block.  Rows (r0, r1) see (cr0, sr0) and (cr1, sr1); columns
// (c0, c1) likewise; a_ij is the entry at (row i, column j).
struct Block {
  float a00, a01, a10, a11;
};
__device__ __forceinline__ Block rotate_block(Block b, float2 r0, float2 r1,
                                              float2 c0, float2 c1) {
  Block o;
  o.a00 = r0.x * c0.x * b.a00 + r0.y * c0.x * b.a10 + r0.x * c0.y * b.a01 + r0.y * c0.y * b.a11;
  o.a01 = r0.x * c1.x * b.a01 + r0.y * c1.x * b.a11 + r0.x * c1.y * b.a00 + r0.y * c1.y * b.a10;
  o.a10 = r1.x * c0.x * b.a10 + r1.y * c0.x * b.a00 + r1.x * c0.y * b.a11 + r1.y * c0.y * b.a01;
  o.a11 = r1.x * c1.x * b.a11 + r1.y * c1.x * b.a01 + r1.x * c1.y * b.a10 + r1.y * c1.y * b.a00;
  return o;
}


// Off-diagonal 2 x 2 block {p_k, q_k} x {p_l, q_l} rotated on both sides:
// the reference's element formula A'[r][c] = rr.x cc.x A[r][c] + rr.y cc.x
// A[r'][c] + rr.x cc.y A[r][c'] + rr.y cc.y A[r'][c'] (p sees (c, -s), q
// (c, s)) evaluated as rows first, then columns: t = J_k^T B, o = t J_l,
// 16 multiply-adds with no coefficient products.
__device__ __forceinline__ void rotate_rows_cols(const Block& b, float ck, float sk, float cl,
                                                 float sl, Block& o) {
  const float t00 = ck * b.a00 - sk * b.a10;
  const float t01 = ck * b.a01 - sk * b.a11;
  const float t10 = ck * b.a10 + sk * b.a00;
  const float t11 = ck * b.a11 + sk * b.a01;
  o.a00 = cl * t00 - sl * t01;
  o.a01 = cl * t01 + sl * t00;
  o.a10 = cl * t10 - sl * t11;
  o.a11 = cl * t11 + sl * t10;
}

}  // namespace jmath
}  // namespace msq
