// Parallel cyclic Jacobi eigensolver, round-robin disjoint pairs.
//
// Reference: makeJacobiSchedule (src/metal_solver.cpp:50-79), symmetricJacobi
// (src/metal_runtime.cpp:584-646), jacobi_rotations_f32 and
// jacobi_similarity_vectors_f32 (metal_mixed_sqrt.metal:1331-1411).
//
// The reference dispatches two kernels per round, ping-ponging A and V through
// device memory (16 (n_pad - 1) dispatches per sweep pair).  Here:
//
//  * batched_jacobi_kernel: one workgroup per matrix for n <= 256.  A lives in
//    LDS as its packed upper triangle (n(n+1)/2 floats, 132 KiB at n = 256) and
//    is rotated in place: the blocks {p_k, q_k} x {p_l, q_l} of one round are
//    disjoint, so each block is read, rotated on both sides and written back
//    by one thread with no second buffer.  V lives in LDS up to n = 128 and in
//    device memory above.  The whole solve -- all sweeps, the FP64 off-norm
//    test after each sweep and the square roots -- is one launch for the whole
//    batch of uneven matrices.
//  * single-matrix global kernels for large n: one launch per round (rotation
//    angles fused into the block update kernel's prologue is not possible since
//    all blocks need all angles, so angles are one tiny launch), A kept exactly
//    symmetric in full storage.
//
// The schedule is the reference's circle method in closed form: in round t the
// participant at slot x is P_t(0) = 0, P_t(x) = 1 + ((x - 1 - t) mod (N - 1))
// for N = n + n % 2, and slot k pairs with slot N - 1 - k.  An odd order gets
// the reference's dummy index n, represented by a zero row and column (its
// rotation is the identity because its off-diagonal entry is zero).
//
// Rotation angles: tau = (a_qq - a_pp) / (2 a_pq), t = sign(tau) / (|tau| +
// sqrt(1 + tau^2)), c = 1 / sqrt(1 + t^2), s = t c, with p < q, evaluated in
// binary64 and rounded once to float (the reference evaluates them in float);
// entries are updated with the reference's element formula
//   A'[r][c] = rr.x cc.x A[r][c] + rr.y cc.x A[r'][c] + rr.x cc.y A[r][c'] +
//              rr.y cc.y A[r'][c']
// where index p sees (c, -s) and index q sees (c, s).
#include <hip/hip_runtime.h>

#include <math.h>

#include "kernels.hpp"

namespace msq {
namespace {

__device__ __forceinline__ int participant(int t, int x, int np) {
  if (x == 0) return 0;
  int r = x - 1 - t;
  if (r < 0) r += np - 1;
  return 1 + r;
}

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

// Rotation (c, s) for the pair p < q from a_pp, a_qq, a_pq, in binary64.
__device__ __forceinline__ float2 rotation_angles(float app, float aqq, float apq) {
  if (apq == 0.0f) return make_float2(1.0f, 0.0f);
  const double tau = (static_cast<double>(aqq) - static_cast<double>(app)) /
                     (2.0 * static_cast<double>(apq));
  const double t = copysign(1.0 / (fabs(tau) + sqrt(1.0 + tau * tau)), tau);
  const double c = 1.0 / sqrt(1.0 + t * t);
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

// Block id -> slots (k, l), k <= l, over an m x m upper triangle.  Rows k and
// m - 1 - k are paired so every unit has m + 1 blocks; an odd m leaves the
// middle row on its own.
__device__ __forceinline__ void block_slots(int b, int m, int& k, int& l) {
  const int half = m / 2;
  const int paired = half * (m + 1);
  if (b < paired) {
    const int rp = b / (m + 1);
    const int j = b - rp * (m + 1);
    if (j < m - rp) {
      k = rp;
      l = rp + j;
    } else {
      k = m - 1 - rp;
      l = k + (j - (m - rp));
    }
  } else {
    k = half;  // only when m is odd
    l = half + (b - paired);
  }
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}
__device__ __forceinline__ float wave_min(float v) {
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off));
  return v;
}

template <int NT, bool V_IN_LDS>
__global__ __launch_bounds__(NT) void batched_jacobi_kernel(
    const JacobiJob* __restrict__ jobs, const double* __restrict__ a64,
    float* __restrict__ vectors, float* __restrict__ diagonal,
    float* __restrict__ roots, JacobiStatus* __restrict__ status) {
  const JacobiJob job = jobs[blockIdx.x];
  const int tid = threadIdx.x;
  const int n = static_cast<int>(job.order);
  const int ld = static_cast<int>(job.ld);
  const int np = n + (n & 1);
  const int m = np / 2;
  const int packed = np * (np + 1) / 2;
  const int blocks = m * (m + 1) / 2;

  extern __shared__ __align__(16) unsigned char smem[];
  float* Ap = reinterpret_cast<float*>(smem);
  float* Vl = Ap + ((packed + 3) & ~3);
  int* slot_p = reinterpret_cast<int*>(Vl + (V_IN_LDS ? n * np : 0));
  int* slot_q = slot_p + 128;
  float2* rot = reinterpret_cast<float2*>(slot_q + 128);
  double* red = reinterpret_cast<double*>(rot + 128);  // 2 * NT / 64
  float* redf = reinterpret_cast<float*>(red + 2 * (NT / 64));
  int* ctrl = reinterpret_cast<int*>(redf + NT / 64);

  const double* A = a64 + job.a_offset;
  float* Vg = vectors + job.v_offset;
  float* V = V_IN_LDS ? Vl : Vg;
  const int vld = V_IN_LDS ? n : ld;

  // Load A (binary64 -> float, as the reference does) and its FP64 norm.
  double norm2 = 0.0;
  for (int e = tid; e < packed; e += NT) {
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
  // V = I over n x np (the dummy column of an odd order stays zero); in device
  // memory the whole ld x ld tile is written so the padding is zero.
  if (V_IN_LDS) {
    for (int e = tid; e < n * np; e += NT) {
      const int i = e % n;
      const int x = e / n;
      Vl[e] = i == x ? 1.0f : 0.0f;
    }
  } else {
    for (int e = tid; e < ld * ld; e += NT) {
      const int i = e % ld;
      const int x = e / ld;
      Vg[e] = (i == x && i < n) ? 1.0f : 0.0f;
    }
  }
  norm2 = wave_sum(norm2);
  if ((tid & 63) == 0) red[tid / 64] = norm2;
  __syncthreads();
  if (tid == 0) {
    double s = 0.0;
    for (int w = 0; w < NT / 64; ++w) s += red[w];
    red[2 * (NT / 64) - 1] = s;  // stash
  }
  __syncthreads();
  const double matrix_norm = sqrt(red[2 * (NT / 64) - 1]);
  __syncthreads();

  const uint32_t fixed = job.fixed_sweeps;
  const uint32_t cap = fixed ? fixed : job.max_sweeps;
  double previous = INFINITY;
  double off = 0.0;
  float min_diag = 0.0f;
  uint32_t sweeps = 0;
  JacobiStatus* st = status + job.status_index;

  for (uint32_t sweep = 0; sweep < cap; ++sweep) {
    for (int t = 0; t < np - 1; ++t) {
      for (int k = tid; k < m; k += NT) {
        const int a = participant(t, k, np);
        const int b = participant(t, np - 1 - k, np);
        const int p = a < b ? a : b;
        const int q = a < b ? b : a;
        slot_p[k] = p;
        slot_q[k] = q;
        rot[k] = rotation_angles(Ap[tri(p, p)], Ap[tri(q, q)], Ap[tri(p, q)]);
      }
      __syncthreads();
      for (int b = tid; b < blocks; b += NT) {
        int k, l;
        block_slots(b, m, k, l);
        const int pk = slot_p[k], qk = slot_q[k];
        const float2 rk = rot[k];
        const float2 rpk = make_float2(rk.x, -rk.y);  // seen by p
        if (k == l) {
          Block blk;
          blk.a00 = Ap[tri(pk, pk)];
          blk.a01 = Ap[tri(pk, qk)];
          blk.a10 = blk.a01;
          blk.a11 = Ap[tri(qk, qk)];
          const Block o = rotate_block(blk, rpk, rk, rpk, rk);
          Ap[tri(pk, pk)] = o.a00;
          Ap[tri(pk, qk)] = o.a01;
          Ap[tri(qk, qk)] = o.a11;
        } else {
          const int pl = slot_p[l], ql = slot_q[l];
          const float2 rl = rot[l];
          const float2 rpl = make_float2(rl.x, -rl.y);
          const int i00 = tri_sym(pk, pl), i01 = tri_sym(pk, ql);
          const int i10 = tri_sym(qk, pl), i11 = tri_sym(qk, ql);
          Block blk;
          blk.a00 = Ap[i00];
          blk.a01 = Ap[i01];
          blk.a10 = Ap[i10];
          blk.a11 = Ap[i11];
          const Block o = rotate_block(blk, rpk, rk, rpl, rl);
          Ap[i00] = o.a00;
          Ap[i01] = o.a01;
          Ap[i10] = o.a10;
          Ap[i11] = o.a11;
        }
      }
      for (int item = tid; item < n * m; item += NT) {
        const int i = item % n;
        const int k = item / n;
        const int p = slot_p[k], q = slot_q[k];
        const float2 r = rot[k];
        float* vp = V + i + static_cast<size_t>(p) * vld;
        float* vq = V + i + static_cast<size_t>(q) * vld;
        const float x = *vp, y = *vq;
        *vp = r.x * x + (-r.y) * y;
        *vq = r.x * y + r.y * x;
      }
      __syncthreads();
    }
    ++sweeps;
    // FP64 off-diagonal norm and the smallest diagonal entry.
    double off2 = 0.0;
    float dmin = INFINITY;
    for (int e = tid; e < packed; e += NT) {
      const int j = tri_col(e);
      const int i = e - (j * (j + 1)) / 2;
      if (j >= n) continue;
      const double v = Ap[e];
      if (i == j) {
        dmin = fminf(dmin, Ap[e]);
      } else {
        off2 += 2.0 * v * v;
      }
    }
    off2 = wave_sum(off2);
    dmin = wave_min(dmin);
    if ((tid & 63) == 0) {
      red[tid / 64] = off2;
      redf[tid / 64] = dmin;
    }
    __syncthreads();
    if (tid == 0) {
      double s = 0.0;
      float d = INFINITY;
      for (int w = 0; w < NT / 64; ++w) {
        s += red[w];
        d = fminf(d, redf[w]);
      }
      off = sqrt(s);
      min_diag = d;
      if (sweeps <= kJacobiHistory) st->off_history[sweeps - 1] = static_cast<float>(off / matrix_norm);
      int stop = 0;
      if (fixed) {
        stop = sweeps >= fixed;
      } else {
        double tol = static_cast<double>(job.off_tol) * matrix_norm;
        if (job.off_tol_lambda > 0.0f && d > 0.0f) {
          const double tl = static_cast<double>(job.off_tol_lambda) * d;
          tol = tl > tol ? tl : tol;
        }
        stop = off <= tol || sweeps >= cap ||
               (off > static_cast<double>(job.stagnation) * previous &&
                off <= static_cast<double>(job.floor_tol) * matrix_norm);
      }
      previous = off;
      ctrl[0] = stop;
      ctrl[1] = *reinterpret_cast<const int*>(&d);
    }
    __syncthreads();
    const int stop = ctrl[0];
    __syncthreads();
    if (stop) break;
  }

  // Outputs: diagonal, square roots, V (padding zero).
  float* dg = diagonal + job.s_offset;
  float* rt = roots + job.s_offset;
  int bad = 0;
  for (int i = tid; i < ld; i += NT) {
    const float d = i < n ? Ap[tri(i, i)] : 0.0f;
    dg[i] = d;
    rt[i] = i < n ? sqrtf(d) : 0.0f;
    if (i < n && !(d > 0.0f && isfinite(d))) bad = 1;
  }
  if (V_IN_LDS) {
    for (int e = tid; e < ld * ld; e += NT) {
      const int i = e % ld;
      const int x = e / ld;
      Vg[e] = (i < n && x < n) ? Vl[i + x * n] : 0.0f;
    }
  } else if (n & 1) {
    for (int i = tid; i < ld; i += NT) Vg[i + static_cast<size_t>(n) * ld] = 0.0f;
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
  }
}

// ----------------------------------------------------------- large n, global
__global__ void global_rotations_kernel(const float* __restrict__ A, int ld,
                                        int np, int t, int* __restrict__ slot_p,
                                        int* __restrict__ slot_q,
                                        float2* __restrict__ rot) {
  const int m = np / 2;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= m) return;
  const int a = participant(t, k, np);
  const int b = participant(t, np - 1 - k, np);
  const int p = a < b ? a : b;
  const int q = a < b ? b : a;
  slot_p[k] = p;
  slot_q[k] = q;
  rot[k] = rotation_angles(A[p + static_cast<size_t>(p) * ld],
                           A[q + static_cast<size_t>(q) * ld],
                           A[p + static_cast<size_t>(q) * ld]);
}

// A blocks (first `blocks` ids) and V items (the rest) of one round.
__global__ void global_round_kernel(float* __restrict__ A, float* __restrict__ V,
                                    int n, int ld, int np,
                                    const int* __restrict__ slot_p,
                                    const int* __restrict__ slot_q,
                                    const float2* __restrict__ rot) {
  const int m = np / 2;
  const long long blocks = static_cast<long long>(m) * (m + 1) / 2;
  const long long id = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (id < blocks) {
    int k, l;
    block_slots(static_cast<int>(id), m, k, l);
    const int pk = slot_p[k], qk = slot_q[k];
    const float2 rk = rot[k];
    const float2 rpk = make_float2(rk.x, -rk.y);
    auto at = [&](int r, int c) -> float& { return A[r + static_cast<size_t>(c) * ld]; };
    if (k == l) {
      Block blk;
      blk.a00 = at(pk, pk);
      blk.a01 = at(pk, qk);
      blk.a10 = blk.a01;
      blk.a11 = at(qk, qk);
      const Block o = rotate_block(blk, rpk, rk, rpk, rk);
      at(pk, pk) = o.a00;
      at(pk, qk) = o.a01;
      at(qk, pk) = o.a01;
      at(qk, qk) = o.a11;
    } else {
      const int pl = slot_p[l], ql = slot_q[l];
      const float2 rl = rot[l];
      const float2 rpl = make_float2(rl.x, -rl.y);
      Block blk;
      blk.a00 = at(pk, pl);
      blk.a01 = at(pk, ql);
      blk.a10 = at(qk, pl);
      blk.a11 = at(qk, ql);
      const Block o = rotate_block(blk, rpk, rk, rpl, rl);
      at(pk, pl) = o.a00;
      at(pl, pk) = o.a00;
      at(pk, ql) = o.a01;
      at(ql, pk) = o.a01;
      at(qk, pl) = o.a10;
      at(pl, qk) = o.a10;
      at(qk, ql) = o.a11;
      at(ql, qk) = o.a11;
    }
    return;
  }
  const long long item = id - blocks;
  if (item >= static_cast<long long>(n) * m) return;
  const int i = static_cast<int>(item % n);
  const int k = static_cast<int>(item / n);
  const int p = slot_p[k], q = slot_q[k];
  const float2 r = rot[k];
  float* vp = V + i + static_cast<size_t>(p) * ld;
  float* vq = V + i + static_cast<size_t>(q) * ld;
  const float x = *vp, y = *vq;
  *vp = r.x * x + (-r.y) * y;
  *vq = r.x * y + r.y * x;
}

__global__ void global_init_kernel(const double* __restrict__ a64, float* __restrict__ A,
                                   float* __restrict__ V, int n, int ld) {
  const size_t total = static_cast<size_t>(ld) * ld;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(e % ld);
    const int j = static_cast<int>(e / ld);
    const bool in = i < n && j < n;
    A[e] = in ? static_cast<float>(a64[e]) : 0.0f;
    V[e] = (in && i == j) ? 1.0f : 0.0f;
  }
}

// Partial sums of the FP64 off-diagonal norm and the diagonal minimum.
__global__ void global_offnorm_kernel(const float* __restrict__ A, int n, int ld,
                                      double* __restrict__ out_off2,
                                      float* __restrict__ out_min) {
  double s = 0.0;
  float d = INFINITY;
  const size_t total = static_cast<size_t>(n) * n;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < total;
       e += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const int i = static_cast<int>(e % n);
    const int j = static_cast<int>(e / n);
    const float v = A[i + static_cast<size_t>(j) * ld];
    if (i == j) {
      d = fminf(d, v);
    } else {
      s += static_cast<double>(v) * v;
    }
  }
  __shared__ double sd[16];
  __shared__ float sf[16];
  s = wave_sum(s);
  d = wave_min(d);
  if ((threadIdx.x & 63) == 0) {
    sd[threadIdx.x / 64] = s;
    sf[threadIdx.x / 64] = d;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    float m = INFINITY;
    for (int w = 0; w < static_cast<int>(blockDim.x / 64); ++w) {
      t += sd[w];
      m = fminf(m, sf[w]);
    }
    out_off2[blockIdx.x] = t;
    out_min[blockIdx.x] = m;
  }
}

__global__ void global_finish_kernel(const float* __restrict__ A, int n, int ld,
                                     float* __restrict__ diagonal,
                                     float* __restrict__ roots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ld) return;
  const float d = i < n ? A[i + static_cast<size_t>(i) * ld] : 0.0f;
  diagonal[i] = d;
  roots[i] = i < n ? sqrtf(d) : 0.0f;
}

template <int NT, bool VL>
hipError_t launch_batched_variant(const JacobiJob* jobs, uint32_t count,
                                  const double* a64, float* vectors, float* diagonal,
                                  float* roots, JacobiStatus* status, size_t lds,
                                  hipStream_t stream) {
  // Dynamic LDS above the default needs an explicit opt-in, sized to this
  // launch (static LDS plus the opt-in must stay within 160 KiB).
  static size_t configured = 0;
  if (lds > configured) {
    hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&batched_jacobi_kernel<NT, VL>),
        hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    configured = lds;
  }
  batched_jacobi_kernel<NT, VL><<<count, NT, lds, stream>>>(jobs, a64, vectors, diagonal,
                                                            roots, status);
  return hipGetLastError();
}

}  // namespace

size_t batched_jacobi_lds_bytes(uint32_t order, bool v_in_lds, int threads) {
  const size_t np = order + (order & 1u);
  const size_t packed = np * (np + 1) / 2;
  size_t bytes = ((packed + 3) & ~size_t{3}) * sizeof(float);
  if (v_in_lds) bytes += static_cast<size_t>(order) * np * sizeof(float);
  bytes += 2 * 128 * sizeof(int) + 128 * sizeof(float2);
  bytes += 2 * (threads / 64) * sizeof(double) + (threads / 64) * sizeof(float);
  bytes += 4 * sizeof(int);
  return bytes;
}

hipError_t launch_batched_jacobi(const JacobiJob* jobs, uint32_t count,
                                 uint32_t max_order, bool v_in_lds, int threads,
                                 const double* a64, float* vectors, float* diagonal,
                                 float* roots, JacobiStatus* status,
                                 hipStream_t stream) {
  if (count == 0) return hipSuccess;
  if (max_order > kBatchedJacobiMaxOrder || (v_in_lds && max_order > 128))
    return hipErrorInvalidValue;
  const size_t lds = batched_jacobi_lds_bytes(max_order, v_in_lds, threads);
  if (lds > 160 * 1024) return hipErrorInvalidValue;
  if (v_in_lds) {
    if (threads == 64) return launch_batched_variant<64, true>(jobs, count, a64, vectors, diagonal, roots, status, lds, stream);
    if (threads == 128) return launch_batched_variant<128, true>(jobs, count, a64, vectors, diagonal, roots, status, lds, stream);
    if (threads == 256) return launch_batched_variant<256, true>(jobs, count, a64, vectors, diagonal, roots, status, lds, stream);
    if (threads == 512) return launch_batched_variant<512, true>(jobs, count, a64, vectors, diagonal, roots, status, lds, stream);
  } else {
    if (threads == 256) return launch_batched_variant<256, false>(jobs, count, a64, vectors, diagonal, roots, status, lds, stream);
    if (threads == 512) return launch_batched_variant<512, false>(jobs, count, a64, vectors, diagonal, roots, status, lds, stream);
    if (threads == 1024) return launch_batched_variant<1024, false>(jobs, count, a64, vectors, diagonal, roots, status, lds, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_global_jacobi_init(const double* a64, float* A, float* V,
                                     uint32_t n, uint32_t ld, hipStream_t stream) {
  const size_t total = static_cast<size_t>(ld) * ld;
  size_t grid = (total + 255) / 256;
  if (grid > 65536) grid = 65536;
  global_init_kernel<<<static_cast<int>(grid), 256, 0, stream>>>(a64, A, V, n, ld);
  return hipGetLastError();
}

hipError_t launch_global_jacobi_round(float* A, float* V, uint32_t n, uint32_t ld,
                                      uint32_t round, int* slot_p, int* slot_q,
                                      float2* rot, hipStream_t stream) {
  const int np = static_cast<int>(n + (n & 1u));
  const int m = np / 2;
  global_rotations_kernel<<<(m + 255) / 256, 256, 0, stream>>>(A, ld, np, round, slot_p,
                                                              slot_q, rot);
  const long long items = static_cast<long long>(m) * (m + 1) / 2 + static_cast<long long>(n) * m;
  global_round_kernel<<<static_cast<int>((items + 255) / 256), 256, 0, stream>>>(
      A, V, n, ld, np, slot_p, slot_q, rot);
  return hipGetLastError();
}

hipError_t launch_global_offnorm(const float* A, uint32_t n, uint32_t ld, double* off2,
                                 float* dmin, int partials, hipStream_t stream) {
  global_offnorm_kernel<<<partials, 256, 0, stream>>>(A, n, ld, off2, dmin);
  return hipGetLastError();
}

hipError_t launch_global_jacobi_finish(const float* A, uint32_t n, uint32_t ld,
                                       float* diagonal, float* roots,
                                       hipStream_t stream) {
  global_finish_kernel<<<(ld + 255) / 256, 256, 0, stream>>>(A, n, ld, diagonal, roots);
  return hipGetLastError();
}

}  // namespace msq
