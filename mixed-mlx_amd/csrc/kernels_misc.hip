// Elementwise and reduction kernels of the solve, batched over uneven
// segments (grid.y = segment).
//
// Reference counterparts: working_from_symmetric_f32 (metal_mixed_sqrt.metal:
// 1911-1925; float-float there, binary64 here; the correction's
// working_add_symmetric_correction, :1955-1970, is fused into the last
// correction product, kernels_gemm.hip), scale_columns_f32 (:21-37),
// transpose_square_f32 (:1413-1421), and the host reduction of
// measureF32Orthogonality (src/metal_runtime.cpp:765-775), done on the device.
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "device_codecs.hpp"

namespace msq {
namespace {

constexpr int kT = 256;

// Grid-stride over the ld x ld tile of segment blockIdx.y.
#define SEGMENT_LOOP(seg, e)                                                        \
  const Segment seg = segs[blockIdx.y];                                             \
  const size_t seg##_n = static_cast<size_t>(seg.ld) * seg.ld;                      \
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;       \
       e < seg##_n; e += static_cast<size_t>(gridDim.x) * blockDim.x)

// ------------------------------------------------------- tile-pair kernels
// Transposing and symmetrising passes work on pairs of 32 x 32 tiles
// {(ti, tj), (tj, ti)}, ti <= tj, so that both tiles are read and written
// along columns (coalesced) through LDS.  One workgroup per pair over the
// whole batch: Segment::reserved holds the number of pairs of the segments
// before it in the list (filled by the host), and a workgroup finds its
// segment by binary search.
constexpr int kTile = 32;
constexpr int kTilePitch = kTile + 1;

struct PairTile {
  Segment sg;
  int ti, tj;  // tile row / column of the first tile, ti <= tj
};

__device__ __forceinline__ PairTile pair_tile(const Segment* __restrict__ segs, uint32_t count) {
  const uint32_t b = blockIdx.x;
  uint32_t lo = 0, hi = count;  // last segment with reserved <= b
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (segs[mid].reserved <= b) lo = mid; else hi = mid;
  }
  PairTile pt;
  pt.sg = segs[lo];
  // Pairs of a segment enumerate ti <= tj column by column: column tj holds
  // tj + 1 pairs.
  const uint32_t r = b - pt.sg.reserved;
  int tj = static_cast<int>((sqrtf(8.0f * static_cast<float>(r) + 1.0f) - 1.0f) * 0.5f);
  while (tj > 0 && static_cast<uint32_t>(tj) * (tj + 1) / 2 > r) --tj;
  while (static_cast<uint32_t>(tj + 1) * (tj + 2) / 2 <= r) ++tj;
  pt.tj = tj;
  pt.ti = static_cast<int>(r - static_cast<uint32_t>(tj) * (tj + 1) / 2);
  return pt;
}

// P <- src[tile (ti, tj)], Q <- src[tile (tj, ti)] (row i of a tile at i,
// column j at j * kTilePitch).
__device__ __forceinline__ void load_pair(const float* __restrict__ src, const PairTile& pt,
                                          float* P, float* Q) {
  const size_t ld = pt.sg.ld;
  const float* m = src + pt.sg.offset;
  for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
    const int i = e % kTile, j = e / kTile;
    P[i + kTilePitch * j] =
        m[(kTile * pt.ti + i) + static_cast<size_t>(kTile * pt.tj + j) * ld];
    Q[i + kTilePitch * j] =
        m[(kTile * pt.tj + i) + static_cast<size_t>(kTile * pt.ti + j) * ld];
  }
  __syncthreads();
}

// x = (c + c^T) / 2 in binary64 (both tiles of the pair).
__global__ __launch_bounds__(256) void symmetrize_to_f64_kernel(const Segment* __restrict__ segs,
                                                                uint32_t count,
                                                                const float* __restrict__ c,
                                                                double* __restrict__ x,
                                                                const int* __restrict__ active) {
  __shared__ float P[kTile * kTilePitch], Q[kTile * kTilePitch];
  const PairTile pt = pair_tile(segs, count);
  if (active && !active[pt.sg.index]) return;
  load_pair(c, pt, P, Q);
  const size_t ld = pt.sg.ld;
  double* xm = x + pt.sg.offset;
  for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
    const int i = e % kTile, j = e / kTile;
    const double s = 0.5 * (static_cast<double>(P[i + kTilePitch * j]) +
                            static_cast<double>(Q[j + kTilePitch * i]));
    xm[(kTile * pt.ti + i) + static_cast<size_t>(kTile * pt.tj + j) * ld] = s;
    if (pt.ti != pt.tj) {
      const double t = 0.5 * (static_cast<double>(Q[i + kTilePitch * j]) +
                              static_cast<double>(P[j + kTilePitch * i]));
      xm[(kTile * pt.tj + i) + static_cast<size_t>(kTile * pt.ti + j) * ld] = t;
    }
  }
}

// out[k + m ld] = V[m][k] * f(s_k): the K-contiguous storage of G = V diag(f(s)),
// f = sqrt (use_sqrt) or identity.  Entries outside the order are zero.
__global__ __launch_bounds__(256) void scale_rows_transpose_kernel(
    const Segment* __restrict__ segs, uint32_t count, const float* __restrict__ v,
    const float* __restrict__ roots, int use_sqrt, float* __restrict__ out,
    const int* __restrict__ active) {
  __shared__ float P[kTile * kTilePitch], Q[kTile * kTilePitch];
  const PairTile pt = pair_tile(segs, count);
  if (active && !active[pt.sg.index]) return;
  load_pair(v, pt, P, Q);
  const size_t ld = pt.sg.ld;
  const int n = static_cast<int>(pt.sg.order);
  float* om = out + pt.sg.offset;
  auto put = [&](int tk, int tm, const float* T) {  // out tile (tk, tm) from V^T
    for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
      const int kk = kTile * tk + e % kTile, mm = kTile * tm + e / kTile;
      float val = 0.0f;
      if (kk < n && mm < n) {
        const float s = roots[pt.sg.vec_offset + kk];
        val = T[(e / kTile) + kTilePitch * (e % kTile)] * (use_sqrt ? sqrtf(s) : s);
      }
      om[kk + static_cast<size_t>(mm) * ld] = val;
    }
  };
  put(pt.ti, pt.tj, Q);  // out tile (ti, tj) = V tile (tj, ti)^T
  if (pt.ti != pt.tj) put(pt.tj, pt.ti, P);
}

// scale_rows_transpose_kernel with the result as BF16 codes (the codec's
// round-to-nearest-even encode of the same float values).
__global__ __launch_bounds__(256) void scale_rows_transpose_bf16_kernel(
    const Segment* __restrict__ segs, uint32_t count, const float* __restrict__ v,
    const float* __restrict__ roots, int use_sqrt, uint16_t* __restrict__ out,
    const int* __restrict__ active) {
  __shared__ float P[kTile * kTilePitch], Q[kTile * kTilePitch];
  const PairTile pt = pair_tile(segs, count);
  if (active && !active[pt.sg.index]) return;
  load_pair(v, pt, P, Q);
  const size_t ld = pt.sg.ld;
  const int n = static_cast<int>(pt.sg.order);
  uint16_t* om = out + pt.sg.offset;
  auto put = [&](int tk, int tm, const float* T) {  // out tile (tk, tm) from V^T
    for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
      const int kk = kTile * tk + e % kTile, mm = kTile * tm + e / kTile;
      float val = 0.0f;
      if (kk < n && mm < n) {
        const float s = roots[pt.sg.vec_offset + kk];
        val = T[(e / kTile) + kTilePitch * (e % kTile)] * (use_sqrt ? sqrtf(s) : s);
      }
      om[kk + static_cast<size_t>(mm) * ld] = encode_bf16(val);
    }
  };
  put(pt.ti, pt.tj, Q);
  if (pt.ti != pt.tj) put(pt.tj, pt.ti, P);
}

__global__ __launch_bounds__(256) void transpose_kernel(const Segment* __restrict__ segs,
                                                        uint32_t count,
                                                        const float* __restrict__ v,
                                                        float* __restrict__ out) {
  __shared__ float P[kTile * kTilePitch], Q[kTile * kTilePitch];
  const PairTile pt = pair_tile(segs, count);
  load_pair(v, pt, P, Q);
  const size_t ld = pt.sg.ld;
  float* om = out + pt.sg.offset;
  for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
    const int i = e % kTile, j = e / kTile;
    om[(kTile * pt.ti + i) + static_cast<size_t>(kTile * pt.tj + j) * ld] = Q[j + kTilePitch * i];
    if (pt.ti != pt.tj)
      om[(kTile * pt.tj + i) + static_cast<size_t>(kTile * pt.ti + j) * ld] =
          P[j + kTilePitch * i];
  }
}

// BF16 codes of V and of V^T from one read of V: the tile pair's two tiles
// are written as they are into cv and transposed into cvt (the codec's
// round-to-nearest-even encode, so the codes equal an encode of V and of a
// transpose of V).
__global__ __launch_bounds__(256) void transpose_encode_bf16_kernel(
    const Segment* __restrict__ segs, uint32_t count, const float* __restrict__ v,
    uint16_t* __restrict__ cv, uint16_t* __restrict__ cvt) {
  __shared__ float P[kTile * kTilePitch], Q[kTile * kTilePitch];
  const PairTile pt = pair_tile(segs, count);
  load_pair(v, pt, P, Q);
  const size_t ld = pt.sg.ld;
  uint16_t* const vm = cv + pt.sg.offset;
  uint16_t* const tm = cvt + pt.sg.offset;
  for (int e = threadIdx.x; e < kTile * kTile; e += blockDim.x) {
    const int i = e % kTile, j = e / kTile;
    const size_t at = (kTile * pt.ti + i) + static_cast<size_t>(kTile * pt.tj + j) * ld;
    vm[at] = encode_bf16(P[i + kTilePitch * j]);
    tm[at] = encode_bf16(Q[j + kTilePitch * i]);
    if (pt.ti != pt.tj) {
      const size_t ta = (kTile * pt.tj + i) + static_cast<size_t>(kTile * pt.ti + j) * ld;
      vm[ta] = encode_bf16(Q[i + kTilePitch * j]);
      tm[ta] = encode_bf16(P[j + kTilePitch * i]);
    }
  }
}

__global__ void zero_f64_kernel(double* p, size_t n) {
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[e] = 0.0;
}

// out = a - b, elementwise.
__global__ void subtract_kernel(const float* __restrict__ a, const float* __restrict__ b,
                                float* __restrict__ out, size_t n) {
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += static_cast<size_t>(gridDim.x) * blockDim.x)
    out[e] = a[e] - b[e];
}

// sum (G_ij - delta_ij)^2 over the order x order block, per segment.
__global__ void orthogonality_kernel(const Segment* __restrict__ segs,
                                     const float* __restrict__ gram,
                                     double* __restrict__ out) {
  double s = 0.0;
  SEGMENT_LOOP(sg, e) {
    const int i = static_cast<int>(e % sg.ld);
    const int j = static_cast<int>(e / sg.ld);
    if (i < static_cast<int>(sg.order) && j < static_cast<int>(sg.order)) {
      const double d = static_cast<double>(gram[sg.offset + e]) - (i == j ? 1.0 : 0.0);
      s += d * d;
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(out + segs[blockIdx.y].index, s);
}

// Segmented amax (finite |v| bit patterns) for per-matrix FP8 scales.
__global__ void segment_amax_kernel(const Segment* __restrict__ segs,
                                    const float* __restrict__ src,
                                    uint32_t* __restrict__ amax) {
  uint32_t m = 0;
  SEGMENT_LOOP(sg, e) {
    const uint32_t a = __float_as_uint(src[sg.offset + e]) & 0x7fffffffu;
    if (a < 0x7f800000u) m = max(m, a);
  }
  for (int off = 32; off > 0; off >>= 1)
    m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), off)));
  if ((threadIdx.x & 63) == 0) atomicMax(amax + segs[blockIdx.y].index, m);
}

__global__ void segment_sigma_kernel(const uint32_t* __restrict__ amax, float* __restrict__ sigma,
                                     uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float a = __uint_as_float(amax[i]);
  sigma[i] = a == 0.0f ? 1.0f : static_cast<float>(448.0 / static_cast<double>(a));
}

// Encodes src * sigma[segment] to E4M3FN over each segment.
__global__ void segment_encode_fp8_kernel(const Segment* __restrict__ segs,
                                          const float* __restrict__ src,
                                          const float* __restrict__ sigma,
                                          uint8_t* __restrict__ dst) {
  const float s = sigma[segs[blockIdx.y].index];
  SEGMENT_LOOP(sg, e) { dst[sg.offset + e] = encode_e4m3(src[sg.offset + e] * s); }
}

__global__ void alpha_kernel(const float* __restrict__ sa, const float* __restrict__ sb,
                             float* __restrict__ alpha, uint32_t count) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  alpha[i] = static_cast<float>(1.0 / (static_cast<double>(sa[i]) * static_cast<double>(sb[i])));
}

// Residual bookkeeping after R = A - X^2: relative residual, forward-error
// estimate ||R||_F / (2 s_min ||X||_F) with ||X||_F^2 = trace(A), the best
// iterate, the stop rules and the next iteration's active flag.
//  * Fixed budget: stop on the residual or estimate target, the budget, or
//    divergence (the residual grew past 4x the initial root's).
//  * Stagnation (stagnation_window > 0; src/metal_solver.cpp:883-902): also
//    stop after `window` corrections in a row that improve the best residual
//    by no more than minimum_relative_improvement; the caller keeps the best
//    iterate (save[i] = 1 when this step's is the new best).
__global__ void convergence_kernel(SolveState* __restrict__ st, const double* __restrict__ norms,
                                   int* __restrict__ active, int* __restrict__ save,
                                   uint32_t count, uint32_t step) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  if (save) save[i] = 0;
  SolveState& s = st[i];
  if (!active[i]) return;
  const double rel = sqrt(norms[i]) / s.matrix_norm;
  const double fwd = sqrt(norms[i]) / (2.0 * s.min_root * s.root_norm);
  if (step < kSolveHistory) {
    s.residuals[step] = rel;
    s.forward_estimates[step] = fwd;
  }
  s.steps = step + 1;
  s.final_residual = rel;
  s.final_forward_estimate = fwd;
  const bool finite = isfinite(rel);
  const double prior = s.best_residual;
  const double improvement = step >= 1 && prior > 0.0 ? (prior - rel) / prior : 0.0;
  if (finite && rel < prior) {
    s.best_residual = rel;
    s.best_iteration = step;
    if (save) save[i] = 1;
  }
  const bool met = rel <= s.residual_tolerance || fwd <= s.forward_tolerance;
  const uint32_t done_corrections = step;  // step 0 is the initial root
  int next = 1;
  if (!finite) {
    s.status = kSolveNonfinite;
    next = 0;
  } else if (met && done_corrections >= s.min_corrections) {
    s.status = kSolveConverged;
    next = 0;
  } else if (s.stagnation_window > 0 && step >= 1 && !(rel <= s.residual_tolerance)) {
    if (improvement > s.minimum_relative_improvement)
      s.stagnation_steps = 0;
    else
      ++s.stagnation_steps;
    if (s.stagnation_steps >= s.stagnation_window) {
      s.status = kSolveStagnated;
      next = 0;
    }
  }
  if (next && done_corrections >= s.max_corrections) {
    s.status = met ? kSolveConverged : kSolveBudgetExhausted;
    next = 0;
  }
  // Diverged: the residual grew well past the initial root's.
  if (next && s.stagnation_window == 0 && step >= 1 && rel > 4.0 * s.residuals[0]) {
    s.status = kSolveDiverged;
    next = 0;
  }
  s.corrections = done_corrections;
  active[i] = next;
}

// Per-matrix solve state from the Jacobi results: ||A||_F, sqrt(trace A) =
// ||A^(1/2)||_F, the smallest root, the resolution test; tolerances and
// budgets from the options.
// allow_refine: a matrix whose only fault is a diagonal entry below the
// resolution (zero or negative included) stays running when its options ask
// for the second Jacobi pass, which decides it.
__global__ void solve_state_kernel(const Segment* __restrict__ segs,
                                   const JacobiStatus* __restrict__ js,
                                   const float* __restrict__ diagonal,
                                   const SolveLimits* __restrict__ limits,
                                   SolveState* __restrict__ st, int* __restrict__ active,
                                   uint32_t count, int allow_refine) {
  // One wave per matrix: the lanes stride over the diagonal, then a
  // butterfly (a fixed order) gives every lane the sums; lane 0 writes.
  const uint32_t i = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = static_cast<int>(threadIdx.x & 63);
  if (i >= count) return;  // whole waves
  const Segment sg = segs[i];
  double trace = 0.0;
  float dmin = INFINITY, dmax = 0.0f;
  int finite = 1;
  for (uint32_t k = static_cast<uint32_t>(lane); k < sg.order; k += 64) {
    const float d = diagonal[sg.vec_offset + k];
    trace += d;
    dmin = fminf(dmin, d);
    dmax = fmaxf(dmax, d);
    finite = finite && isfinite(d);
  }
  for (int o = 32; o > 0; o >>= 1) {
    trace += __shfl_xor(trace, o);
    dmin = fminf(dmin, __shfl_xor(dmin, o));
    dmax = fmaxf(dmax, __shfl_xor(dmax, o));
    finite = finite & __shfl_xor(finite, o);
  }
  if (lane != 0) return;
  const SolveLimits& lim = limits[i];
  SolveState s = {};
  s.matrix_norm = js[i].matrix_norm;
  s.root_norm = sqrt(trace > 0.0 ? trace : 0.0);
  s.min_root = dmin > 0.0f ? sqrt(static_cast<double>(dmin)) : 0.0;
  s.residual_tolerance = lim.residual_tolerance;
  s.forward_tolerance = lim.forward_tolerance;
  s.minimum_relative_improvement = lim.minimum_relative_improvement;
  s.min_corrections = lim.min_corrections;
  s.max_corrections = lim.max_corrections;
  s.stagnation_window = lim.stagnation_window;
  s.best_residual = INFINITY;
  s.basis_orthogonality = 0.0;
  s.fp32_orthogonality = NAN;
  s.frozen_min_root = s.min_root;
  const bool norm_ok = isfinite(s.matrix_norm) && s.matrix_norm > 0.0;
  const bool valid = (js[i].flags & kJacobiInvalidSpectrum) == 0 && dmin > 0.0f && norm_ok;
  s.unresolved =
      finite && norm_ok && dmax > 0.0f && lim.resolution > 0.0f && dmin < lim.resolution * dmax
          ? 1u
          : 0u;
  const bool running = valid || (allow_refine && lim.refine && s.unresolved);
  s.status = running ? kSolveRunning : kSolveInvalidSpectrum;
  st[i] = s;
  active[i] = running ? 1 : 0;
}

// measureF32Orthogonality's acceptance test (src/metal_solver.cpp:217-236):
// the basis tier is kept when ||V_t^T V_t - I||_F / sqrt(n) <= tolerance.
__global__ void basis_escalation_kernel(const Segment* __restrict__ segs, uint32_t count,
                                        const double* __restrict__ orth2, double tolerance,
                                        const SolveLimits* __restrict__ limits,
                                        const int* __restrict__ active, int* __restrict__ esc,
                                        SolveState* __restrict__ st) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const Segment sg = segs[k];
  const uint32_t i = sg.index;
  const double tol =
      limits && limits[i].basis_tolerance > 0.0 ? limits[i].basis_tolerance : tolerance;
  const double o = sqrt(orth2[i] / static_cast<double>(sg.order));
  const int e = (active[i] && !(o <= tol)) ? 1 : 0;
  esc[i] = e;
  st[i].basis_orthogonality = o;
  st[i].escalated = static_cast<uint32_t>(e);
}

__global__ void record_orthogonality_kernel(const Segment* __restrict__ segs, uint32_t count,
                                            const double* __restrict__ orth2,
                                            SolveState* __restrict__ st) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const Segment sg = segs[k];
  st[sg.index].fp32_orthogonality = sqrt(orth2[sg.index] / static_cast<double>(sg.order));
}

__global__ void reject_spectrum_kernel(const Segment* __restrict__ segs, uint32_t count,
                                       const int* __restrict__ rejected,
                                       SolveState* __restrict__ st, int* __restrict__ active) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const uint32_t i = segs[k].index;
  if (rejected[i] && active[i]) {
    st[i].status = kSolveInvalidSpectrum;
    active[i] = 0;
  }
}

// Smallest frozen root of each segment (the roots X0 is formed with).
__global__ void frozen_roots_kernel(const Segment* __restrict__ segs, uint32_t count,
                                    const float* __restrict__ qroots,
                                    const int* __restrict__ rejected,
                                    SolveState* __restrict__ st) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= count) return;
  const Segment sg = segs[k];
  float m = INFINITY;
  for (uint32_t r = 0; r < sg.order; ++r) m = fminf(m, qroots[sg.vec_offset + r]);
  st[sg.index].frozen_min_root = m;
  st[sg.index].spectrum_rejected = rejected ? static_cast<uint32_t>(rejected[sg.index]) : 0u;
}

__global__ void copy_tiles_f64_kernel(const Segment* __restrict__ segs,
                                      const double* __restrict__ src, double* __restrict__ dst,
                                      const int* __restrict__ flag) {
  if (flag && !flag[segs[blockIdx.y].index]) return;
  SEGMENT_LOOP(sg, e) { dst[sg.offset + e] = src[sg.offset + e]; }
}

// The upper triangle (i <= j) of each active segment's order: x_ij and x_ji
// both become ((x_ij + d_ij) + (x_ji + d_ji)) / 2, the diagonal x_ii + d_ii.
__global__ void add_symmetrize_f64_kernel(const Segment* __restrict__ segs,
                                          double* __restrict__ x, const float* __restrict__ d,
                                          const int* __restrict__ active) {
  if (active && !active[segs[blockIdx.y].index]) return;
  SEGMENT_LOOP(sg, e) {
    const uint32_t i = static_cast<uint32_t>(e % sg.ld), j = static_cast<uint32_t>(e / sg.ld);
    if (i > j || j >= sg.order) continue;
    const size_t ij = sg.offset + e, ji = sg.offset + j + static_cast<size_t>(i) * sg.ld;
    if (i == j) {
      x[ij] = x[ij] + static_cast<double>(d[ij]);
    } else {
      const double a = x[ij] + static_cast<double>(d[ij]);
      const double b = x[ji] + static_cast<double>(d[ji]);
      const double v = 0.5 * (a + b);
      x[ij] = v;
      x[ji] = v;
    }
  }
}

__global__ void sylvester_divide_kernel(const Segment* __restrict__ segs,
                                        const float* __restrict__ in,
                                        const float* __restrict__ roots, float* __restrict__ out,
                                        const int* __restrict__ active) {
  if (active && !active[segs[blockIdx.y].index]) return;
  SEGMENT_LOOP(sg, e) {
    const uint32_t i = static_cast<uint32_t>(e % sg.ld), j = static_cast<uint32_t>(e / sg.ld);
    float v = 0.0f;
    if (i < sg.order && j < sg.order)
      v = in[sg.offset + e] / (roots[sg.vec_offset + i] + roots[sg.vec_offset + j]);
    out[sg.offset + e] = v;
  }
}

__global__ void scale_columns_kernel(const Segment* __restrict__ segs, const float* __restrict__ in,
                                     const float* __restrict__ roots, float* __restrict__ out,
                                     const int* __restrict__ active) {
  if (active && !active[segs[blockIdx.y].index]) return;
  SEGMENT_LOOP(sg, e) {
    const uint32_t i = static_cast<uint32_t>(e % sg.ld), j = static_cast<uint32_t>(e / sg.ld);
    out[sg.offset + e] =
        i < sg.order && j < sg.order ? in[sg.offset + e] * roots[sg.vec_offset + j] : 0.0f;
  }
}

__global__ void copy_words_kernel(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                                  size_t n) {
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[e] = src[e];
}

__global__ void widen_f64_kernel(const float* __restrict__ src, double* __restrict__ dst,
                                 size_t n) {
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[e] = static_cast<double>(src[e]);
}

template <typename T>
__global__ void move_tiles_kernel(const Segment* __restrict__ from, const Segment* __restrict__ to,
                                  const T* __restrict__ src, T* __restrict__ dst) {
  const Segment f = from[blockIdx.y], t = to[blockIdx.y];
  const size_t n = static_cast<size_t>(f.ld) * f.ld;
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += static_cast<size_t>(gridDim.x) * blockDim.x)
    dst[t.offset + e] = src[f.offset + e];
}

__global__ void move_vectors_kernel(const Segment* __restrict__ from,
                                    const Segment* __restrict__ to, const float* __restrict__ src,
                                    float* __restrict__ dst) {
  const Segment f = from[blockIdx.y], t = to[blockIdx.y];
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < f.ld; e += gridDim.x * blockDim.x)
    dst[t.vec_offset + e] = src[f.vec_offset + e];
}

__global__ void fill_u32_kernel(uint32_t* p, size_t n, uint32_t value) {
  for (size_t e = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += static_cast<size_t>(gridDim.x) * blockDim.x)
    p[e] = value;
}

// sum of squares of the order x order block of a binary64 segment.
__global__ void sumsq_f64_kernel(const Segment* __restrict__ segs, const double* __restrict__ a,
                                 double* __restrict__ out) {
  double s = 0.0;
  SEGMENT_LOOP(sg, e) {
    const int i = static_cast<int>(e % sg.ld);
    const int j = static_cast<int>(e / sg.ld);
    if (i < static_cast<int>(sg.order) && j < static_cast<int>(sg.order)) {
      const double v = a[sg.offset + e];
      s += v * v;
    }
  }
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(out + segs[blockIdx.y].index, s);
}

int seg_grid_x(uint32_t max_ld) {
  const size_t n = static_cast<size_t>(max_ld) * max_ld;
  size_t g = (n + kT * 4 - 1) / (kT * 4);
  if (g < 1) g = 1;
  if (g > 1024) g = 1024;
  return static_cast<int>(g);
}

}  // namespace

hipError_t launch_symmetrize_to_f64(const Segment* segs, uint32_t count, uint32_t pairs,
                                    const float* c, double* x, const int* active,
                                    hipStream_t stream) {
  if (!count || !pairs) return hipSuccess;
  symmetrize_to_f64_kernel<<<pairs, kT, 0, stream>>>(segs, count, c, x, active);
  return hipGetLastError();
}
hipError_t launch_transpose_encode_bf16(const Segment* segs, uint32_t count, uint32_t pairs,
                                        const float* v, uint16_t* cv, uint16_t* cvt,
                                        hipStream_t stream) {
  if (!count || !pairs) return hipSuccess;
  transpose_encode_bf16_kernel<<<pairs, kT, 0, stream>>>(segs, count, v, cv, cvt);
  return hipGetLastError();
}
hipError_t launch_scale_rows_transpose_bf16(const Segment* segs, uint32_t count, uint32_t pairs,
                                            const float* v, const float* roots, int use_sqrt,
                                            uint16_t* out, const int* active, hipStream_t stream) {
  if (!count || !pairs) return hipSuccess;
  scale_rows_transpose_bf16_kernel<<<pairs, kT, 0, stream>>>(segs, count, v, roots, use_sqrt,
                                                             out, active);
  return hipGetLastError();
}
hipError_t launch_scale_rows_transpose(const Segment* segs, uint32_t count, uint32_t pairs,
                                       const float* v, const float* roots, int use_sqrt,
                                       float* out, const int* active, hipStream_t stream) {
  if (!count || !pairs) return hipSuccess;
  scale_rows_transpose_kernel<<<pairs, kT, 0, stream>>>(segs, count, v, roots, use_sqrt, out,
                                                        active);
  return hipGetLastError();
}
hipError_t launch_transpose(const Segment* segs, uint32_t count, uint32_t pairs, const float* v,
                            float* out, hipStream_t stream) {
  if (!count || !pairs) return hipSuccess;
  transpose_kernel<<<pairs, kT, 0, stream>>>(segs, count, v, out);
  return hipGetLastError();
}
hipError_t launch_subtract(const float* a, const float* b, float* out, size_t n,
                           hipStream_t stream) {
  if (!n) return hipSuccess;
  size_t g = (n + kT - 1) / kT;
  if (g > 16384) g = 16384;
  subtract_kernel<<<static_cast<int>(g), kT, 0, stream>>>(a, b, out, n);
  return hipGetLastError();
}
hipError_t launch_zero_f64(double* p, size_t n, hipStream_t stream) {
  if (!n) return hipSuccess;
  size_t g = (n + kT - 1) / kT;
  if (g > 4096) g = 4096;
  zero_f64_kernel<<<static_cast<int>(g), kT, 0, stream>>>(p, n);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, size_t n, uint32_t value, hipStream_t stream) {
  if (!n) return hipSuccess;
  size_t g = (n + kT - 1) / kT;
  if (g > 4096) g = 4096;
  fill_u32_kernel<<<static_cast<int>(g), kT, 0, stream>>>(p, n, value);
  return hipGetLastError();
}
hipError_t launch_sumsq_f64(const Segment* segs, uint32_t count, uint32_t max_ld, const double* a,
                            double* out, hipStream_t stream) {
  if (!count) return hipSuccess;
  sumsq_f64_kernel<<<dim3(seg_grid_x(max_ld), count), kT, 0, stream>>>(segs, a, out);
  return hipGetLastError();
}
hipError_t launch_orthogonality(const Segment* segs, uint32_t count, uint32_t max_ld,
                                const float* gram, double* out, hipStream_t stream) {
  if (!count) return hipSuccess;
  orthogonality_kernel<<<dim3(seg_grid_x(max_ld), count), kT, 0, stream>>>(segs, gram, out);
  return hipGetLastError();
}
hipError_t launch_segment_fp8(const Segment* segs, uint32_t count, uint32_t max_ld,
                              const float* src, uint32_t* amax, float* sigma, uint8_t* dst,
                              hipStream_t stream) {
  if (!count) return hipSuccess;
  segment_amax_kernel<<<dim3(seg_grid_x(max_ld), count), kT, 0, stream>>>(segs, src, amax);
  segment_sigma_kernel<<<(count + kT - 1) / kT, kT, 0, stream>>>(amax, sigma, count);
  segment_encode_fp8_kernel<<<dim3(seg_grid_x(max_ld), count), kT, 0, stream>>>(segs, src, sigma,
                                                                                dst);
  return hipGetLastError();
}
hipError_t launch_alpha(const float* sa, const float* sb, float* alpha, uint32_t count,
                        hipStream_t stream) {
  if (!count) return hipSuccess;
  alpha_kernel<<<(count + kT - 1) / kT, kT, 0, stream>>>(sa, sb, alpha, count);
  return hipGetLastError();
}
hipError_t launch_solve_state(const Segment* segs, const JacobiStatus* js,
                              const float* diagonal, const SolveLimits* limits, SolveState* st,
                              int* active, uint32_t count, bool allow_refine,
                              hipStream_t stream) {
  if (!count) return hipSuccess;
  constexpr uint32_t kPerBlock = kT / 64;  // one wave per matrix
  solve_state_kernel<<<(count + kPerBlock - 1) / kPerBlock, kT, 0, stream>>>(
      segs, js, diagonal, limits, st, active, count, allow_refine ? 1 : 0);
  return hipGetLastError();
}
hipError_t launch_basis_escalation(const Segment* segs, uint32_t count, const double* orth2,
                                   double tolerance, const SolveLimits* limits, const int* active,
                                   int* esc, SolveState* st, hipStream_t stream) {
  if (!count) return hipSuccess;
  basis_escalation_kernel<<<(count + kT - 1) / kT, kT, 0, stream>>>(segs, count, orth2, tolerance,
                                                                   limits, active, esc, st);
  return hipGetLastError();
}
hipError_t launch_convergence(SolveState* st, const double* norms, int* active, int* save,
                              uint32_t count, uint32_t step, hipStream_t stream) {
  if (!count) return hipSuccess;
  convergence_kernel<<<(count + kT - 1) / kT, kT, 0, stream>>>(st, norms, active, save, count,
                                                              step);
  return hipGetLastError();
}
hipError_t launch_record_orthogonality(const Segment* segs, uint32_t count, const double* orth2,
                                       SolveState* st, hipStream_t stream) {
  if (!count) return hipSuccess;
  record_orthogonality_kernel<<<(count + kT - 1) / kT, kT, 0, stream>>>(segs, count, orth2, st);
  return hipGetLastError();
}
hipError_t launch_reject_spectrum(const Segment* segs, uint32_t count, const int* rejected,
                                  SolveState* st, int* active, hipStream_t stream) {
  if (!count) return hipSuccess;
  reject_spectrum_kernel<<<(count + kT - 1) / kT, kT, 0, stream>>>(segs, count, rejected, st,
                                                                  active);
  return hipGetLastError();
}
hipError_t launch_frozen_roots(const Segment* segs, uint32_t count, const float* qroots,
                               const int* rejected, SolveState* st, hipStream_t stream) {
  if (!count) return hipSuccess;
  frozen_roots_kernel<<<(count + kT - 1) / kT, kT, 0, stream>>>(segs, count, qroots, rejected, st);
  return hipGetLastError();
}
hipError_t launch_copy_tiles_f64(const Segment* segs, uint32_t count, uint32_t max_ld,
                                 const double* src, double* dst, const int* flag,
                                 hipStream_t stream) {
  if (!count) return hipSuccess;
  copy_tiles_f64_kernel<<<dim3(seg_grid_x(max_ld), count), kT, 0, stream>>>(segs, src, dst, flag);
  return hipGetLastError();
}
hipError_t launch_add_symmetrize_f64(const Segment* segs, uint32_t count, uint32_t max_ld,
                                     double* x, const float* d, const int* active,
                                     hipStream_t stream) {
  if (!count) return hipSuccess;
  add_symmetrize_f64_kernel<<<dim3(seg_grid_x(max_ld), count), kT, 0, stream>>>(segs, x, d,
                                                                               active);
  return hipGetLastError();
}
hipError_t launch_sylvester_divide(const Segment* segs, uint32_t count, uint32_t max_ld,
                                   const float* in, const float* roots, float* out,
                                   const int* active, hipStream_t stream) {
  if (!count) return hipSuccess;
  sylvester_divide_kernel<<<dim3(seg_grid_x(max_ld), count), kT, 0, stream>>>(segs, in, roots, out,
                                                                            active);
  return hipGetLastError();
}
hipError_t launch_scale_columns(const Segment* segs, uint32_t count, uint32_t max_ld,
                                const float* in, const float* roots, float* out,
                                const int* active, hipStream_t stream) {
  if (!count) return hipSuccess;
  scale_columns_kernel<<<dim3(seg_grid_x(max_ld), count), kT, 0, stream>>>(segs, in, roots, out,
                                                                         active);
  return hipGetLastError();
}
hipError_t launch_copy_words(const void* src, void* dst, size_t words, hipStream_t stream) {
  if (!words) return hipSuccess;
  size_t g = (words + kT - 1) / kT;
  if (g > 16384) g = 16384;
  copy_words_kernel<<<static_cast<int>(g), kT, 0, stream>>>(static_cast<const uint32_t*>(src),
                                                           static_cast<uint32_t*>(dst), words);
  return hipGetLastError();
}
hipError_t launch_widen_f64(const float* src, double* dst, size_t n, hipStream_t stream) {
  if (!n) return hipSuccess;
  size_t g = (n + kT - 1) / kT;
  if (g > 16384) g = 16384;
  widen_f64_kernel<<<static_cast<int>(g), kT, 0, stream>>>(src, dst, n);
  return hipGetLastError();
}
hipError_t launch_move_tiles(const Segment* from, const Segment* to, uint32_t count,
                             uint32_t max_ld, const void* src, void* dst, size_t elem_bytes,
                             hipStream_t stream) {
  if (!count) return hipSuccess;
  const dim3 grid(seg_grid_x(max_ld), count);
  if (elem_bytes == 8)
    move_tiles_kernel<double><<<grid, kT, 0, stream>>>(from, to, static_cast<const double*>(src),
                                                      static_cast<double*>(dst));
  else if (elem_bytes == 4)
    move_tiles_kernel<float><<<grid, kT, 0, stream>>>(from, to, static_cast<const float*>(src),
                                                     static_cast<float*>(dst));
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}
hipError_t launch_move_vectors(const Segment* from, const Segment* to, uint32_t count,
                               uint32_t max_ld, const float* src, float* dst,
                               hipStream_t stream) {
  if (!count) return hipSuccess;
  move_vectors_kernel<<<dim3((max_ld + kT - 1) / kT, count), kT, 0, stream>>>(from, to, src, dst);
  return hipGetLastError();
}

}  // namespace msq
