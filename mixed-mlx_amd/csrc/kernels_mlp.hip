// Learned tier selector: per-matrix features and a small MLP (forward,
// backward, Adam) trained on the GPU.
//
// No counterpart in the reference.  The selector reads cheap features of each
// SPD matrix -- order, Gershgorin bounds, diagonal dominance, power-iteration
// estimates of lambda_max and lambda_min -- and predicts, for each of the
// solver configurations it chooses between, whether that configuration meets
// the accuracy target.
//
// Network: 16 inputs -> 64 (ReLU) -> 64 (ReLU) -> 8 logits, float32.
// Parameter vector (5768 floats): W1[64][16], b1[64], W2[64][64], b2[64],
// W3[8][64], b3[8], each weight matrix row-major (W[j][i] multiplies input i
// into output j).
// Loss: masked binary cross-entropy with logits, averaged over the mask.
#include <hip/hip_runtime.h>

#include <math.h>

#include "device_codecs.hpp"
#include "kernels.hpp"

namespace msq {
namespace {

constexpr int kIn = kMlpInputs, kHid = kMlpHidden, kOut = kMlpOutputs;
constexpr int kW1 = 0, kB1 = kW1 + kHid * kIn, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid,
              kW3 = kB2 + kHid, kB3 = kW3 + kOut * kHid, kParams = kB3 + kOut;
static_assert(kParams == kMlpParams, "parameter layout");

// ------------------------------------------------------------------ features
// One workgroup per matrix.  Pass 1 reads A (binary64) once for the
// diagonal, Gershgorin, dominance and norm features.  The power iterations
// run on A rounded to float (an A-perturbation of 2^-24 relative, far below
// what the features resolve) with binary64 accumulation:
//  * orders up to 256 (features_regs_kernel): A lives in registers, four
//    lanes per row (lane p of row i holds columns [p C, p C + C), C = ROWS /
//    4), so the 40 steps never touch memory again; one instantiation per
//    order range (1..64, 65..128, 129..192, 193..256) with 4 ROWS lanes;
//  * orders 257..512 (features_global_kernel): one row per lane, A re-read
//    from memory each step.
// Each launch runs over its range's own segment list (a workgroup given a
// matrix outside its range returns at once).
constexpr int kFeatGlobalThreads = 512;
constexpr int kStats = 13;

// Statistic s of the per-row features: 0..6 and 12 are sums (trace, ||A||_F^2,
// off-diagonal ||.||_F^2, sum and sum of squares of log10 d, dominance sum,
// weak-row count, start-vector norm^2), 7 / 9 / 11 minima (diagonal,
// Gershgorin low, dominance), 8 / 10 maxima (diagonal, Gershgorin high).
__device__ __forceinline__ double stat_combine(int s, double a, double b) {
  if (s == 7 || s == 9 || s == 11) return fmin(a, b);
  if (s == 8 || s == 10) return fmax(a, b);
  return a + b;
}

// The kStats statistics of a workgroup in a fixed order: a butterfly within
// each wave (each lane combines the same two values, so every lane of a wave
// holds the same result), then thread s < kStats combines statistic s over
// the NW wave partials in wave order into out[s] (visible after the caller's
// next barrier).
template <int NW>
__device__ void reduce_stats(double v[kStats], double* part, double* out) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1)
#pragma unroll
    for (int s = 0; s < kStats; ++s) v[s] = stat_combine(s, v[s], __shfl_xor(v[s], off));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int s = 0; s < kStats; ++s) part[wave * kStats + s] = v[s];
  __syncthreads();
  if (threadIdx.x < kStats) {
    const int s = threadIdx.x;
    double acc = part[s];
    for (int w = 1; w < NW; ++w) acc = stat_combine(s, acc, part[w * kStats + s]);
    out[s] = acc;
  }
}

// Per-row statistics of row d (diagonal), r (off-diagonal absolute row sum)
// into v (neutral elements elsewhere); `own`: this lane reports the row.
__device__ __forceinline__ void row_stats(bool own, double d, double r, double fro, double off,
                                          double v[kStats]) {
#pragma unroll
  for (int s = 0; s < kStats; ++s) v[s] = 0.0;
  v[7] = v[9] = v[11] = INFINITY;
  v[8] = v[10] = -INFINITY;
  v[1] = fro;
  v[2] = off;
  if (own) {
    v[0] = d;
    v[7] = v[8] = d;
    v[9] = d - r;
    v[10] = d + r;
    const double ld10 = log10(fmax(d, 1e-300));
    v[3] = ld10;
    v[4] = ld10 * ld10;
    const double dom = fmin(fmax(d / fmax(r, 1e-300), 1e-6), 1e6);
    v[5] = dom;
    v[11] = dom;
    v[6] = d < r ? 1.0 : 0.0;
  }
}

// Sum over the four lanes of a row (lanes 4 r .. 4 r + 3), by DPP quad
// permutations (lane ^ 1, then lane ^ 2): each lane adds the same two values,
// so all four hold the same sum.
__device__ __forceinline__ double dpp_quad(double v, int ctrl_xor1) {
  const int lo = __double2loint(v), hi = __double2hiint(v);
  const int plo = ctrl_xor1 ? __builtin_amdgcn_mov_dpp(lo, 0xb1, 0xf, 0xf, false)
                            : __builtin_amdgcn_mov_dpp(lo, 0x4e, 0xf, 0xf, false);
  const int phi = ctrl_xor1 ? __builtin_amdgcn_mov_dpp(hi, 0xb1, 0xf, 0xf, false)
                            : __builtin_amdgcn_mov_dpp(hi, 0x4e, 0xf, 0xf, false);
  return __hiloint2double(phi, plo);
}
__device__ __forceinline__ double quad_sum(double v) {
  v += dpp_quad(v, 1);  // quad_perm [1, 0, 3, 2]
  v += dpp_quad(v, 0);  // quad_perm [2, 3, 0, 1]
  return v;
}

// Start vector of the lambda_min iteration: a fixed hash of the index in
// [-0.5, 0.5) (any vector with a component along the bottom eigenvector).
__device__ __forceinline__ double start_vector(int j) {
  const uint32_t h = static_cast<uint32_t>(j + 1) * 2654435761u;
  return static_cast<double>((h >> 16) & 0xffffu) / 65536.0 - 0.5;
}

// The 16 features from the reduced statistics and the power iterations'
// last two Rayleigh quotients of each phase (thread 0).
__device__ void write_features(const double* st, int n, double lmax, double lmax_prev, double mu,
                               double mu_prev, double c, float* f) {
  const double tr = st[0], fro = st[1], off = st[2], logd = st[3], logd2 = st[4];
  const double domsum = st[5], weak = st[6], dmin = st[7], dmax = st[8], gl = st[9];
  const double gh = st[10], dommin = st[11];
  const double lmin = c - mu;
  const double t = tr / n;
  const double lmax_c = fmax(lmax, 1e-300);
  const double lmin_c = fmax(lmin, lmax_c * 1e-16);
  const double mean_logd = logd / n;
  f[0] = static_cast<float>(n / 256.0);
  f[1] = static_cast<float>(log2(static_cast<double>(n)) / 8.0);
  f[2] = static_cast<float>(log10(lmax_c / lmin_c));
  f[3] = static_cast<float>(log10(lmax_c / t));
  f[4] = static_cast<float>(log10(lmin_c / t));
  f[5] = static_cast<float>(fmin(fmax(gl / lmax_c, -16.0), 1.0));
  f[6] = static_cast<float>(log10(fmax(gh, 1e-300) / lmax_c));
  f[7] = static_cast<float>(log10(fmax(dmin, 1e-300) / fmax(dmax, 1e-300)));
  f[8] = static_cast<float>(sqrt(fmax(logd2 / n - mean_logd * mean_logd, 0.0)));
  f[9] = static_cast<float>(log10(dommin));
  f[10] = static_cast<float>(log10(domsum / n));
  f[11] = static_cast<float>(fro > 0.0 ? sqrt(off / fro) : 0.0);
  f[12] = static_cast<float>(log10(fmax(sqrt(fro) * sqrt(static_cast<double>(n)) / tr, 1e-300)));
  f[13] = static_cast<float>(log10(fabs(lmax - lmax_prev) / lmax_c + 1e-30));
  f[14] = static_cast<float>(log10(fabs(mu - mu_prev) / fmax(fabs(mu), 1e-300) + 1e-30));
  f[15] = static_cast<float>(weak / n);
}

// Orders (ROWS - 64, ROWS], A (binary64, column-major, pitch ld) held in
// registers as float.  The power steps keep the vector unnormalised in LDS
// (double-buffered) with its scale s = 1 / ||y|| in registers: step k reads
// buffer k mod 2, forms y = s A b (or s (c b - A b) for lambda_min), writes y
// and the wave partials of x^T y and y^T y to the other buffer, and meets the
// other waves at one barrier; every lane then sums the NW partials in wave
// order.  The last lambda_max step writes the start vector instead of y.
// For orders 193..256 the lane's last C - R columns sit in LDS instead
// (float4 k-major per lane, conflict-free): 64 register columns per lane do
// not fit the 128 registers of a 1024-lane workgroup beside the rest.
template <int ROWS>
__global__ __launch_bounds__(4 * ROWS) void features_regs_kernel(
    const Segment* __restrict__ segs, const double* __restrict__ a64,
    float* __restrict__ features) {
  constexpr int T = 4 * ROWS;
  constexpr int NW = T / 64;
  constexpr int C = ROWS / 4;                 // columns per lane
  constexpr int R = ROWS == 256 ? 40 : C;     // of which in registers
  constexpr int LQ = (C - R) / 4;             // float4 groups in LDS
  static_assert(C % 8 == 0 && R % 8 == 0, "chunks of 8 columns");
  // The vector b, double-buffered, each lane-phase's C columns padded by two
  // doubles (16 bytes) so that the four phases' broadcast reads of a wave
  // fall in different LDS banks: column j at j + 2 (j / C).
  constexpr int XP = ROWS + 8;
  __shared__ __align__(16) double xb[2][XP];
  __shared__ float4 al[LQ > 0 ? LQ : 1][LQ > 0 ? T : 1];
  __shared__ double pw[2][NW][2];
  __shared__ double part[NW * kStats];
  __shared__ double stats[kStats];
  const Segment sg = segs[blockIdx.x];
  const int n = static_cast<int>(sg.order);
  if (n > ROWS || n <= ROWS - 64) return;
  const int ld = static_cast<int>(sg.ld);
  const double* A = a64 + sg.offset;
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int row = tid >> 2, phase = tid & 3, col0 = phase * C;
  const bool row_ok = row < n;

  // Pass 1: this lane's columns of its row, rounded to float for the power
  // steps; diagonal, off-diagonal absolute row sum, Frobenius norms.  Chunks
  // of 8 columns, kept apart (sched_barrier) so that the loads in flight stay
  // within the register budget.
  float a[R];
  double diag = 0.0, rsum = 0.0, fro = 0.0, off = 0.0;
#pragma unroll
  for (int k0 = 0; k0 < C; k0 += 8) {
    float t[8];
    // Unconditional in-bounds loads (clamped indices) of the chunk, all
    // issued before one empty asm that keeps them out of a branch (the
    // selects below zero what lies outside the order).
    double raw[8];
#pragma unroll
    for (int k = 0; k < 8; ++k)
      raw[k] = A[min(row, n - 1) + static_cast<size_t>(min(col0 + k0 + k, n - 1)) * ld];
    asm volatile("" : "+v"(raw[0]), "+v"(raw[1]), "+v"(raw[2]), "+v"(raw[3]), "+v"(raw[4]),
                 "+v"(raw[5]), "+v"(raw[6]), "+v"(raw[7]));
#pragma unroll
    for (int k = k0; k < k0 + 8; ++k) {
      const int j = col0 + k;
      const double v = (row_ok && j < n) ? raw[k - k0] : 0.0;
      // The float is formed here (pinned), so the binary64 value dies.
      float f = static_cast<float>(v);
      asm volatile("" : "+v"(f));
      if (k < R) {
        a[k < R ? k : 0] = f;
      } else {
        t[k - k0] = f;
      }
      fro += v * v;
      if (j == row) {
        diag = v;
      } else {
        rsum += fabs(v);
        off += v * v;
      }
    }
    if (k0 >= R) {
      al[(k0 - R) / 4][tid] = make_float4(t[0], t[1], t[2], t[3]);
      al[(k0 - R) / 4 + 1][tid] = make_float4(t[4], t[5], t[6], t[7]);
    }
    asm volatile("" : "+v"(diag), "+v"(rsum), "+v"(fro), "+v"(off));
    __builtin_amdgcn_sched_barrier(0);
  }
  diag = quad_sum(diag);
  rsum = quad_sum(rsum);
  // Pinned here (empty asm): the statistics are reduced after the power
  // steps, and without the pin their sums sink there too, keeping every
  // loaded binary64 value alive through the loop.
  asm volatile("" : "+v"(diag), "+v"(rsum), "+v"(fro), "+v"(off));
  const bool own = row_ok && phase == 0;
  // The start vector's squared norm now (its wave partials, summed by every
  // lane when the lambda_min phase starts); the other statistics after the
  // power steps, when A's registers are free.
  {
    double sv = own ? start_vector(row) * start_vector(row) : 0.0;
#pragma unroll
    for (int o = 32; o >= 4; o >>= 1) sv += __shfl_xor(sv, o);
    if (lane == 0) part[wave * kStats + 12] = sv;
  }

  // Power iterations (x = s b, b the buffer): kPowerIterationsMax steps with
  // y = A x from x = 1 / sqrt(n), then kPowerIterationsMin steps with
  // y = c x - A x, c just above lambda_max, from the start vector.  Beyond
  // the order the buffers stay zero.
  for (int i = tid; i < XP; i += T) {
    const int j = i - 2 * (i / (C + 2));  // the column padded index i holds
    const bool col = i % (C + 2) < C;
    xb[0][i] = col && j < n ? 1.0 : 0.0;
    xb[1][i] = 0.0;
  }
  __syncthreads();
  double s = 1.0 / sqrt(static_cast<double>(n));
  double c = 0.0, lmax = 0.0, lmax_prev = 0.0, mu = 0.0, mu_prev = 0.0;
  int cur = 0;
  const int total_iters = kPowerIterationsMax + kPowerIterationsMin;
#pragma unroll 1
  for (int it = 0; it < total_iters; ++it) {
    // The float copy of A is pinned for each step (empty asm, no
    // instruction): otherwise its widening to binary64 is hoisted out of the
    // loop and the doubled copy no longer fits the register file.
#pragma unroll
    for (int k = 0; k < R; ++k) asm volatile("" : "+v"(a[k]));
    const double* xc = xb[cur];
    double zs[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k0 = 0; k0 < C; k0 += 8) {
      if (col0 + k0 < n) {
        float w[8];
        if (k0 < R) {
#pragma unroll
          for (int h = 0; h < 8; ++h) w[h] = a[k0 < R ? k0 + h : 0];
        } else {
          const float4 lo = al[(k0 - R) / 4][tid], hi = al[(k0 - R) / 4 + 1][tid];
          w[0] = lo.x; w[1] = lo.y; w[2] = lo.z; w[3] = lo.w;
          w[4] = hi.x; w[5] = hi.y; w[6] = hi.z; w[7] = hi.w;
        }
        const double2* xv = reinterpret_cast<const double2*>(xc + col0 + 2 * phase + k0);
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          const double2 x2 = xv[h];
          zs[(2 * h) & 3] += static_cast<double>(w[2 * h]) * x2.x;
          zs[(2 * h + 1) & 3] += static_cast<double>(w[2 * h + 1]) * x2.y;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    const double z = quad_sum((zs[0] + zs[1]) + (zs[2] + zs[3]));
    const int prow = row + 2 * (row / C);
    const double br = row_ok ? xc[prow] : 0.0;
    const double y = c != 0.0 ? s * (c * br - z) : s * z;
    double q = 0.0, yy = 0.0;
    if (row_ok && phase == 0) {
      q = (s * br) * y;
      yy = y * y;
    }
    // Only lanes 4 r carry values: the butterfly over offsets 32 .. 4.
#pragma unroll
    for (int o = 32; o >= 4; o >>= 1) {
      q += __shfl_xor(q, o);
      yy += __shfl_xor(yy, o);
    }
    const int nxt = cur ^ 1;
    if (row_ok && phase == 0)
      xb[nxt][prow] = it == kPowerIterationsMax - 1 ? start_vector(row) : y;
    if (lane == 0) {
      pw[nxt][wave][0] = q;
      pw[nxt][wave][1] = yy;
    }
    __syncthreads();
    double qs = 0.0, ys = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      qs += pw[nxt][w][0];
      ys += pw[nxt][w][1];
    }
    if (it < kPowerIterationsMax) {
      lmax_prev = lmax;
      lmax = qs;
    } else {
      mu_prev = mu;
      mu = qs;
    }
    if (it == kPowerIterationsMax - 1) {
      c = 1.0001 * fmax(lmax, 1e-300);
      double sv = 0.0;
      for (int w = 0; w < NW; ++w) sv += part[w * kStats + 12];
      s = 1.0 / sqrt(sv);
    } else {
      s = ys > 0.0 ? 1.0 / sqrt(ys) : 0.0;
    }
    cur = nxt;
  }
  {
    double v[kStats];
    row_stats(own, diag, rsum, fro, off, v);
    reduce_stats<NW>(v, part, stats);
  }
  __syncthreads();
  if (tid == 0)
    write_features(stats, n, lmax, lmax_prev, mu, mu_prev, c,
                   features + static_cast<size_t>(sg.index) * kIn);
}

// Block-wide sums in a fixed order (a butterfly within each wave, then a
// butterfly over the NW wave partials, which every lane reads back).
template <int NW>
__device__ void block_sum2(double& a, double& b, double* red) {
  static_assert(NW >= 1 && NW <= 64 && (NW & (NW - 1)) == 0, "waves: a power of two");
  for (int off = 32; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) {
    red[2 * wave] = a;
    red[2 * wave + 1] = b;
  }
  __syncthreads();
  const int w = lane & (NW - 1);
  a = red[2 * w];
  b = red[2 * w + 1];
  for (int off = NW / 2; off > 0; off >>= 1) {
    a += __shfl_xor(a, off);
    b += __shfl_xor(b, off);
  }
}

// Orders 257..512: one row per lane, A re-read from memory each step.
__global__ __launch_bounds__(kFeatGlobalThreads) void features_global_kernel(
    const Segment* __restrict__ segs, const double* __restrict__ a64,
    float* __restrict__ features) {
  constexpr int T = kFeatGlobalThreads;
  constexpr int NW = T / 64;
  __shared__ double x[kMaxFeatureOrder];  // power-step vector
  __shared__ double red[2 * NW];
  __shared__ double part[NW * kStats];
  __shared__ double stats[kStats];
  const Segment sg = segs[blockIdx.x];
  const int n = static_cast<int>(sg.order);
  if (n <= 256 || n > kMaxFeatureOrder) return;
  const int ld = static_cast<int>(sg.ld);
  const double* A = a64 + sg.offset;
  const int tid = threadIdx.x;
  const int row = tid;
  const bool row_ok = row < n;

  double diag = 0.0, rsum = 0.0, fro = 0.0, off = 0.0;
  for (int j = 0; j < n && row_ok; ++j) {
    const double v = A[row + static_cast<size_t>(j) * ld];
    fro += v * v;
    if (j == row) {
      diag = v;
    } else {
      rsum += fabs(v);
      off += v * v;
    }
  }
  {
    double v[kStats];
    row_stats(row_ok, diag, rsum, fro, off, v);
    v[12] = 0.0;
    reduce_stats<NW>(v, part, stats);
  }

  const double inv_sqrt_n = 1.0 / sqrt(static_cast<double>(n));
  for (int i = tid; i < kMaxFeatureOrder; i += T) x[i] = i < n ? inv_sqrt_n : 0.0;
  __syncthreads();
  double lmax = 0.0, lmax_prev = 0.0, mu = 0.0, mu_prev = 0.0, c = 0.0;
  const int total_iters = kPowerIterationsMax + kPowerIterationsMin;
#pragma unroll 1
  for (int it = 0; it < total_iters; ++it) {
    // Every read of x in a step precedes the barriers of its reduction, so
    // the new x is written in place after it.
    if (it == kPowerIterationsMax) {
      c = 1.0001 * fmax(lmax, 1e-300);
      double snorm = 0.0, unused = 0.0;
      for (int i = tid; i < n; i += T) snorm += start_vector(i) * start_vector(i);
      block_sum2<NW>(snorm, unused, red);
      for (int i = tid; i < n; i += T) x[i] = start_vector(i) / sqrt(snorm);
      __syncthreads();
    }
    double y = 0.0;
    if (row_ok) {
      for (int j = 0; j < n; ++j)
        y += static_cast<double>(static_cast<float>(A[row + static_cast<size_t>(j) * ld])) * x[j];
    }
    double q = 0.0, yy = 0.0;
    if (row_ok) {
      if (c != 0.0) y = c * x[row] - y;
      q = x[row] * y;
      yy = y * y;
    }
    block_sum2<NW>(q, yy, red);
    if (it < kPowerIterationsMax) {
      lmax_prev = lmax;
      lmax = q;
    } else {
      mu_prev = mu;
      mu = q;
    }
    const double inv = yy > 0.0 ? 1.0 / sqrt(yy) : 0.0;
    if (row_ok) x[row] = y * inv;
    __syncthreads();
  }
  if (tid == 0)
    write_features(stats, n, lmax, lmax_prev, mu, mu_prev, c,
                   features + static_cast<size_t>(sg.index) * kIn);
}

// ----------------------------------------------------------------------- MLP
constexpr int kMlpThreads = 256;

__device__ __forceinline__ void stage_params(const float* __restrict__ params, float* p) {
  for (int e = threadIdx.x; e < kParams; e += blockDim.x) p[e] = params[e];
  __syncthreads();
}

// One thread per sample.  h1 / h2 (post-ReLU) are stored when non-null.
__global__ __launch_bounds__(kMlpThreads) void mlp_forward_kernel(
    const float* __restrict__ params, const float* __restrict__ x, int batch,
    float* __restrict__ logits, float* __restrict__ h1_out, float* __restrict__ h2_out) {
  __shared__ float p[kParams];
  stage_params(params, p);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  float xi[kIn];
#pragma unroll
  for (int i = 0; i < kIn; ++i) xi[i] = x[static_cast<size_t>(b) * kIn + i];
  float h1[kHid];
#pragma unroll
  for (int j = 0; j < kHid; ++j) {
    float acc = p[kB1 + j];
#pragma unroll
    for (int i = 0; i < kIn; ++i) acc = fmaf(p[kW1 + j * kIn + i], xi[i], acc);
    h1[j] = fmaxf(acc, 0.0f);
    if (h1_out) h1_out[static_cast<size_t>(b) * kHid + j] = h1[j];
  }
  float z[kOut];
#pragma unroll
  for (int k = 0; k < kOut; ++k) z[k] = p[kB3 + k];
  for (int j = 0; j < kHid; ++j) {
    float acc = p[kB2 + j];
#pragma unroll
    for (int i = 0; i < kHid; ++i) acc = fmaf(p[kW2 + j * kHid + i], h1[i], acc);
    const float h2 = fmaxf(acc, 0.0f);
    if (h2_out) h2_out[static_cast<size_t>(b) * kHid + j] = h2;
#pragma unroll
    for (int k = 0; k < kOut; ++k) z[k] = fmaf(p[kW3 + k * kHid + j], h2, z[k]);
  }
#pragma unroll
  for (int k = 0; k < kOut; ++k) logits[static_cast<size_t>(b) * kOut + k] = z[k];
}

// BF16 inference: weights stored in BF16, inputs and hidden activations
// rounded to BF16 (round to nearest even, after the ReLU), products
// accumulated in float with fused multiply-adds in index order (a product of
// two BF16 values is exact in float, so each step rounds once), logits in
// float.  One thread per sample, the same order of operations as
// mlp_forward_kernel.
__global__ __launch_bounds__(kMlpThreads) void mlp_forward_bf16_kernel(
    const uint16_t* __restrict__ params, const float* __restrict__ x, int batch,
    float* __restrict__ logits) {
  __shared__ float p[kParams];
  for (int e = threadIdx.x; e < kParams; e += blockDim.x) p[e] = decode_bf16(params[e]);
  __syncthreads();
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  float xi[kIn];
#pragma unroll
  for (int i = 0; i < kIn; ++i)
    xi[i] = decode_bf16(encode_bf16(x[static_cast<size_t>(b) * kIn + i]));
  float h1[kHid];
#pragma unroll
  for (int j = 0; j < kHid; ++j) {
    float acc = p[kB1 + j];
#pragma unroll
    for (int i = 0; i < kIn; ++i) acc = fmaf(p[kW1 + j * kIn + i], xi[i], acc);
    h1[j] = decode_bf16(encode_bf16(fmaxf(acc, 0.0f)));
  }
  float z[kOut];
#pragma unroll
  for (int k = 0; k < kOut; ++k) z[k] = p[kB3 + k];
  for (int j = 0; j < kHid; ++j) {
    float acc = p[kB2 + j];
#pragma unroll
    for (int i = 0; i < kHid; ++i) acc = fmaf(p[kW2 + j * kHid + i], h1[i], acc);
    const float h2 = decode_bf16(encode_bf16(fmaxf(acc, 0.0f)));
#pragma unroll
    for (int k = 0; k < kOut; ++k) z[k] = fmaf(p[kW3 + k * kHid + j], h2, z[k]);
  }
#pragma unroll
  for (int k = 0; k < kOut; ++k) logits[static_cast<size_t>(b) * kOut + k] = z[k];
}

// Per-sample deltas: dz = mask (sigmoid(z) - y) / mask_total,
// dh2 = (W3^T dz) o [h2 > 0], dh1 = (W2^T dh2) o [h1 > 0]; and the sample's
// loss sum_k mask (softplus(z) - y z).
__global__ __launch_bounds__(kMlpThreads) void mlp_delta_kernel(
    const float* __restrict__ params, const float* __restrict__ logits,
    const float* __restrict__ h1, const float* __restrict__ h2,
    const float* __restrict__ targets, const float* __restrict__ mask, int batch,
    float inv_total, float* __restrict__ dz_out, float* __restrict__ dh1_out,
    float* __restrict__ dh2_out, float* __restrict__ loss_out) {
  __shared__ float p[kParams];
  stage_params(params, p);
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  float dz[kOut];
  float loss = 0.0f;
#pragma unroll
  for (int k = 0; k < kOut; ++k) {
    const size_t e = static_cast<size_t>(b) * kOut + k;
    const float z = logits[e], y = targets[e], m = mask[e];
    const float sig = 1.0f / (1.0f + expf(-z));
    dz[k] = m * (sig - y) * inv_total;
    // softplus(z) - y z, evaluated without overflow
    loss += m * (fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z))) - y * z);
    dz_out[e] = dz[k];
  }
  loss_out[b] = loss;
  float dh2[kHid];
#pragma unroll
  for (int j = 0; j < kHid; ++j) {
    float acc = 0.0f;
#pragma unroll
    for (int k = 0; k < kOut; ++k) acc = fmaf(p[kW3 + k * kHid + j], dz[k], acc);
    dh2[j] = h2[static_cast<size_t>(b) * kHid + j] > 0.0f ? acc : 0.0f;
    dh2_out[static_cast<size_t>(b) * kHid + j] = dh2[j];
  }
  for (int i = 0; i < kHid; ++i) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < kHid; ++j) acc = fmaf(p[kW2 + j * kHid + i], dh2[j], acc);
    dh1_out[static_cast<size_t>(b) * kHid + i] = h1[static_cast<size_t>(b) * kHid + i] > 0.0f ? acc : 0.0f;
  }
}

// Parameter gradients, one thread per parameter, summing over the batch in
// sample order (deterministic).
__global__ __launch_bounds__(kMlpThreads) void mlp_grad_kernel(
    const float* __restrict__ x, const float* __restrict__ h1, const float* __restrict__ h2,
    const float* __restrict__ dz, const float* __restrict__ dh1, const float* __restrict__ dh2,
    int batch, float* __restrict__ grads) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kParams) return;
  const float* delta;  // per-sample output delta, row stride dstride
  const float* act;    // per-sample input activation (null for a bias)
  int dstride, astride, j, i;
  if (q < kB1) {
    j = (q - kW1) / kIn; i = (q - kW1) % kIn;
    delta = dh1; dstride = kHid; act = x; astride = kIn;
  } else if (q < kW2) {
    j = q - kB1; i = 0;
    delta = dh1; dstride = kHid; act = nullptr; astride = 0;
  } else if (q < kB2) {
    j = (q - kW2) / kHid; i = (q - kW2) % kHid;
    delta = dh2; dstride = kHid; act = h1; astride = kHid;
  } else if (q < kW3) {
    j = q - kB2; i = 0;
    delta = dh2; dstride = kHid; act = nullptr; astride = 0;
  } else if (q < kB3) {
    j = (q - kW3) / kHid; i = (q - kW3) % kHid;
    delta = dz; dstride = kOut; act = h2; astride = kHid;
  } else {
    j = q - kB3; i = 0;
    delta = dz; dstride = kOut; act = nullptr; astride = 0;
  }
  float g = 0.0f;
  if (act) {
    for (int b = 0; b < batch; ++b)
      g = fmaf(delta[static_cast<size_t>(b) * dstride + j], act[static_cast<size_t>(b) * astride + i], g);
  } else {
    for (int b = 0; b < batch; ++b) g += delta[static_cast<size_t>(b) * dstride + j];
  }
  grads[q] = g;
}

__global__ void mlp_adam_kernel(float* __restrict__ params, const float* __restrict__ grads,
                                float* __restrict__ m1, float* __restrict__ m2, int count,
                                float lr, float beta1, float beta2, float eps, float decay,
                                float corr1, float corr2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  const float g = grads[q] + decay * params[q];
  const float a = beta1 * m1[q] + (1.0f - beta1) * g;
  const float v = beta2 * m2[q] + (1.0f - beta2) * g * g;
  m1[q] = a;
  m2[q] = v;
  params[q] -= lr * (a / corr1) / (sqrtf(v / corr2) + eps);
}

}  // namespace

int feature_rows_for_order(uint32_t order) {
  return order <= 64 ? 64 : order <= 128 ? 128 : order <= 192 ? 192 : order <= 256 ? 256 : 0;
}

hipError_t launch_features(int rows, const Segment* segs, uint32_t count, const double* a64,
                           float* features, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  switch (rows) {
    case 64: features_regs_kernel<64><<<count, 4 * 64, 0, stream>>>(segs, a64, features); break;
    case 128: features_regs_kernel<128><<<count, 4 * 128, 0, stream>>>(segs, a64, features); break;
    case 192: features_regs_kernel<192><<<count, 4 * 192, 0, stream>>>(segs, a64, features); break;
    case 256: features_regs_kernel<256><<<count, 4 * 256, 0, stream>>>(segs, a64, features); break;
    case 0:
      features_global_kernel<<<count, kFeatGlobalThreads, 0, stream>>>(segs, a64, features);
      break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_mlp_forward(const float* params, const float* x, int batch, float* logits,
                              float* h1, float* h2, hipStream_t stream) {
  if (batch <= 0) return hipSuccess;
  mlp_forward_kernel<<<(batch + kMlpThreads - 1) / kMlpThreads, kMlpThreads, 0, stream>>>(
      params, x, batch, logits, h1, h2);
  return hipGetLastError();
}

hipError_t launch_mlp_forward_bf16(const uint16_t* params, const float* x, int batch,
                                   float* logits, hipStream_t stream) {
  if (batch <= 0) return hipSuccess;
  mlp_forward_bf16_kernel<<<(batch + kMlpThreads - 1) / kMlpThreads, kMlpThreads, 0, stream>>>(
      params, x, batch, logits);
  return hipGetLastError();
}

hipError_t launch_mlp_backward(const float* params, const float* x, const float* h1,
                               const float* h2, const float* logits, const float* targets,
                               const float* mask, int batch, float mask_total, float* dz,
                               float* dh1, float* dh2, float* losses, float* grads,
                               hipStream_t stream) {
  if (batch <= 0) return hipSuccess;
  const float inv = mask_total > 0.0f ? 1.0f / mask_total : 0.0f;
  mlp_delta_kernel<<<(batch + kMlpThreads - 1) / kMlpThreads, kMlpThreads, 0, stream>>>(
      params, logits, h1, h2, targets, mask, batch, inv, dz, dh1, dh2, losses);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  mlp_grad_kernel<<<(kParams + kMlpThreads - 1) / kMlpThreads, kMlpThreads, 0, stream>>>(
      x, h1, h2, dz, dh1, dh2, batch, grads);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float* params, const float* grads, float* m1, float* m2, int count,
                           float lr, float beta1, float beta2, float eps, float decay, int step,
                           hipStream_t stream) {
  if (count <= 0) return hipSuccess;
  const float corr1 = 1.0f - powf(beta1, static_cast<float>(step));
  const float corr2 = 1.0f - powf(beta2, static_cast<float>(step));
  mlp_adam_kernel<<<(count + 255) / 256, 256, 0, stream>>>(params, grads, m1, m2, count, lr,
                                                          beta1, beta2, eps, decay, corr1, corr2);
  return hipGetLastError();
}

}  // namespace msq
