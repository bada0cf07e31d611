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
//  * orders up to 256: A lives in registers, four lanes per row (lane p of
//    row i holds columns 4 k + p; two rows per lane for orders 129..256),
//    so the 40 steps never touch memory again;
//    one instantiation per order range (1..64, 65..128, 129..256) with
//    4 ROWS lanes, so small matrices take small workgroups;
//  * orders 257..512: one row per lane, A re-read from memory each step.
// Each instantiation runs over its range's own segment list (a workgroup
// given a matrix outside its range returns at once).
constexpr int kFeatGlobalThreads = 512;

// Register path: rows of A per lane.  Orders 129..256 hold two rows per lane
// (512 lanes): with one row on 1024 lanes, A's 64 floats per lane leave too
// few of the 128 registers a 1024-lane workgroup gets and the rest spills.
__host__ __device__ constexpr int feat_rows_per_lane(int rows) { return rows == 256 ? 2 : 1; }
__host__ __device__ constexpr int feat_threads(int rows) {
  return rows ? 4 * rows / feat_rows_per_lane(rows) : kFeatGlobalThreads;
}

// Block-wide reductions in a fixed order (a butterfly within each wave, then
// a butterfly over the NW wave partials, which every lane reads back): the
// results do not depend on timing, and few values are live at a time.
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
template <int NW>
__device__ double block_sum(double v, double* red) {
  double z = 0.0;
  block_sum2<NW>(v, z, red);
  return v;
}
template <int NW>
__device__ double block_min(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  v = red[lane & (NW - 1)];
  for (int off = NW / 2; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off));
  return v;
}
template <int NW>
__device__ double block_max(double v, double* red) { return -block_min<NW>(-v, red); }
// Sum over the four lanes of a row (lanes 4 r .. 4 r + 3).
__device__ __forceinline__ double quad_sum(double v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  return v;
}

// Start vector of the lambda_min iteration: a fixed hash of the index in
// [-0.5, 0.5) (any vector with a component along the bottom eigenvector).
__device__ __forceinline__ double start_vector(int j) {
  const uint32_t h = static_cast<uint32_t>(j + 1) * 2654435761u;
  return static_cast<double>((h >> 16) & 0xffffu) / 65536.0 - 0.5;
}

// A binary64 column-major (pitch ld), order <= 512.  ROWS: orders
// (ROWS / 2, ROWS] on the register path, or 0 for orders 257..512.
template <int ROWS>
__global__ __launch_bounds__(feat_threads(ROWS)) void features_kernel(
    const Segment* __restrict__ segs, const double* __restrict__ a64,
    float* __restrict__ features) {
  constexpr bool REGS = ROWS != 0;
  constexpr int T = feat_threads(ROWS);
  constexpr int NW = T / 64;
  constexpr int COLS = REGS ? ROWS / 4 : 1;  // register values per lane
  __shared__ double x[kMaxFeatureOrder];  // power-step vector
  __shared__ double red[2 * NW];
  const Segment sg = segs[blockIdx.x];
  const int n = static_cast<int>(sg.order);
  if (REGS ? (n > ROWS || (ROWS > 64 && n <= ROWS / 2)) : n <= 256) return;
  const int ld = static_cast<int>(sg.ld);
  const double* A = a64 + sg.offset;
  const int tid = threadIdx.x;
  // Rows and column phase of this lane: register path (rows tid / 4 + r T / 4,
  // r < RPL, columns 4 k + tid % 4), else row tid over all columns.
  constexpr int RPL = REGS ? feat_rows_per_lane(ROWS) : 1;
  const int row0 = REGS ? tid >> 2 : tid;
  const int phase = REGS ? tid & 3 : 0;
  int rows[RPL];
  bool rows_ok[RPL];
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    rows[r] = row0 + r * (T / 4);
    rows_ok[r] = rows[r] < n;
  }

  // Pass 1: diagonal, off-diagonal absolute row sums, Frobenius norms.
  float a[RPL][COLS];
  double diag[RPL], rsum[RPL];
  double fro = 0.0, off = 0.0;
#pragma unroll
  for (int r = 0; r < RPL; ++r) diag[r] = rsum[r] = 0.0;
  if constexpr (REGS) {
    // Chunks of 8 columns, kept apart (sched_barrier) so that the loads in
    // flight stay within the register budget.
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      const int row = rows[r];
#pragma unroll
      for (int k0 = 0; k0 < COLS; k0 += 8) {
#pragma unroll
        for (int k = k0; k < k0 + 8; ++k) {
          // Unconditional in-bounds load (clamped indices), then a select; the
          // empty asm keeps the load out of a branch (the compiler would
          // otherwise make each load conditional, one basic block per column).
          const int j = 4 * k + phase;
          double raw = A[min(row, n - 1) + static_cast<size_t>(min(j, n - 1)) * ld];
          asm volatile("" : "+v"(raw));
          const double v = (rows_ok[r] && j < n) ? raw : 0.0;
          a[r][k] = static_cast<float>(v);
          fro += v * v;
          if (j == row) {
            diag[r] = v;
          } else {
            rsum[r] += fabs(v);
            off += v * v;
          }
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      diag[r] = quad_sum(diag[r]);
      rsum[r] = quad_sum(rsum[r]);
    }
  } else {
    const int row = rows[0];
    for (int j = 0; j < n && rows_ok[0]; ++j) {
      const double v = A[row + static_cast<size_t>(j) * ld];
      fro += v * v;
      if (j == row) {
        diag[0] = v;
      } else {
        rsum[0] += fabs(v);
        off += v * v;
      }
    }
  }
  // Per-row features from one lane of each row.
  double tr = 0.0, dmin = INFINITY, dmax = -INFINITY, gl = INFINITY, gh = -INFINITY;
  double logd = 0.0, logd2 = 0.0, dommin = INFINITY, domsum = 0.0, weak = 0.0;
#pragma unroll
  for (int r = 0; r < RPL; ++r) {
    if (rows_ok[r] && phase == 0) {
      const double d = diag[r];
      tr += d;
      dmin = fmin(dmin, d);
      dmax = fmax(dmax, d);
      gl = fmin(gl, d - rsum[r]);
      gh = fmax(gh, d + rsum[r]);
      const double ld10 = log10(fmax(d, 1e-300));
      logd += ld10;
      logd2 += ld10 * ld10;
      const double dom = fmin(fmax(d / fmax(rsum[r], 1e-300), 1e-6), 1e6);
      dommin = fmin(dommin, dom);
      domsum += dom;
      weak += d < rsum[r] ? 1.0 : 0.0;
    }
  }
  block_sum2<NW>(tr, fro, red);
  block_sum2<NW>(off, logd, red);
  block_sum2<NW>(logd2, domsum, red);
  weak = block_sum<NW>(weak, red);
  dmin = block_min<NW>(dmin, red);
  dmax = block_max<NW>(dmax, red);
  gl = block_min<NW>(gl, red);
  gh = block_max<NW>(gh, red);
  dommin = block_min<NW>(dommin, red);
  // The reduced statistics are read again only by thread 0 at the end: kept
  // in LDS meanwhile, so that the power steps get the registers (the binary64
  // values would otherwise stay live in every lane).
  __shared__ double stats[12];
  if (tid == 0) {
    stats[0] = tr;
    stats[1] = fro;
    stats[2] = off;
    stats[3] = logd;
    stats[4] = logd2;
    stats[5] = domsum;
    stats[6] = weak;
    stats[7] = dmin;
    stats[8] = dmax;
    stats[9] = gl;
    stats[10] = gh;
    stats[11] = dommin;
  }

  // Power iterations: kPowerIterationsMax steps of x <- y / ||y|| with
  // y = A x from x = 1 / sqrt(n) (lambda_max), then kPowerIterationsMin steps
  // with y = c x - A x, c just above lambda_max, from the start vector (top of
  // c I - A: lambda_min = c - mu).  The last two Rayleigh quotients x^T y of
  // each phase are kept.  x beyond the order stays zero (the register path
  // reads columns up to the next multiple of 4).
  const double inv_sqrt_n = 1.0 / sqrt(static_cast<double>(n));
  for (int i = tid; i < kMaxFeatureOrder; i += T) x[i] = i < n ? inv_sqrt_n : 0.0;
  __syncthreads();
  double lmax = 0.0, lmax_prev = 0.0, mu = 0.0, mu_prev = 0.0, c = 0.0;
  const int total_iters = kPowerIterationsMax + kPowerIterationsMin;
#pragma unroll 1
  for (int it = 0; it < total_iters; ++it) {
    // Every read of x in a step precedes the barriers of its reduction, so
    // the new x is written in place after it.
    double* xc = x;
    if (it == kPowerIterationsMax) {
      c = 1.0001 * fmax(lmax, 1e-300);
      double snorm = 0.0;
      for (int i = tid; i < n; i += T) snorm += start_vector(i) * start_vector(i);
      snorm = block_sum<NW>(snorm, red);
      for (int i = tid; i < n; i += T) xc[i] = start_vector(i) / sqrt(snorm);
      __syncthreads();
    }
    double y[RPL];
    if constexpr (REGS) {
      // The register copy of A is pinned as float for each step (empty asm, no
      // instruction): otherwise its widening to binary64 is hoisted out of the
      // loop and the doubled copy no longer fits the register file.  Four
      // partial sums per row (columns 16 k0 + 4 u + p, u = k mod 4) shorten
      // the dependent chain of binary64 FMAs; each x value read serves every
      // row of the lane.
#pragma unroll
      for (int r = 0; r < RPL; ++r)
#pragma unroll
        for (int k = 0; k < COLS; ++k) asm volatile("" : "+v"(a[r][k]));
      double ys[RPL][4];
#pragma unroll
      for (int r = 0; r < RPL; ++r) ys[r][0] = ys[r][1] = ys[r][2] = ys[r][3] = 0.0;
#pragma unroll
      for (int r = 0; r < RPL; ++r) {
#pragma unroll
        for (int k0 = 0; k0 < COLS; k0 += 8) {
          if (4 * k0 < n) {
#pragma unroll
            for (int k = k0; k < k0 + 8; ++k)
              ys[r][k & 3] += static_cast<double>(a[r][k]) * xc[4 * k + phase];
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
#pragma unroll
      for (int r = 0; r < RPL; ++r)
        y[r] = quad_sum((ys[r][0] + ys[r][1]) + (ys[r][2] + ys[r][3]));
    } else {
      y[0] = 0.0;
      if (rows_ok[0]) {
        for (int j = 0; j < n; ++j)
          y[0] += static_cast<double>(static_cast<float>(A[rows[0] + static_cast<size_t>(j) * ld])) *
                  xc[j];
      }
    }
    double q = 0.0, yy = 0.0;
#pragma unroll
    for (int r = 0; r < RPL; ++r) {
      if (c != 0.0 && rows_ok[r]) y[r] = c * xc[rows[r]] - y[r];
      if (rows_ok[r] && phase == 0) {
        q += xc[rows[r]] * y[r];
        yy += y[r] * y[r];
      }
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
#pragma unroll
    for (int r = 0; r < RPL; ++r)
      if (rows_ok[r] && phase == 0) xc[rows[r]] = y[r] * inv;
    __syncthreads();
  }
  const double lmin = c - mu;

  if (tid == 0) {
    tr = stats[0];
    fro = stats[1];
    off = stats[2];
    logd = stats[3];
    logd2 = stats[4];
    domsum = stats[5];
    weak = stats[6];
    dmin = stats[7];
    dmax = stats[8];
    gl = stats[9];
    gh = stats[10];
    dommin = stats[11];
    const double t = tr / n;
    const double lmax_c = fmax(lmax, 1e-300);
    const double lmin_c = fmax(lmin, lmax_c * 1e-16);
    const double mean_logd = logd / n;
    float* f = features + static_cast<size_t>(sg.index) * kIn;
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
  return order <= 64 ? 64 : order <= 128 ? 128 : order <= 256 ? 256 : 0;
}

hipError_t launch_features(int rows, const Segment* segs, uint32_t count, const double* a64,
                           float* features, hipStream_t stream) {
  if (count == 0) return hipSuccess;
  switch (rows) {
    case 64: features_kernel<64><<<count, feat_threads(64), 0, stream>>>(segs, a64, features); break;
    case 128: features_kernel<128><<<count, feat_threads(128), 0, stream>>>(segs, a64, features); break;
    case 256: features_kernel<256><<<count, feat_threads(256), 0, stream>>>(segs, a64, features); break;
    case 0: features_kernel<0><<<count, feat_threads(0), 0, stream>>>(segs, a64, features); break;
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
