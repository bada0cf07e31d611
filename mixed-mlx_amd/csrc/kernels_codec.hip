// Codec kernels: FP32 <-> BF16 / FP8 E4M3FN / FP8 E5M2 / FP4 E2M1 / MXFP8 /
// MXFP4 on linear (column-major, reference-layout) arrays, plus the per-tensor
// amax-scaled FP8 E4M3 variant.
//
// The reference's GPU codecs run one thread per element (one per byte for FP4,
// one per 32-element block for MX, with a serial block max;
// metal_mixed_sqrt.metal:831-1111, :1839-1886).  Here every thread owns eight
// consecutive elements: two 16-byte loads, one 4-16 byte store, and for MX a
// 32-element block spans four adjacent lanes whose maxima meet in two xor
// shuffles.  All launches are grid-stride and HBM-bound.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_codecs.hpp"
#include "kernels.hpp"

namespace msq {
namespace {

constexpr int kThreads = 256;

__device__ __forceinline__ void load8(const float* __restrict__ src, size_t i,
                                      size_t n, float v[8]) {
  if (i + 8 <= n) {
    const float4 a = *reinterpret_cast<const float4*>(src + i);
    const float4 b = *reinterpret_cast<const float4*>(src + i + 4);
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
    v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = i + j < n ? src[i + j] : 0.0f;
  }
}
__device__ __forceinline__ void store8(float* __restrict__ dst, size_t i,
                                       size_t n, const float v[8]) {
  if (i + 8 <= n) {
    *reinterpret_cast<float4*>(dst + i) = make_float4(v[0], v[1], v[2], v[3]);
    *reinterpret_cast<float4*>(dst + i + 4) = make_float4(v[4], v[5], v[6], v[7]);
  } else {
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (i + j < n) dst[i + j] = v[j];
  }
}

// The float inputs are 16-byte aligned only when the base is; the host checks
// alignment and all library buffers are allocated aligned.

__global__ void encode_bf16_kernel(const float* __restrict__ src,
                                   uint16_t* __restrict__ dst, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 8;
  for (size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
       i < n; i += stride) {
    float v[8];
    load8(src, i, n, v);
    uint16_t c[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) c[j] = encode_bf16(v[j]);
    if (i + 8 <= n) {
      uint4 w;
      w.x = c[0] | (static_cast<uint32_t>(c[1]) << 16);
      w.y = c[2] | (static_cast<uint32_t>(c[3]) << 16);
      w.z = c[4] | (static_cast<uint32_t>(c[5]) << 16);
      w.w = c[6] | (static_cast<uint32_t>(c[7]) << 16);
      *reinterpret_cast<uint4*>(dst + i) = w;
    } else {
      for (int j = 0; j < 8; ++j)
        if (i + j < n) dst[i + j] = c[j];
    }
  }
}

__global__ void decode_bf16_kernel(const uint16_t* __restrict__ src,
                                   float* __restrict__ dst, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 8;
  for (size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
       i < n; i += stride) {
    float v[8];
    if (i + 8 <= n) {
      const uint4 w = *reinterpret_cast<const uint4*>(src + i);
      const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        v[2 * j] = decode_bf16(static_cast<uint16_t>(ws[j] & 0xffffu));
        v[2 * j + 1] = decode_bf16(static_cast<uint16_t>(ws[j] >> 16));
      }
    } else {
      for (int j = 0; j < 8; ++j) v[j] = i + j < n ? decode_bf16(src[i + j]) : 0.0f;
    }
    store8(dst, i, n, v);
  }
}

// E4M3FN codes of eight floats, packed (element j in byte j), by the
// hardware conversion v_cvt_pk_fp8_f32 (OCP E4M3FN, round to nearest even),
// with encode_e4m3's rules for what the conversion does differently:
// magnitudes above 448 (infinity included) are clamped to 448 first (code
// 0x7e), and a NaN becomes sign | 0x7f.
__device__ __forceinline__ uint2 encode_e4m3_x8(const float (&x)[8]) {
  float c[8];
  bool nan = false;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    c[j] = fabsf(x[j]) > 448.0f ? copysignf(448.0f, x[j]) : x[j];
    nan = nan || x[j] != x[j];
  }
  int lo = __builtin_amdgcn_cvt_pk_fp8_f32(c[0], c[1], 0, false);
  lo = __builtin_amdgcn_cvt_pk_fp8_f32(c[2], c[3], lo, true);
  int hi = __builtin_amdgcn_cvt_pk_fp8_f32(c[4], c[5], 0, false);
  hi = __builtin_amdgcn_cvt_pk_fp8_f32(c[6], c[7], hi, true);
  uint2 w = make_uint2(static_cast<uint32_t>(lo), static_cast<uint32_t>(hi));
  if (nan) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (x[j] == x[j]) continue;
      const uint32_t code = ((f32_bits(x[j]) >> 24) & 0x80u) | 0x7fu;
      const int s = 8 * (j & 3);
      uint32_t& word = j < 4 ? w.x : w.y;
      word = (word & ~(0xffu << s)) | (code << s);
    }
  }
  return w;
}

// FP8, scalar or per-tensor scaled.  fmt: 0 E4M3FN, 1 E5M2.  When `sigma` is
// non-null the value encoded is v * (*sigma) (one IEEE product).
__global__ void encode_fp8_kernel(const float* __restrict__ src,
                                  uint8_t* __restrict__ dst, size_t n, int fmt,
                                  const float* __restrict__ sigma) {
  const float scale = sigma ? *sigma : 1.0f;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 8;
  for (size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
       i < n; i += stride) {
    float v[8];
    load8(src, i, n, v);
    if (fmt == 0 && i + 8 <= n) {
      float x[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) x[j] = sigma ? v[j] * scale : v[j];
      *reinterpret_cast<uint2*>(dst + i) = encode_e4m3_x8(x);
      continue;
    }
    uint8_t c[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float x = sigma ? v[j] * scale : v[j];
      c[j] = fmt == 0 ? encode_e4m3(x) : encode_e5m2(x);
    }
    if (i + 8 <= n) {
      uint2 w;
      w.x = c[0] | (c[1] << 8) | (c[2] << 16) | (static_cast<uint32_t>(c[3]) << 24);
      w.y = c[4] | (c[5] << 8) | (c[6] << 16) | (static_cast<uint32_t>(c[7]) << 24);
      *reinterpret_cast<uint2*>(dst + i) = w;
    } else {
      for (int j = 0; j < 8; ++j)
        if (i + j < n) dst[i + j] = c[j];
    }
  }
}

__global__ void decode_fp8_kernel(const uint8_t* __restrict__ src,
                                  float* __restrict__ dst, size_t n, int fmt,
                                  const float* __restrict__ sigma) {
  const float scale = sigma ? *sigma : 1.0f;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 8;
  for (size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
       i < n; i += stride) {
    uint8_t c[8];
    if (i + 8 <= n) {
      const uint2 w = *reinterpret_cast<const uint2*>(src + i);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        c[j] = (w.x >> (8 * j)) & 0xffu;
        c[j + 4] = (w.y >> (8 * j)) & 0xffu;
      }
    } else {
      for (int j = 0; j < 8; ++j) c[j] = i + j < n ? src[i + j] : 0;
    }
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float x = fmt == 0 ? decode_e4m3(c[j]) : decode_e5m2(c[j]);
      v[j] = sigma ? x / scale : x;  // correctly rounded, as the oracle
    }
    store8(dst, i, n, v);
  }
}

__global__ void encode_fp4_kernel(const float* __restrict__ src,
                                  uint8_t* __restrict__ dst, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 8;
  for (size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
       i < n; i += stride) {
    float v[8];
    load8(src, i, n, v);
    uint32_t w = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      // element beyond n packs as zero nibble (packFp4 leaves the upper nibble 0)
      const uint32_t c = i + j < n ? encode_e2m1(v[j]) : 0u;
      w |= c << (4 * j);
    }
    if (i + 8 <= n) {
      *reinterpret_cast<uint32_t*>(dst + i / 2) = w;
    } else {
      for (size_t j = 0; j < 8 && i + j < n; j += 2) dst[(i + j) / 2] = (w >> (4 * j)) & 0xffu;
    }
  }
}

// fmt: 0 FP4 E2M1 scalar, 1 MXFP4 E2M1.
__global__ void decode_fp4_kernel(const uint8_t* __restrict__ src,
                                  const uint8_t* __restrict__ scales,
                                  float* __restrict__ dst, size_t n) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 8;
  for (size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
       i < n; i += stride) {
    uint32_t w;
    if (i + 8 <= n) {
      w = *reinterpret_cast<const uint32_t*>(src + i / 2);
    } else {
      w = 0;
      for (size_t j = 0; j < 8 && i + j < n; j += 2)
        w |= static_cast<uint32_t>(src[(i + j) / 2]) << (4 * j);
    }
    const int e = scales ? static_cast<int>(scales[i / 32]) - 127 : 0;
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float x = decode_e2m1((w >> (4 * j)) & 0xfu);
      v[j] = scales ? ldexp_exact(x, e) : x;
    }
    store8(dst, i, n, v);
  }
}

// MX encode on the reference linear layout: block b covers elements
// [32 b, 32 b + 32) (a partial last block covers the tail).  Four adjacent
// lanes own one block.  fmt: 0 MXFP8 E4M3FN (emax 8), 1 MXFP8 E5M2 (emax 15),
// 2 MXFP4 E2M1 (emax 2).
__global__ void encode_mx_kernel(const float* __restrict__ src,
                                 uint8_t* __restrict__ dst,
                                 uint8_t* __restrict__ scales, size_t n, int fmt) {
  const size_t blocks = (n + 31) / 32;
  const size_t lanes_total = blocks * 4;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  // Every lane of a wave runs the same number of iterations so that the
  // shuffles below always see their partners.
  const size_t base0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t iters = (lanes_total + stride - 1) / stride;
  for (size_t it = 0; it < iters; ++it) {
    const size_t lane = base0 + it * stride;
    const bool active = lane < lanes_total;
    const size_t i = lane * 8;
    float v[8];
    uint32_t m = 0;
    if (active) {
      load8(src, i, n, v);
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (i + j < n) m = max(m, finite_abs_bits(v[j]));
    }
    m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), 1, 4)));
    m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), 2, 4)));
    if (!active) continue;
    const int emax = fmt == 0 ? 8 : (fmt == 1 ? 15 : 2);
    const int e = mx_scale_exponent(m, emax);
    if ((lane & 3) == 0) scales[lane / 4] = static_cast<uint8_t>(e + 127);
    if (fmt == 2) {
      uint32_t w = 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const uint32_t c = i + j < n ? encode_e2m1(v[j], e) : 0u;
        w |= c << (4 * j);
      }
      if (i + 8 <= n) {
        *reinterpret_cast<uint32_t*>(dst + i / 2) = w;
      } else {
        for (size_t j = 0; j < 8 && i + j < n; j += 2) dst[(i + j) / 2] = (w >> (4 * j)) & 0xffu;
      }
    } else {
      uint8_t c[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) c[j] = fmt == 0 ? encode_e4m3(v[j], e) : encode_e5m2(v[j], e);
      if (i + 8 <= n) {
        uint2 w;
        w.x = c[0] | (c[1] << 8) | (c[2] << 16) | (static_cast<uint32_t>(c[3]) << 24);
        w.y = c[4] | (c[5] << 8) | (c[6] << 16) | (static_cast<uint32_t>(c[7]) << 24);
        *reinterpret_cast<uint2*>(dst + i) = w;
      } else {
        for (int j = 0; j < 8; ++j)
          if (i + j < n) dst[i + j] = c[j];
      }
    }
  }
}

// MXFP8 decode (fmt 0 E4M3FN, 1 E5M2).
__global__ void decode_mxfp8_kernel(const uint8_t* __restrict__ src,
                                    const uint8_t* __restrict__ scales,
                                    float* __restrict__ dst, size_t n, int fmt) {
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 8;
  for (size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
       i < n; i += stride) {
    uint8_t c[8];
    if (i + 8 <= n) {
      const uint2 w = *reinterpret_cast<const uint2*>(src + i);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        c[j] = (w.x >> (8 * j)) & 0xffu;
        c[j + 4] = (w.y >> (8 * j)) & 0xffu;
      }
    } else {
      for (int j = 0; j < 8; ++j) c[j] = i + j < n ? src[i + j] : 0;
    }
    const int e = static_cast<int>(scales[i / 32]) - 127;
    float v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      v[j] = ldexp_exact(fmt == 0 ? decode_e4m3(c[j]) : decode_e5m2(c[j]), e);
    store8(dst, i, n, v);
  }
}

// amax over finite |v| as a float bit pattern (atomicMax on the pattern orders
// like the value for non-negative floats).  *out must be zeroed by the caller.
__global__ void amax_kernel(const float* __restrict__ src, size_t n,
                            uint32_t* __restrict__ out) {
  uint32_t m = 0;
  // Two groups of eight per iteration, both loaded before either is reduced
  // (four 16-byte loads in flight per lane: the grid is capped, so each lane
  // walks the tensor in several steps).
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 8;
  for (size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
       i < n; i += 2 * stride) {
    float v[8], w[8];
    load8(src, i, n, v);
    const size_t i2 = i + stride;
    if (i2 < n) load8(src, i2, n, w);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (i + j < n) m = max(m, finite_abs_bits(v[j]));
    if (i2 < n) {
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (i2 + j < n) m = max(m, finite_abs_bits(w[j]));
    }
  }
  for (int off = 32; off > 0; off >>= 1)
    m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), off)));
  // One atomic per block: the grid is capped, so a large tensor costs a few
  // hundred atomics on the one address rather than one per wave.
  __shared__ uint32_t wave_max[kThreads / 64];
  if ((threadIdx.x & 63) == 0) wave_max[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t b = 0;
    for (int w = 0; w < static_cast<int>(blockDim.x) / 64; ++w) b = max(b, wave_max[w]);
    atomicMax(out, b);
  }
}

// sigma = (float)(448 / (double)amax), 1 when amax is 0.
__global__ void fp8_sigma_kernel(const uint32_t* __restrict__ amax_bits,
                                 float* __restrict__ sigma) {
  const float amax = bits_f32(*amax_bits);
  *sigma = amax == 0.0f ? 1.0f : static_cast<float>(448.0 / static_cast<double>(amax));
}

__global__ void zero_u32_kernel(uint32_t* p) { *p = 0u; }

// Stochastic encode on the linear layout (the reference's *_stochastic
// quantize kernels): element i rounds with the hash of (i, seed, stream).
// Eight elements per thread; an MX block spans four adjacent lanes (its scale
// as in the nearest mode).  fmt: the Format value.
__global__ void encode_stochastic_kernel(const float* __restrict__ src, uint8_t* __restrict__ dst,
                                         uint8_t* __restrict__ scales, size_t n, int fmt,
                                         uint32_t seed, uint32_t stream) {
  const size_t lanes_total = (n + 7) / 8;
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
  const size_t base0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool mx = fmt == static_cast<int>(Format::Mxfp8E4M3Fn) ||
                  fmt == static_cast<int>(Format::Mxfp8E5M2) ||
                  fmt == static_cast<int>(Format::Mxfp4E2M1);
  // Whole waves run the same iterations (the MX shuffles see their partners):
  // the lane count is padded to a multiple of four lanes per block.
  const size_t lanes_padded = mx ? (n + 31) / 32 * 4 : lanes_total;
  const size_t iters = (lanes_padded + stride - 1) / stride;
  for (size_t it = 0; it < iters; ++it) {
    const size_t lane = base0 + it * stride;
    const bool active = lane < lanes_padded;
    const size_t i = lane * 8;
    float v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t m = 0;
    if (active && i < n) {
      load8(src, i, n, v);
#pragma unroll
      for (int j = 0; j < 8; ++j)
        if (i + j < n) m = max(m, finite_abs_bits(v[j]));
    }
    int e = 0;
    if (mx) {
      m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), 1, 4)));
      m = max(m, static_cast<uint32_t>(__shfl_xor(static_cast<int>(m), 2, 4)));
      const int emax = fmt == static_cast<int>(Format::Mxfp8E4M3Fn)
                           ? 8
                           : (fmt == static_cast<int>(Format::Mxfp8E5M2) ? 15 : 2);
      e = mx_scale_exponent(m, emax);
      if (active && i < n && (lane & 3) == 0) scales[lane / 4] = static_cast<uint8_t>(e + 127);
    }
    if (!active || i >= n) continue;
    const float scale = pow2f(e);
    const uint32_t idx = static_cast<uint32_t>(i);
    if (fmt == static_cast<int>(Format::Bf16)) {
      uint16_t* d = reinterpret_cast<uint16_t*>(dst);
      for (int j = 0; j < 8 && i + j < n; ++j)
        d[i + j] = encode_bf16_stochastic(v[j], idx + j, seed, stream);
    } else if (fmt == static_cast<int>(Format::Fp4E2M1) ||
               fmt == static_cast<int>(Format::Mxfp4E2M1)) {
      uint32_t w = 0;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        const float x = mx ? v[j] / scale : v[j];
        const uint32_t c = i + j < n ? encode_e2m1_stochastic(x, idx + j, seed, stream) : 0u;
        w |= c << (4 * j);
      }
      for (size_t j = 0; j < 8 && i + j < n; j += 2) dst[(i + j) / 2] = (w >> (4 * j)) & 0xffu;
    } else {
      const bool e4 = fmt == static_cast<int>(Format::Fp8E4M3Fn) ||
                      fmt == static_cast<int>(Format::Mxfp8E4M3Fn);
      for (int j = 0; j < 8 && i + j < n; ++j) {
        const float x = mx ? v[j] / scale : v[j];
        dst[i + j] = e4 ? encode_e4m3_stochastic(x, idx + j, seed, stream)
                        : encode_e5m2_stochastic(x, idx + j, seed, stream);
      }
    }
  }
}

// Log-affine spectrum encode: one thread per block of kLogBlock elements
// (metadata, then its codes).  e4m3: 8-bit E4M3FN codes, else packed E2M1
// nibbles (a block is four whole bytes).
__global__ void encode_log_kernel(const float* __restrict__ src, uint8_t* __restrict__ dst,
                                  float* __restrict__ metadata, size_t n, int e4m3,
                                  int stochastic, uint32_t seed, uint32_t stream) {
  const size_t blocks = (n + kLogBlock - 1) / kLogBlock;
  const float maximum = e4m3 ? 448.0f : 6.0f;
  for (size_t b = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; b < blocks;
       b += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const uint32_t begin = static_cast<uint32_t>(b * kLogBlock);
    const uint32_t end = static_cast<uint32_t>(min(n, b * kLogBlock + kLogBlock));
    float mn, span;
    log_block_metadata(src, begin, end, mn, span);
    metadata[2 * b] = mn;
    metadata[2 * b + 1] = span;
    uint8_t nib[kLogBlock] = {};
    for (uint32_t k = begin; k < end; ++k) {
      const float x = log_normalized(src[k], mn, span, maximum);
      uint8_t c;
      if (e4m3)
        c = stochastic ? encode_e4m3_stochastic(x, k, seed, stream) : encode_e4m3(x);
      else
        c = stochastic ? encode_e2m1_stochastic(x, k, seed, stream) : encode_e2m1(x);
      if (e4m3)
        dst[k] = c;
      else
        nib[k - begin] = c;
    }
    if (!e4m3)
      for (uint32_t k = begin; k < end; k += 2)
        dst[k / 2] = static_cast<uint8_t>(nib[k - begin] |
                                          (k + 1 < end ? nib[k + 1 - begin] << 4 : 0));
  }
}

__global__ void decode_log_kernel(const uint8_t* __restrict__ codes,
                                  const float* __restrict__ metadata, float* __restrict__ dst,
                                  size_t n, int e4m3) {
  const float maximum = e4m3 ? 448.0f : 6.0f;
  for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += static_cast<size_t>(gridDim.x) * blockDim.x) {
    const size_t b = i / kLogBlock;
    const float d = e4m3 ? decode_e4m3(codes[i]) : decode_e2m1((codes[i / 2] >> (4 * (i & 1))) & 0xfu);
    dst[i] = log_decode(d, metadata[2 * b], metadata[2 * b + 1], maximum);
  }
}

// Element i of values[0, n) (one matrix's spectrum, the reference's
// quantizeAndDecode of an order-length vector) rounded to `precision` and
// decoded: block formats recompute their block's scale or metadata here.
__device__ float quantize_decode_at(int precision, const float* __restrict__ v, uint32_t n,
                                    uint32_t i, int stochastic, uint32_t seed, uint32_t stream) {
  const float x = v[i];
  switch (precision) {
    case 1:  // BF16
      return decode_bf16(stochastic ? encode_bf16_stochastic(x, i, seed, stream) : encode_bf16(x));
    case 2:  // FP8 E4M3FN
      return decode_e4m3(stochastic ? encode_e4m3_stochastic(x, i, seed, stream) : encode_e4m3(x));
    case 3:  // FP8 E5M2
      return decode_e5m2(stochastic ? encode_e5m2_stochastic(x, i, seed, stream) : encode_e5m2(x));
    case 6:  // FP4 E2M1
      return decode_e2m1(stochastic ? encode_e2m1_stochastic(x, i, seed, stream) : encode_e2m1(x));
    case 4:
    case 5:
    case 7: {  // MX: the block's scale from its largest finite magnitude
      const uint32_t b0 = i / 32 * 32, b1 = min(n, b0 + 32);
      uint32_t m = 0;
      for (uint32_t k = b0; k < b1; ++k) m = max(m, finite_abs_bits(v[k]));
      const int emax = precision == 4 ? 8 : (precision == 5 ? 15 : 2);
      const int e = mx_scale_exponent(m, emax);
      float d;
      if (precision == 7)
        d = decode_e2m1(stochastic ? encode_e2m1_stochastic(x / pow2f(e), i, seed, stream)
                                   : encode_e2m1(x, e));
      else if (precision == 4)
        d = decode_e4m3(stochastic ? encode_e4m3_stochastic(x / pow2f(e), i, seed, stream)
                                   : encode_e4m3(x, e));
      else
        d = decode_e5m2(stochastic ? encode_e5m2_stochastic(x / pow2f(e), i, seed, stream)
                                   : encode_e5m2(x, e));
      return ldexp_exact(d, e);
    }
    case 8:
    case 9: {  // log-affine E4M3FN / E2M1, blocks of kLogBlock
      const uint32_t b0 = i / kLogBlock * kLogBlock, b1 = min(n, b0 + kLogBlock);
      float mn, span;
      log_block_metadata(v, b0, b1, mn, span);
      const float maximum = precision == 8 ? 448.0f : 6.0f;
      const float y = log_normalized(x, mn, span, maximum);
      const float d = precision == 8
                          ? decode_e4m3(stochastic ? encode_e4m3_stochastic(y, i, seed, stream)
                                                   : encode_e4m3(y))
                          : decode_e2m1(stochastic ? encode_e2m1_stochastic(y, i, seed, stream)
                                                   : encode_e2m1(y));
      return log_decode(d, mn, span, maximum);
    }
    default:
      return x;
  }
}

// Spectrum tier of a batched solve: each segment's roots (its order-length
// vector) rounded to `precision` and decoded into qroots; a matrix with a
// decoded root that is not positive and finite is rejected (rejected[index]
// = 1) and keeps its float roots, as the reference's spectrum acceptance
// (src/metal_solver.cpp:240-263).  active: per matrix, or null.
__global__ __launch_bounds__(256) void segment_spectrum_kernel(
    const Segment* __restrict__ segs, int precision, const float* __restrict__ roots,
    float* __restrict__ qroots, int* __restrict__ rejected, const int* __restrict__ active,
    int stochastic, uint32_t seed, uint32_t stream) {
  const Segment sg = segs[blockIdx.x];
  if (active && !active[sg.index]) return;
  const float* v = roots + sg.vec_offset;
  float* q = qroots + sg.vec_offset;
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < sg.order; i += blockDim.x) {
    const float d = quantize_decode_at(precision, v, sg.order, i, stochastic, seed, stream);
    q[i] = d;
    if (!(d > 0.0f && d <= 3.4028234663852886e38f)) bad = 1;
  }
  __syncthreads();
  const int reject = bad;
  for (uint32_t i = threadIdx.x; i < sg.ld; i += blockDim.x)
    q[i] = i < sg.order ? (reject ? v[i] : q[i]) : 0.0f;
  if (threadIdx.x == 0) rejected[sg.index] = reject;
}

// --------------------------------------------------- compact matrix storage
// The native-quantized correction stores each matrix as the reference does:
// its n x n elements in column-major linear order (MX blocks of 32
// consecutive elements, FP4 pairs packed from element 0), from compact
// element offset coff[k] (a multiple of 64, so every matrix starts on a whole
// byte and scale block).  The float side is the batch layout (ld x ld, zero
// outside the order), read or written transposed when asked.
__device__ __forceinline__ float compact_source(const float* __restrict__ src, const Segment& sg,
                                                uint32_t e, int transposed) {
  const uint32_t n = sg.order;
  const uint32_t r = e % n, c = e / n;
  return transposed ? src[sg.offset + c + static_cast<size_t>(r) * sg.ld]
                    : src[sg.offset + r + static_cast<size_t>(c) * sg.ld];
}

// One thread per 32-element block of a segment's compact storage.  precision:
// an msq_precision value of the reference codecs (1..7).
__global__ void compact_quantize_kernel(const Segment* __restrict__ segs,
                                        const uint64_t* __restrict__ coff, int precision,
                                        const float* __restrict__ src, int src_transposed,
                                        uint8_t* __restrict__ codes, uint8_t* __restrict__ scales,
                                        const int* __restrict__ active) {
  const Segment sg = segs[blockIdx.y];
  if (active && !active[sg.index]) return;
  const uint32_t total = sg.order * sg.order;
  const uint32_t blocks = (total + 31) / 32;
  const uint64_t base = coff[blockIdx.y];
  for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < blocks; b += gridDim.x * blockDim.x) {
    const uint32_t e0 = 32 * b, e1 = min(total, e0 + 32);
    float v[32];
    uint32_t m = 0;
    for (uint32_t e = e0; e < e1; ++e) {
      v[e - e0] = compact_source(src, sg, e, src_transposed);
      m = max(m, finite_abs_bits(v[e - e0]));
    }
    int sc = 0;
    if (precision == 4 || precision == 5 || precision == 7) {
      sc = mx_scale_exponent(m, precision == 4 ? 8 : (precision == 5 ? 15 : 2));
      scales[base / 32 + b] = static_cast<uint8_t>(sc + 127);
    }
    for (uint32_t e = e0; e < e1; ++e) {
      const float x = v[e - e0];
      const uint64_t at = base + e;
      switch (precision) {
        case 1: reinterpret_cast<uint16_t*>(codes)[at] = encode_bf16(x); break;
        case 2: codes[at] = encode_e4m3(x); break;
        case 3: codes[at] = encode_e5m2(x); break;
        case 4: codes[at] = encode_e4m3(x, sc); break;
        case 5: codes[at] = encode_e5m2(x, sc); break;
        case 6:
        case 7: {
          if ((e & 1u) == 0) {
            const uint8_t lo = encode_e2m1(x, precision == 7 ? sc : 0);
            const uint8_t hi = e + 1 < e1 ? encode_e2m1(v[e + 1 - e0], precision == 7 ? sc : 0) : 0;
            codes[at / 2] = static_cast<uint8_t>(lo | (hi << 4));
          }
          break;
        }
        default: break;
      }
    }
  }
}

__device__ __forceinline__ float compact_decode(int precision, const uint8_t* __restrict__ codes,
                                                const uint8_t* __restrict__ scales, uint64_t at) {
  switch (precision) {
    case 1: return decode_bf16(reinterpret_cast<const uint16_t*>(codes)[at]);
    case 2: return decode_e4m3(codes[at]);
    case 3: return decode_e5m2(codes[at]);
    case 4: return ldexp_exact(decode_e4m3(codes[at]), static_cast<int>(scales[at / 32]) - 127);
    case 5: return ldexp_exact(decode_e5m2(codes[at]), static_cast<int>(scales[at / 32]) - 127);
    case 6: return decode_e2m1((codes[at / 2] >> (4 * (at & 1))) & 0xfu);
    case 7:
      return ldexp_exact(decode_e2m1((codes[at / 2] >> (4 * (at & 1))) & 0xfu),
                         static_cast<int>(scales[at / 32]) - 127);
    default: return 0.0f;
  }
}

// Decodes each segment's compact storage into its ld x ld float tile (zero
// outside the order), transposed when asked.
__global__ void compact_dequantize_kernel(const Segment* __restrict__ segs,
                                          const uint64_t* __restrict__ coff, int precision,
                                          const uint8_t* __restrict__ codes,
                                          const uint8_t* __restrict__ scales,
                                          float* __restrict__ dst, int dst_transposed,
                                          const int* __restrict__ active) {
  const Segment sg = segs[blockIdx.y];
  if (active && !active[sg.index]) return;
  const uint32_t n = sg.order, ld = sg.ld;
  const uint64_t base = coff[blockIdx.y];
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < ld * ld; p += gridDim.x * blockDim.x) {
    const uint32_t r = p % ld, c = p / ld;
    float v = 0.0f;
    if (r < n && c < n) {
      const uint32_t e = dst_transposed ? c + r * n : r + c * n;
      v = compact_decode(precision, codes, scales, base + e);
    }
    dst[sg.offset + p] = v;
  }
}

int grid_for(size_t n, int per_thread) {
  const size_t threads = (n + per_thread - 1) / per_thread;
  size_t blocks = (threads + kThreads - 1) / kThreads;
  if (blocks < 1) blocks = 1;
  if (blocks > 65536) blocks = 65536;
  return static_cast<int>(blocks);
}

}  // namespace

hipError_t launch_encode(Format fmt, const float* src, void* values,
                         uint8_t* scales, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  switch (fmt) {
    case Format::Bf16:
      encode_bf16_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(
          src, static_cast<uint16_t*>(values), n);
      break;
    case Format::Fp8E4M3Fn:
    case Format::Fp8E5M2:
      encode_fp8_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(
          src, static_cast<uint8_t*>(values), n, fmt == Format::Fp8E4M3Fn ? 0 : 1,
          nullptr);
      break;
    case Format::Fp4E2M1:
      encode_fp4_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(
          src, static_cast<uint8_t*>(values), n);
      break;
    case Format::Mxfp8E4M3Fn:
    case Format::Mxfp8E5M2:
    case Format::Mxfp4E2M1: {
      const int f = fmt == Format::Mxfp8E4M3Fn ? 0 : (fmt == Format::Mxfp8E5M2 ? 1 : 2);
      const size_t lanes = (n + 31) / 32 * 4;
      encode_mx_kernel<<<grid_for(lanes, 1), kThreads, 0, stream>>>(
          src, static_cast<uint8_t*>(values), scales, n, f);
      break;
    }
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_decode(Format fmt, const void* values, const uint8_t* scales,
                         float* dst, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  switch (fmt) {
    case Format::Bf16:
      decode_bf16_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(
          static_cast<const uint16_t*>(values), dst, n);
      break;
    case Format::Fp8E4M3Fn:
    case Format::Fp8E5M2:
      decode_fp8_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(
          static_cast<const uint8_t*>(values), dst, n,
          fmt == Format::Fp8E4M3Fn ? 0 : 1, nullptr);
      break;
    case Format::Fp4E2M1:
      decode_fp4_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(
          static_cast<const uint8_t*>(values), nullptr, dst, n);
      break;
    case Format::Mxfp4E2M1:
      decode_fp4_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(
          static_cast<const uint8_t*>(values), scales, dst, n);
      break;
    case Format::Mxfp8E4M3Fn:
    case Format::Mxfp8E5M2:
      decode_mxfp8_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(
          static_cast<const uint8_t*>(values), scales, dst, n,
          fmt == Format::Mxfp8E4M3Fn ? 0 : 1);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_fp8_amax_sigma(const float* src, size_t n, uint32_t* amax_bits,
                                 float* sigma, hipStream_t stream) {
  zero_u32_kernel<<<1, 1, 0, stream>>>(amax_bits);
  if (n > 0) amax_kernel<<<std::min(grid_for(n, 8), 1024), kThreads, 0, stream>>>(src, n, amax_bits);
  fp8_sigma_kernel<<<1, 1, 0, stream>>>(amax_bits, sigma);
  return hipGetLastError();
}

hipError_t launch_encode_fp8_scaled(const float* src, uint8_t* values,
                                    const float* sigma, size_t n,
                                    hipStream_t stream) {
  if (n == 0) return hipSuccess;
  encode_fp8_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(src, values, n, 0, sigma);
  return hipGetLastError();
}

hipError_t launch_decode_fp8_scaled(const uint8_t* values, const float* sigma,
                                    float* dst, size_t n, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  decode_fp8_kernel<<<grid_for(n, 8), kThreads, 0, stream>>>(values, dst, n, 0, sigma);
  return hipGetLastError();
}

hipError_t launch_encode_stochastic(Format fmt, const float* src, void* values, uint8_t* scales,
                                   size_t n, uint32_t seed, uint32_t stream, hipStream_t s) {
  if (n == 0) return hipSuccess;
  if (fmt == Format::Fp32) return hipErrorInvalidValue;
  const bool mx = fmt == Format::Mxfp8E4M3Fn || fmt == Format::Mxfp8E5M2 ||
                  fmt == Format::Mxfp4E2M1;
  const size_t lanes = mx ? (n + 31) / 32 * 4 : (n + 7) / 8;
  encode_stochastic_kernel<<<grid_for(lanes, 1), kThreads, 0, s>>>(
      src, static_cast<uint8_t*>(values), scales, n, static_cast<int>(fmt), seed, stream);
  return hipGetLastError();
}

hipError_t launch_encode_log(bool e4m3, const float* src, uint8_t* values, float* metadata,
                             size_t n, bool stochastic, uint32_t seed, uint32_t stream,
                             hipStream_t s) {
  if (n == 0) return hipSuccess;
  const size_t blocks = (n + kLogBlock - 1) / kLogBlock;
  encode_log_kernel<<<grid_for(blocks, 1), kThreads, 0, s>>>(src, values, metadata, n,
                                                              e4m3 ? 1 : 0, stochastic ? 1 : 0,
                                                              seed, stream);
  return hipGetLastError();
}

hipError_t launch_decode_log(bool e4m3, const uint8_t* values, const float* metadata, float* dst,
                             size_t n, hipStream_t s) {
  if (n == 0) return hipSuccess;
  decode_log_kernel<<<grid_for(n, 1), kThreads, 0, s>>>(values, metadata, dst, n, e4m3 ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_segment_spectrum(const Segment* segs, uint32_t count, int precision,
                                   const float* roots, float* qroots, int* rejected,
                                   const int* active, bool stochastic, uint32_t seed,
                                   uint32_t stream, hipStream_t s) {
  if (count == 0) return hipSuccess;
  segment_spectrum_kernel<<<count, 256, 0, s>>>(segs, precision, roots, qroots, rejected, active,
                                                stochastic ? 1 : 0, seed, stream);
  return hipGetLastError();
}

hipError_t launch_compact_quantize(const Segment* segs, const uint64_t* coff, uint32_t count,
                                  uint32_t max_order, int precision, const float* src,
                                  bool src_transposed, uint8_t* codes, uint8_t* scales,
                                  const int* active, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const uint32_t blocks = (max_order * max_order + 31) / 32;
  const uint32_t gx = std::max<uint32_t>(1, std::min<uint32_t>((blocks + 255) / 256, 1024));
  compact_quantize_kernel<<<dim3(gx, count), 256, 0, s>>>(segs, coff, precision, src,
                                                           src_transposed ? 1 : 0, codes, scales,
                                                           active);
  return hipGetLastError();
}

hipError_t launch_compact_dequantize(const Segment* segs, const uint64_t* coff, uint32_t count,
                                    uint32_t max_ld, int precision, const uint8_t* codes,
                                    const uint8_t* scales, float* dst, bool dst_transposed,
                                    const int* active, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const uint32_t elems = max_ld * max_ld;
  const uint32_t gx = std::max<uint32_t>(1, std::min<uint32_t>((elems + 255) / 256, 1024));
  compact_dequantize_kernel<<<dim3(gx, count), 256, 0, s>>>(segs, coff, precision, codes, scales,
                                                             dst, dst_transposed ? 1 : 0, active);
  return hipGetLastError();
}

}  // namespace msq
