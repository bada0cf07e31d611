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
  const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x * 8;
  for (size_t i = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
       i < n; i += stride) {
    float v[8];
    load8(src, i, n, v);
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (i + j < n) m = max(m, finite_abs_bits(v[j]));
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

}  // namespace msq
