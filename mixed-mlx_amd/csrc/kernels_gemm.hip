// Batched square matrix products on the matrix cores, for every matrix of an
// uneven batch in one launch.
//
// Reference: matmul_tiled_{f32,bf16} / matmul_tiled_scaled_* and the low-factor
// products (metal_mixed_sqrt.metal:162-366, :1493-1600), which are 16x16
// threadgroup tiles of scalar FMAs, and the float-float residual product
// working_residual (:1927-1942), an O(n) loop per entry.  Here every product
// is an MFMA tile loop:
//
//   type    instruction                              k per MFMA   operand
//   F64     v_mfma_f64_16x16x4_f64                    4           double
//   F32     v_mfma_f32_16x16x4_f32                    4           float
//   Bf16    v_mfma_f32_16x16x32_bf16                  32          8 x bf16
//   Fp8     v_mfma_scale_f32_16x16x128_f8f6f4 (fp8)   128         32 x E4M3FN, unit scales
//   Mxfp4   v_mfma_scale_f32_16x16x128_f8f6f4 (fp4)   128         32 x E2M1 + E8M0
//
// Both operands are stored K-contiguous (A as A^T column-major, B column-
// major), so a 16-byte global load feeds LDS directly and every lane's
// fragment is one contiguous LDS read.  An MXFP4 lane fragment is exactly one
// 32-element block along K, so its E8M0 scale is one byte per lane.
// A workgroup computes a (16 F WM) x (16 F WN) tile, each wave a 16 F x 16 F
// block as F x F MFMA tiles; tiles of all matrices form one flat grid (the
// host's tile list), so small and large orders share a launch.
#include <hip/hip_runtime.h>

#include "device_codecs.hpp"
#include "kernels.hpp"

namespace msq {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

template <GemmType T>
struct Traits;
// KS: K elements per stage; MPS: MFMA instructions per stage along K, each
// covering KS / MPS.  Wider stages (128 bytes of a row for BF16 and FP32)
// measured slower at n = 4096: the larger staging buffers cost occupancy.
template <>
struct Traits<GemmType::F64> {
  static constexpr int KS = 16, BITS = 64, MPS = 4;
  typedef f64x4 Acc;
  typedef double Out;
};
template <>
struct Traits<GemmType::F32> {
  static constexpr int KS = 16, BITS = 32, MPS = 4;
  typedef f32x4 Acc;
  typedef float Out;
};
template <>
struct Traits<GemmType::Bf16> {
  static constexpr int KS = 32, BITS = 16, MPS = 1;
  typedef f32x4 Acc;
  typedef float Out;
};
template <>
struct Traits<GemmType::Fp8> {
  static constexpr int KS = 128, BITS = 8, MPS = 1;
  typedef f32x4 Acc;
  typedef float Out;
};
template <>
struct Traits<GemmType::Mxfp4> {
  static constexpr int KS = 128, BITS = 4, MPS = 1;
  typedef f32x4 Acc;
  typedef float Out;
};

// One MFMA sub-step on fragments read from LDS row pointers.
template <GemmType T>
__device__ __forceinline__ void mma_step(typename Traits<T>::Acc& acc, const unsigned char* arow,
                                         const unsigned char* brow, int sub, int quad,
                                         int sa, int sb) {
  if constexpr (T == GemmType::F64) {
    const double a = reinterpret_cast<const double*>(arow)[4 * sub + quad];
    const double b = reinterpret_cast<const double*>(brow)[4 * sub + quad];
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  } else if constexpr (T == GemmType::F32) {
    const float a = reinterpret_cast<const float*>(arow)[4 * sub + quad];
    const float b = reinterpret_cast<const float*>(brow)[4 * sub + quad];
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  } else if constexpr (T == GemmType::Bf16) {
    // 32 K per instruction: lane quad q holds K 8 q .. 8 q + 7 (16 bytes).
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(arow + 64 * sub + 16 * quad);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(brow + 64 * sub + 16 * quad);
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  } else if constexpr (T == GemmType::Fp8) {
    // The f8f6f4 instruction with E4M3 operands (format 0) and unit scales
    // (E8M0 127): 128 K per instruction, lane quad q holds K 32 q .. 32 q + 31.
    const uint4* pa = reinterpret_cast<const uint4*>(arow + 128 * sub + 32 * quad);
    const uint4* pb = reinterpret_cast<const uint4*>(brow + 128 * sub + 32 * quad);
    const uint4 a0 = pa[0], a1 = pa[1], b0 = pb[0], b1 = pb[1];
    i32x8 a = {static_cast<int>(a0.x), static_cast<int>(a0.y), static_cast<int>(a0.z),
               static_cast<int>(a0.w), static_cast<int>(a1.x), static_cast<int>(a1.y),
               static_cast<int>(a1.z), static_cast<int>(a1.w)};
    i32x8 b = {static_cast<int>(b0.x), static_cast<int>(b0.y), static_cast<int>(b0.z),
               static_cast<int>(b0.w), static_cast<int>(b1.x), static_cast<int>(b1.y),
               static_cast<int>(b1.z), static_cast<int>(b1.w)};
    acc = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, acc, 0, 0, 0, 127, 0, 127);
  } else {
    // 128 K per instruction: lane quad q holds the E2M1 block K 32 q .. 32 q + 31.
    const uint4 a4 = *reinterpret_cast<const uint4*>(arow + 64 * sub + 16 * quad);
    const uint4 b4 = *reinterpret_cast<const uint4*>(brow + 64 * sub + 16 * quad);
    i32x8 a = {static_cast<int>(a4.x), static_cast<int>(a4.y), static_cast<int>(a4.z),
               static_cast<int>(a4.w), 0, 0, 0, 0};
    i32x8 b = {static_cast<int>(b4.x), static_cast<int>(b4.y), static_cast<int>(b4.z),
               static_cast<int>(b4.w), 0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, acc, 4, 4, 0, sa, 0, sb);
  }
}

template <GemmType T, int WM, int WN, int F, Epilogue EPI>
__global__ __launch_bounds__(64 * WM * WN) void gemm_kernel(GemmArgs args) {
  typedef Traits<T> Tr;
  constexpr int NT = 64 * WM * WN;
  constexpr int BM = 16 * F * WM, BN = 16 * F * WN;
  constexpr int RB = Tr::KS * Tr::BITS / 8;  // bytes of one row per stage
  constexpr int RP = RB + 16;                // padded LDS row pitch
  constexpr int CHUNKS = RB / 16;
  constexpr bool MX = T == GemmType::Mxfp4;

  // XCD-aware tile order: workgroups are dealt to the 8 XCDs round-robin
  // (workgroup b to XCD b mod 8), so XCD x takes the x-th contiguous eighth
  // of the tile list.  A matrix's tiles are adjacent in the list, so its
  // operand panels are read into one XCD's L2 and reused there.
  const uint32_t nb = gridDim.x, q = nb / 8, rem = nb % 8;
  const uint32_t xcd = blockIdx.x % 8, k = blockIdx.x / 8;
  const uint32_t tile = (xcd < rem ? xcd * (q + 1) : rem * (q + 1) + (xcd - rem) * q) + k;
  const GemmJob job = args.jobs[args.tile_job[tile]];
  if (args.active && !args.active[job.status_index]) return;
  const int tm = static_cast<int>(args.tile_pos[tile] >> 16) * BM;
  const int tn = static_cast<int>(args.tile_pos[tile] & 0xffffu) * BN;
  const int ld = static_cast<int>(job.ld);
  const int order = static_cast<int>(job.order);

  typedef typename Tr::Out Out;
  // LDS: the K-step stages of A and B, reused after the K loop for the
  // output tile when the epilogue goes through LDS (column pitch TP).
  constexpr int TP = BM + 1;
  constexpr int STAGE_BYTES = (BM + BN) * RP;
  constexpr int TILE_BYTES = BN * TP * static_cast<int>(sizeof(Out));
  constexpr bool STAGED = EPI != Epilogue::Gram && TILE_BYTES <= 36 * 1024;
  constexpr int SMEM = STAGED && TILE_BYTES > STAGE_BYTES ? TILE_BYTES : STAGE_BYTES;
  __shared__ __align__(16) unsigned char smem[SMEM];
  unsigned char* const As = smem;
  unsigned char* const Bs = smem + BM * RP;
  constexpr int NBLK = Tr::KS / 32;  // MX: 32-element scale blocks per row per stage
  __shared__ uint8_t Ss[MX ? (BM + BN) * NBLK : 1];

  const unsigned char* abase;
  const unsigned char* bbase;
  if constexpr (MX) {
    abase = static_cast<const unsigned char*>(args.a) + job.a_offset / 2;
    bbase = static_cast<const unsigned char*>(args.b) + job.b_offset / 2;
  } else {
    abase = static_cast<const unsigned char*>(args.a) + job.a_offset * (Tr::BITS / 8);
    bbase = static_cast<const unsigned char*>(args.b) + job.b_offset * (Tr::BITS / 8);
  }
  const uint8_t* asc = MX ? args.a_scales + job.a_offset / 32 : nullptr;
  const uint8_t* bsc = MX ? args.b_scales + job.b_offset / 32 : nullptr;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = (wave % WM) * 16 * F;
  const int wn = (wave / WM) * 16 * F;
  const int quad = lane >> 4;
  const int row16 = lane & 15;

  typename Tr::Acc acc[F][F];
#pragma unroll
  for (int i = 0; i < F; ++i)
#pragma unroll
    for (int j = 0; j < F; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[i][j][r] = 0;

  const int kend = order < ld ? ((order + 31) & ~31) : ld;
  const size_t row_bytes = MX ? static_cast<size_t>(ld) / 2
                              : static_cast<size_t>(ld) * (Tr::BITS / 8);
  // Software pipeline: the 16-byte chunks of the next K step are loaded into
  // registers while the MFMAs of this step run on LDS.
  constexpr int TOTAL = (BM + BN) * CHUNKS;
  constexpr int CPT = (TOTAL + NT - 1) / NT;  // chunks per thread per step
  uint4 pre[CPT];
  auto global_load = [&](int kk) {
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int c = tid + u * NT;
      pre[u] = make_uint4(0, 0, 0, 0);
      if (c < TOTAL) {
        const int r = c / CHUNKS;
        const int ch = c % CHUNKS;
        const bool is_a = r < BM;
        const int g = (is_a ? tm : tn) + (is_a ? r : r - BM);
        const int k0 = kk + ch * (128 / Tr::BITS);  // first element of the chunk
        if (g < ld && k0 < kend) {
          const unsigned char* src = (is_a ? abase : bbase) + static_cast<size_t>(g) * row_bytes +
                                     (MX ? static_cast<size_t>(k0) / 2
                                         : static_cast<size_t>(k0) * (Tr::BITS / 8));
          pre[u] = *reinterpret_cast<const uint4*>(src);
        }
      }
    }
  };
  global_load(0);
  for (int kk = 0; kk < kend; kk += Tr::KS) {
    // Stage A (BM rows) and B (BN rows) of this K step.
#pragma unroll
    for (int u = 0; u < CPT; ++u) {
      const int c = tid + u * NT;
      if (c < TOTAL) {
        const int r = c / CHUNKS;
        const int ch = c % CHUNKS;
        const bool is_a = r < BM;
        *reinterpret_cast<uint4*>((is_a ? As : Bs) + (is_a ? r : r - BM) * RP + ch * 16) = pre[u];
      }
    }
    if constexpr (MX) {
      for (int c = tid; c < (BM + BN) * NBLK; c += NT) {
        const int r = c / NBLK;
        const int blk = c % NBLK;
        const bool is_a = r < BM;
        const int rr = is_a ? r : r - BM;
        const int g = (is_a ? tm : tn) + rr;
        const int k0 = kk + 32 * blk;
        uint8_t s = 127;
        if (g < ld && k0 < kend)
          s = (is_a ? asc : bsc)[(static_cast<size_t>(g) * ld + k0) / 32];
        Ss[r * NBLK + blk] = s;
      }
    }
    __syncthreads();
    if (kk + Tr::KS < kend) global_load(kk + Tr::KS);
#pragma unroll
    for (int i = 0; i < F; ++i) {
      const int ar = wm + 16 * i + row16;
      const unsigned char* arow = As + ar * RP;
#pragma unroll
      for (int j = 0; j < F; ++j) {
        const int br = wn + 16 * j + row16;
        const unsigned char* brow = Bs + br * RP;
#pragma unroll
        for (int s = 0; s < Tr::MPS; ++s) {
          // Sub-steps past the order's K range hold zeros: skipped.
          if (s > 0 && kk + s * (Tr::KS / Tr::MPS) >= kend) break;
          int sa = 0, sb = 0;
          if constexpr (MX) {
            sa = Ss[ar * NBLK + 4 * s + quad];
            sb = Ss[(BM + br) * NBLK + 4 * s + quad];
          }
          mma_step<T>(acc[i][j], arow, brow, s, quad, sa, sb);
        }
      }
    }
    __syncthreads();
  }

  // Epilogue.  The f32-accumulating MFMAs leave C[row = 4 quad + r][col =
  // row16] in accumulator r of a lane; v_mfma_f64_16x16x4_f64 leaves
  // C[row = 4 r + quad][col = row16].
  const float alpha = args.alpha ? args.alpha[job.status_index] : 1.0f;
  Out* C = static_cast<Out*>(args.c) + job.c_offset;
  double norm2 = 0.0;
  const bool mirror_tile = args.mirror && tm != tn;
  // Entry (row, col) of the product, v, written at (row, col) ("main") or,
  // for the symmetric and transposed forms, at (col, row) ("mirror").
  auto entry_main = [&](int row, int col, Out v) {
    const size_t idx = static_cast<size_t>(row) + static_cast<size_t>(col) * ld;
    const bool in = row < order && col < order;
    if constexpr (EPI == Epilogue::Residual) {
      const double* A = static_cast<const double*>(args.aux) + job.aux_offset;
      const double res = A[idx] - static_cast<double>(v);
      if (args.c) C[idx] = static_cast<Out>(res);
      if (args.c_float) args.c_float[job.c_offset + idx] = static_cast<float>(res);
      if (args.c_bf16) args.c_bf16[job.c_offset + idx] = encode_bf16(static_cast<float>(res));
      if (in) norm2 += res * res;
    } else if constexpr (EPI == Epilogue::Gram) {
      if (!in) return;
      const double d = static_cast<double>(static_cast<float>(v) * alpha) - (row == col ? 1.0 : 0.0);
      norm2 += (mirror_tile ? 2.0 : 1.0) * d * d;
    } else if constexpr (EPI == Epilogue::AddSym) {
      // Upper entries only (row <= col), added at (row, col) here and at
      // (col, row) by the mirror pass, so X stays exactly symmetric.
      if (row > col || col >= order) return;
      double* X = static_cast<double*>(args.c) + job.c_offset;
      X[idx] += static_cast<double>(static_cast<float>(v) * alpha);
    } else if constexpr (T == GemmType::F64) {
      if (!args.transpose_c) C[idx] = v;
    } else {
      if (args.transpose_c) return;
      if (EPI == Epilogue::Hadamard && args.mirror && row > col) return;
      float w = static_cast<float>(v) * alpha;
      if constexpr (EPI == Epilogue::Hadamard) {
        const float* s = static_cast<const float*>(args.aux) + job.aux_offset;
        w = in ? w * (1.0f / (s[row] + s[col])) : 0.0f;
      }
      if (EPI == Epilogue::Accumulate)
        C[idx] += args.c_scale * w;
      else if (args.c_bf16)
        args.c_bf16[job.c_offset + idx] = encode_bf16(w);
      else
        C[idx] = w;
    }
  };
  auto entry_mirror = [&](int row, int col, Out v) {
    const size_t midx = static_cast<size_t>(col) + static_cast<size_t>(row) * ld;
    const bool in = row < order && col < order;
    if constexpr (EPI == Epilogue::Residual) {
      // A_ji - (X X)_ij: (X X)_ji = (X X)_ij for the solve's symmetric X.
      if (!mirror_tile) return;
      const double* A = static_cast<const double*>(args.aux) + job.aux_offset;
      const double res = A[midx] - static_cast<double>(v);
      if (args.c) C[midx] = static_cast<Out>(res);
      if (args.c_float) args.c_float[job.c_offset + midx] = static_cast<float>(res);
      if (args.c_bf16) args.c_bf16[job.c_offset + midx] = encode_bf16(static_cast<float>(res));
      if (in) norm2 += res * res;
    } else if constexpr (EPI == Epilogue::AddSym) {
      if (row >= col || col >= order) return;
      double* X = static_cast<double*>(args.c) + job.c_offset;
      X[midx] += static_cast<double>(static_cast<float>(v) * alpha);
    } else if constexpr (EPI == Epilogue::Gram) {
      return;
    } else if constexpr (T == GemmType::F64) {
      if (args.transpose_c) C[midx] = v;
    } else {
      // Transposed store, or the mirror of a symmetric Hadamard product.
      const bool sym = EPI == Epilogue::Hadamard && args.mirror;
      if (!args.transpose_c && !(sym && row < col)) return;
      float w = static_cast<float>(v) * alpha;
      if constexpr (EPI == Epilogue::Hadamard) {
        const float* s = static_cast<const float*>(args.aux) + job.aux_offset;
        w = in ? w * (1.0f / (s[row] + s[col])) : 0.0f;
      }
      if (EPI == Epilogue::Accumulate)
        C[midx] += args.c_scale * w;
      else if (args.c_bf16)
        args.c_bf16[job.c_offset + midx] = encode_bf16(w);
      else
        C[midx] = w;
    }
  };
  if constexpr (STAGED) {
    // Through LDS: the tile goes to LDS column-major, then every thread
    // walks it along columns for the (row, col) writes and along rows for
    // the (col, row) ones, so both are coalesced in global memory.
    Out* Tt = reinterpret_cast<Out*>(smem);
#pragma unroll
    for (int i = 0; i < F; ++i)
#pragma unroll
      for (int j = 0; j < F; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int lr = wm + 16 * i + (T == GemmType::F64 ? 4 * r + quad : 4 * quad + r);
          const int lc = wn + 16 * j + row16;
          Tt[lc * TP + lr] = static_cast<Out>(acc[i][j][r]);
        }
    __syncthreads();
    for (int e = tid; e < BM * BN; e += NT) {
      const int lr = e % BM, lc = e / BM;
      const int row = tm + lr, col = tn + lc;
      if (row < ld && col < ld) entry_main(row, col, Tt[lc * TP + lr]);
    }
    const bool second = (EPI == Epilogue::Residual && mirror_tile) || EPI == Epilogue::AddSym ||
                        args.transpose_c || (EPI == Epilogue::Hadamard && args.mirror);
    if (second) {
      for (int e = tid; e < BM * BN; e += NT) {
        const int lc = e % BN, lr = e / BN;
        const int row = tm + lr, col = tn + lc;
        if (row < ld && col < ld) entry_mirror(row, col, Tt[lc * TP + lr]);
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < F; ++i) {
#pragma unroll
      for (int j = 0; j < F; ++j) {
        const int col = tn + wn + 16 * j + row16;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = tm + wm + 16 * i + (T == GemmType::F64 ? 4 * r + quad : 4 * quad + r);
          if (row >= ld || col >= ld) continue;
          const Out v = static_cast<Out>(acc[i][j][r]);
          entry_main(row, col, v);
          entry_mirror(row, col, v);
        }
      }
    }
  }
  if constexpr (EPI == Epilogue::Residual || EPI == Epilogue::Gram) {
    // A fixed butterfly within the wave, then one slot per wave.
    for (int off = 32; off > 0; off >>= 1) norm2 += __shfl_xor(norm2, off);
    if (lane == 0)
      args.partials[args.partial_base + static_cast<size_t>(tile) * (NT / 64) + wave] = norm2;
  }
}

__global__ void gemm_norms_kernel(const GemmJob* __restrict__ jobs,
                                  const uint32_t* __restrict__ first,
                                  const uint32_t* __restrict__ count, uint32_t njobs,
                                  const double* __restrict__ partials,
                                  const int* __restrict__ active, double* __restrict__ norms) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= njobs) return;
  const uint32_t i = jobs[k].status_index;
  if (active && !active[i]) return;
  double s = 0.0;
  for (uint32_t q = 0; q < count[k]; ++q) s += partials[static_cast<size_t>(first[k]) + q];
  norms[i] = s;
}

template <GemmType T, Epilogue E>
hipError_t launch_typed(int tile, const GemmArgs& args, uint32_t tiles, hipStream_t stream) {
  // 32: one wave of 2 x 2 MFMA blocks; 64: four such waves; 128: four
  // waves of 4 x 4 blocks (each fragment read from LDS serves four MFMAs).
  if (tile == 32) {
    gemm_kernel<T, 1, 1, 2, E><<<tiles, 64, 0, stream>>>(args);
  } else if (tile == 64) {
    gemm_kernel<T, 2, 2, 2, E><<<tiles, 256, 0, stream>>>(args);
  } else if (tile == 128) {
    gemm_kernel<T, 2, 2, 4, E><<<tiles, 256, 0, stream>>>(args);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <GemmType T>
hipError_t launch_epilogue(Epilogue epi, int tile, const GemmArgs& args, uint32_t tiles,
                           hipStream_t stream) {
  switch (epi) {
    case Epilogue::Store: return launch_typed<T, Epilogue::Store>(tile, args, tiles, stream);
    case Epilogue::Hadamard: return launch_typed<T, Epilogue::Hadamard>(tile, args, tiles, stream);
    case Epilogue::AddSym: return launch_typed<T, Epilogue::AddSym>(tile, args, tiles, stream);
    case Epilogue::Gram: return launch_typed<T, Epilogue::Gram>(tile, args, tiles, stream);
    case Epilogue::Accumulate:
      return launch_typed<T, Epilogue::Accumulate>(tile, args, tiles, stream);
    default: return hipErrorInvalidValue;
  }
}

hipError_t dispatch_epilogue(GemmType type, Epilogue epi, int tile, const GemmArgs& args,
                             uint32_t tiles, hipStream_t stream) {
  switch (type) {
    case GemmType::F32: return launch_epilogue<GemmType::F32>(epi, tile, args, tiles, stream);
    case GemmType::Bf16: return launch_epilogue<GemmType::Bf16>(epi, tile, args, tiles, stream);
    case GemmType::Fp8: return launch_epilogue<GemmType::Fp8>(epi, tile, args, tiles, stream);
    case GemmType::Mxfp4: return launch_epilogue<GemmType::Mxfp4>(epi, tile, args, tiles, stream);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

int gemm_waves(int tile) { return tile == 32 ? 1 : 4; }

hipError_t launch_gemm_norms(const GemmJob* jobs, const uint32_t* first, const uint32_t* count,
                             uint32_t njobs, const double* partials, const int* active,
                             double* norms, hipStream_t stream) {
  if (njobs == 0) return hipSuccess;
  gemm_norms_kernel<<<(njobs + 255) / 256, 256, 0, stream>>>(jobs, first, count, njobs, partials,
                                                              active, norms);
  return hipGetLastError();
}

hipError_t launch_gemm(GemmType type, Epilogue epi, int tile, const GemmArgs& args,
                       uint32_t tiles, hipStream_t stream) {
  if (tiles == 0) return hipSuccess;
  if (epi == Epilogue::Residual) {
    if (type != GemmType::F64) return hipErrorInvalidValue;
    return launch_typed<GemmType::F64, Epilogue::Residual>(tile, args, tiles, stream);
  }
  if (type == GemmType::F64)
    return epi == Epilogue::Store
               ? launch_typed<GemmType::F64, Epilogue::Store>(tile, args, tiles, stream)
               : hipErrorInvalidValue;
  return dispatch_epilogue(type, epi, tile, args, tiles, stream);
}

}  // namespace msq
