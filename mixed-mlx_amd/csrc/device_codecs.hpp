// Storage codecs for BF16, FP8 E4M3FN / E5M2, FP4 E2M1 and the MX block
// formats, usable from host and device code.
//
// Semantics are those of the reference host codecs (src/storage_codecs.cpp):
//   * BF16: round to nearest even on the bit pattern, NaN keeps a quiet bit
//     (:16-28);
//   * E4M3FN: NaN -> 0x7f, zero keeps its sign, |v| > 448 or inf -> 0x7e,
//     otherwise nearest finite code with ties to the even code (:96-108);
//   * E5M2: NaN -> 0x7e, inf -> 0x7c, finite |v| > 57344 -> 0x7b (:110-125);
//   * FP4 E2M1: NaN and zero -> signed zero, |v| > 6 or inf -> 0x7 (:127-136);
//   * MX: scale exponent clamp(ilogb(max finite |v|) - emax, -127, 128), stored
//     biased by 127, elements encoded as v / 2^exponent (:148-198).
// The reference finds the nearest code by a scan over all codes; here the same
// result comes from round-half-even on the float's bit pattern (normal codes)
// and rintf of an exact power-of-two multiple (subnormal codes): a few integer
// operations and selects per element, no branches.
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define MSQ_HD __host__ __device__ __forceinline__
#else
#define MSQ_HD inline
#endif

namespace msq {

MSQ_HD uint32_t f32_bits(float v) {
  union {
    float f;
    uint32_t u;
  } x;
  x.f = v;
  return x.u;
}
MSQ_HD float bits_f32(uint32_t u) {
  union {
    float f;
    uint32_t u;
  } x;
  x.u = u;
  return x.f;
}

// ---------------------------------------------------------------- BF16
MSQ_HD uint16_t encode_bf16(float value) {
  const uint32_t bits = f32_bits(value);
  uint32_t rounded = bits >> 16;
  const uint32_t remainder = bits & 0xffffu;
  if (remainder > 0x8000u || (remainder == 0x8000u && (rounded & 1u) != 0)) {
    ++rounded;
  }
  if ((bits & 0x7f800000u) == 0x7f800000u && (bits & 0x007fffffu) != 0) {
    rounded |= 0x0040u;
  }
  return static_cast<uint16_t>(rounded);
}
MSQ_HD float decode_bf16(uint16_t code) {
  return bits_f32(static_cast<uint32_t>(code) << 16);
}

// x * 2^k exactly as std::ldexp does for these few-bit values: the product is
// representable unless it overflows (then inf); two steps keep 2^k finite.
MSQ_HD float ldexp_exact(float x, int k) {
  if (k > 127) {
    x *= bits_f32(0x7f000000u);  // 2^127
    k -= 127;
  } else if (k < -126) {
    x *= bits_f32(0x00800000u);  // 2^-126
    k += 126;
  }
  return x * bits_f32(static_cast<uint32_t>(k + 127) << 23);
}

// ------------------------------------------------- minifloat round to nearest
// Round to nearest, ties to even, of v / 2^scale_exp to a minifloat whose
// codes count up monotonically with value.  Normal codes: the float's
// exponent and top mantissa bits, rounded in the integer domain (adding
// half a quantum minus one plus the kept lowest bit carries exactly when the
// dropped bits are above half, or at half with an odd kept bit; a carry walks
// into the next binade), then rebased to the minifloat's exponent bias.
// Subnormal codes: |v| as an integer multiple of the smallest subnormal,
// rounded by rintf (ties to even; the product by a power of two is exact).
// The scaling by 2^-scale_exp is exact wherever the result can be non-zero:
// a scaled value below 2^-126 (where a float product could round) is far
// below half the smallest subnormal of every format here, so its code is 0.
MSQ_HD float scale_pow2(float v, int scale_exp) {
  return scale_exp ? ldexp_exact(v, -scale_exp) : v;
}

// E4M3FN of v / 2^scale_exp (scale_exp = 0 for the scalar codec): above 448
// (infinity included) saturates to 0x7e, NaN is 0x7f, the sign is kept.
MSQ_HD uint8_t encode_e4m3(float v, int scale_exp = 0) {
  const uint32_t sign = (f32_bits(v) >> 24) & 0x80u;
  const uint32_t a = f32_bits(scale_pow2(v, scale_exp)) & 0x7fffffffu;
  const uint32_t sub = a < 0x3c800000u ? a : 0x3c800000u;  // below 2^-6
  uint32_t code = a >= 0x3c800000u ? ((a + 0x7ffffu + ((a >> 20) & 1u)) >> 20) - 960u
                                   : static_cast<uint32_t>(rintf(bits_f32(sub) * 512.0f));
  if (a > 0x43e00000u) code = 0x7eu;
  if (a > 0x7f800000u) code = 0x7fu;
  return static_cast<uint8_t>(sign | code);
}

// E5M2 of v / 2^scale_exp: finite values above 57344 saturate to 0x7b,
// infinity stays 0x7c, NaN is 0x7e.
MSQ_HD uint8_t encode_e5m2(float v, int scale_exp = 0) {
  const uint32_t bits = f32_bits(v);
  const uint32_t sign = (bits >> 24) & 0x80u;
  const uint32_t av = bits & 0x7fffffffu;
  const uint32_t a = f32_bits(scale_pow2(v, scale_exp)) & 0x7fffffffu;
  const uint32_t sub = a < 0x38800000u ? a : 0x38800000u;  // below 2^-14
  uint32_t code = a >= 0x38800000u ? ((a + 0xfffffu + ((a >> 21) & 1u)) >> 21) - 448u
                                   : static_cast<uint32_t>(rintf(bits_f32(sub) * 65536.0f));
  if (a > 0x47600000u) code = 0x7bu;
  if (av == 0x7f800000u) code = 0x7cu;
  if (av > 0x7f800000u) code = 0x7eu;
  return static_cast<uint8_t>(sign | code);
}

// FP4 E2M1 of v / 2^scale_exp (a nibble in the low four bits): above 6
// (infinity included) saturates to 0x7, NaN encodes as a signed zero.
MSQ_HD uint8_t encode_e2m1(float v, int scale_exp = 0) {
  const uint32_t bits = f32_bits(v);
  const uint32_t sign = (bits >> 28) & 0x8u;
  const uint32_t a = f32_bits(scale_pow2(v, scale_exp)) & 0x7fffffffu;
  const uint32_t sub = a < 0x3f800000u ? a : 0x3f800000u;  // below 1
  uint32_t code = a >= 0x3f800000u ? ((a + 0x1fffffu + ((a >> 22) & 1u)) >> 22) - 252u
                                   : static_cast<uint32_t>(rintf(bits_f32(sub) * 2.0f));
  if (a > 0x40c00000u) code = 0x7u;
  if ((bits & 0x7fffffffu) > 0x7f800000u) code = 0u;
  return static_cast<uint8_t>(sign | code);
}

// ------------------------------------------------------------------ decoders
MSQ_HD float decode_e4m3(uint8_t code) {
  const uint32_t sign = static_cast<uint32_t>(code & 0x80u) << 24;
  const uint32_t e = (code >> 3) & 0xfu;
  const uint32_t f = code & 0x7u;
  if (e == 0xfu && f == 0x7u) return bits_f32(sign | 0x7fc00000u);
  if (e == 0) {
    // f * 2^-9, exact
    return bits_f32(sign | f32_bits(static_cast<float>(f) * 0.001953125f));
  }
  return bits_f32(sign | ((e + 120u) << 23) | (f << 20));
}
MSQ_HD float decode_e5m2(uint8_t code) {
  const uint32_t sign = static_cast<uint32_t>(code & 0x80u) << 24;
  const uint32_t e = (code >> 2) & 0x1fu;
  const uint32_t f = code & 0x3u;
  if (e == 0x1fu) {
    return bits_f32(sign | (f == 0 ? 0x7f800000u : 0x7fc00000u));
  }
  if (e == 0) {
    return bits_f32(sign | f32_bits(static_cast<float>(f) * 1.52587890625e-05f));
  }
  return bits_f32(sign | ((e + 112u) << 23) | (f << 21));
}
MSQ_HD float decode_e2m1(uint8_t nibble) {
  const uint32_t sign = static_cast<uint32_t>(nibble & 0x8u) << 28;
  const uint32_t e = (nibble >> 1) & 0x3u;
  const uint32_t f = nibble & 0x1u;
  if (e == 0) return bits_f32(sign | (f ? 0x3f000000u : 0u));  // 0 or 0.5
  return bits_f32(sign | ((e + 126u) << 23) | (f << 22));
}

// MX scale exponent from the largest finite magnitude's bit pattern (bit
// patterns of non-negative floats order like their values).
MSQ_HD int mx_scale_exponent(uint32_t max_abs_bits, int emax) {
  if (max_abs_bits == 0) return 0;
  int ilogb;
  const uint32_t ef = max_abs_bits >> 23;
  if (ef == 0) {
    ilogb = (31 - __builtin_clz(max_abs_bits & 0x7fffffu)) - 149;
  } else {
    ilogb = static_cast<int>(ef) - 127;
  }
  int e = ilogb - emax;
  return e < -127 ? -127 : (e > 128 ? 128 : e);
}
// |v| bits if finite, else 0 (non-finite values do not enter the block max).
MSQ_HD uint32_t finite_abs_bits(float v) {
  const uint32_t a = f32_bits(v) & 0x7fffffffu;
  return a < 0x7f800000u ? a : 0u;
}

// ------------------------------------------------------ stochastic rounding
// The reference's stochastic mode (metal_mixed_sqrt.metal:394-405, :462-483,
// :545-567, :740-762, :1888-1910): a value between two adjacent codes rounds
// up with probability (|v| - lower) / (upper - lower), decided by a counter
// hash of (element index, seed, stream); values the nearest mode maps
// exactly, zeros, NaN and the saturating range use the nearest-mode encoder.
// The probability is exact in float (the numerator is exact by Sterbenz's
// lemma or is |v| itself, the denominator a power of two), so the decision is
// the same on every IEEE implementation.
MSQ_HD uint32_t stochastic_hash(uint32_t index, uint32_t seed, uint32_t stream) {
  const uint32_t state = index * 747796405u + seed + stream * 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}
MSQ_HD bool stochastic_round_up(float probability, uint32_t index, uint32_t seed,
                                uint32_t stream) {
  // 24 hash bits as a float in [0, 1): exact.
  const float sample =
      static_cast<float>(stochastic_hash(index, seed, stream) >> 8u) * 5.9604644775390625e-08f;
  return sample < probability;
}

MSQ_HD uint16_t encode_bf16_stochastic(float value, uint32_t index, uint32_t seed,
                                       uint32_t stream) {
  const uint32_t bits = f32_bits(value);
  if ((bits & 0x7f800000u) == 0x7f800000u || (bits & 0x7fffffffu) == 0u)
    return encode_bf16(value);
  const uint32_t discarded = bits & 0xffffu;
  uint32_t rounded = bits & 0xffff0000u;
  if ((stochastic_hash(index, seed, stream) & 0xffffu) < discarded) rounded += 0x00010000u;
  return static_cast<uint16_t>(rounded >> 16);
}

MSQ_HD uint8_t encode_e4m3_stochastic(float value, uint32_t index, uint32_t seed,
                                      uint32_t stream) {
  const uint32_t bits = f32_bits(value);
  const uint8_t sign = (bits >> 31) ? 0x80u : 0u;
  const float mag = bits_f32(bits & 0x7fffffffu);
  if (!(mag == mag) || mag == 0.0f || mag >= 448.0f) return encode_e4m3(value);
  const uint8_t nearest = encode_e4m3(mag);
  const float nv = decode_e4m3(nearest);
  if (nv == mag) return sign | nearest;
  const uint8_t lower = nv < mag ? nearest : static_cast<uint8_t>(nearest - 1u);
  const uint8_t upper = static_cast<uint8_t>(lower + 1u);
  const float lv = decode_e4m3(lower), uv = decode_e4m3(upper);
  const float p = (mag - lv) / (uv - lv);
  return sign | (stochastic_round_up(p, index, seed, stream) ? upper : lower);
}

MSQ_HD uint8_t encode_e5m2_stochastic(float value, uint32_t index, uint32_t seed,
                                      uint32_t stream) {
  const uint32_t bits = f32_bits(value);
  const uint8_t sign = (bits >> 31) ? 0x80u : 0u;
  const float mag = bits_f32(bits & 0x7fffffffu);
  if (!(mag == mag) || mag == 0.0f || mag >= 57344.0f) return encode_e5m2(value);  // inf too
  const uint8_t nearest = encode_e5m2(mag);
  const float nv = decode_e5m2(nearest);
  if (nv == mag) return sign | nearest;
  const uint8_t lower = nv < mag ? nearest : static_cast<uint8_t>(nearest - 1u);
  const uint8_t upper = static_cast<uint8_t>(lower + 1u);
  const float lv = decode_e5m2(lower), uv = decode_e5m2(upper);
  const float p = (mag - lv) / (uv - lv);
  return sign | (stochastic_round_up(p, index, seed, stream) ? upper : lower);
}

// FP4 E2M1 nibble.
MSQ_HD uint8_t encode_e2m1_stochastic(float value, uint32_t index, uint32_t seed,
                                      uint32_t stream) {
  const uint32_t bits = f32_bits(value);
  const uint8_t sign = (bits >> 31) ? 0x8u : 0u;
  const float mag = bits_f32(bits & 0x7fffffffu);
  if (!(mag == mag) || mag == 0.0f || mag >= 6.0f) return encode_e2m1(value);  // inf too
  uint8_t lower = 0;
  while (lower < 7u && decode_e2m1(static_cast<uint8_t>(lower + 1u)) <= mag) ++lower;
  const float lv = decode_e2m1(lower);
  if (lv == mag || lower == 7u) return sign | lower;
  const uint8_t upper = static_cast<uint8_t>(lower + 1u);
  const float uv = decode_e2m1(upper);
  const float p = (mag - lv) / (uv - lv);
  return sign | (stochastic_round_up(p, index, seed, stream) ? upper : lower);
}

// 2^e as a float, inf for e = 128 (the MX scale byte 255).
MSQ_HD float pow2f(int e) { return ldexp_exact(1.0f, e); }

// ------------------------------------------------------ log-affine spectrum
// The reference's log-affine spectrum storage for FP8 E4M3FN and FP4 E2M1
// (precisions *_log_s; metal_mixed_sqrt.metal:1689-1817, blocks of 8): per
// block the float pair (min log2 s, span), span the block's log2 range
// rounded down by one ulp when min + span would exceed the block maximum;
// code = encode(M (log2 s - min) / span), M = 448 (E4M3) or 6 (E2M1), 0 for a
// constant block; decode exp2(min + span * code / M).  exp2 is evaluated in
// binary64 and rounded to float; log2 in binary64 and rounded toward -inf, so
// that exp2 of a block's largest log2 never exceeds its largest value (to
// nearest, log2 of the largest float is 128 and the top code would decode to
// infinity; the reference's one-ulp step of the span aims at the same bound).
constexpr int kLogBlock = 8;
MSQ_HD float log2_f32(float x) {
  const double l = ::log2(static_cast<double>(x));
  float f = static_cast<float>(l);
  if (static_cast<double>(f) > l) f = bits_f32(f > 0.0f ? f32_bits(f) - 1u : f32_bits(f) + 1u);
  return f;
}
MSQ_HD float exp2_f32(float y) { return static_cast<float>(::exp2(static_cast<double>(y))); }
// The next float toward zero (span > 0, finite).
MSQ_HD float toward_zero_f32(float v) {
  const uint32_t b = f32_bits(v);
  return (b & 0x7fffffffu) == 0u ? v : bits_f32(b - 1u);
}
// Metadata (min, span) of the block of values[begin, end).
MSQ_HD void log_block_metadata(const float* values, uint32_t begin, uint32_t end, float& minimum,
                               float& span) {
  float mn = INFINITY, mx = -INFINITY;
  for (uint32_t k = begin; k < end; ++k) {
    const float l = log2_f32(values[k]);
    mn = l < mn ? l : mn;  // the reference's min / max (NaN-ignoring as fmin / fmax)
    mx = l > mx ? l : mx;
  }
  float sp = mx - mn;
  if (mn + sp > mx) sp = toward_zero_f32(sp);
  minimum = mn;
  span = sp;
}
MSQ_HD float log_normalized(float value, float minimum, float span, float maximum) {
  return span == 0.0f ? 0.0f : maximum * (log2_f32(value) - minimum) / span;
}
MSQ_HD float log_decode(float decoded_code, float minimum, float span, float maximum) {
  return exp2_f32(minimum + span * decoded_code / maximum);
}

}  // namespace msq
