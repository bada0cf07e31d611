/*
 * CPU oracle for the storage codecs (test infrastructure only).
 *
 * This file restates, in plain C, the host codecs of the reference
 * (src/storage_codecs.cpp).  It is the checker that the HIP codec kernels are
 * compared against bit for bit; it is never linked into the product library.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Each function follows the reference algorithm, not a faster equivalent:
 * encoding is a brute-force scan for the nearest finite code with ties going to
 * the even code (storage_codecs.cpp:79-94), with the reference's explicit
 * saturation and NaN rules (:96-136); the MX scale uses ilogb of the block's
 * largest finite magnitude (:148-162).  The device kernels use integer
 * round-to-nearest-even instead, so agreement between the two is a real check.
 *
 * Pinned by: the reference's own known answers (storage_codecs.cpp:200-247,
 * metal_checks.cpp:195-203 and :286-292, tests/matrix_codec.cpp:59-76),
 * torch's float8/bfloat16 casts on in-range values, and exhaustive decode
 * tables checked against the closed-form format definitions.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <stddef.h>

/* storage_codecs.cpp:16-28 */
uint16_t oracle_float_to_bf16(float value) {
  uint32_t bits;
  memcpy(&bits, &value, sizeof(bits));
  uint32_t rounded = bits >> 16;
  const uint32_t remainder = bits & 0xffffu;
  if (remainder > 0x8000u || (remainder == 0x8000u && (rounded & 1u) != 0)) {
    ++rounded;
  }
  if ((bits & 0x7f800000u) == 0x7f800000u && (bits & 0x007fffffu) != 0) {
    rounded |= 0x0040u;
  }
  return (uint16_t)rounded;
}

/* storage_codecs.cpp:30-35 */
float oracle_bf16_to_float(uint16_t value) {
  const uint32_t bits = (uint32_t)value << 16;
  float result;
  memcpy(&result, &bits, sizeof(result));
  return result;
}

/* storage_codecs.cpp:37-49 */
float oracle_fp8_e4m3fn_to_float(uint8_t value) {
  const float sign = (value & 0x80u) == 0 ? 1.0f : -1.0f;
  const uint8_t exponent = (value >> 3) & 0x0fu;
  const uint8_t fraction = value & 0x07u;
  if (exponent == 0x0fu && fraction == 0x07u) {
    return copysignf(NAN, sign);
  }
  const float magnitude =
      exponent == 0 ? ldexpf((float)fraction, -9)
                    : ldexpf(1.0f + (float)fraction / 8.0f, (int)exponent - 7);
  return copysignf(magnitude, sign);
}

/* storage_codecs.cpp:51-66 */
float oracle_fp8_e5m2_to_float(uint8_t value) {
  const float sign = (value & 0x80u) == 0 ? 1.0f : -1.0f;
  const uint8_t exponent = (value >> 2) & 0x1fu;
  const uint8_t fraction = value & 0x03u;
  if (exponent == 0x1fu) {
    if (fraction == 0) {
      return copysignf(INFINITY, sign);
    }
    return copysignf(NAN, sign);
  }
  const float magnitude =
      exponent == 0 ? ldexpf((float)fraction, -16)
                    : ldexpf(1.0f + (float)fraction / 4.0f, (int)exponent - 15);
  return copysignf(magnitude, sign);
}

/* storage_codecs.cpp:68-77 */
float oracle_fp4_e2m1_to_float(uint8_t value) {
  const float sign = (value & 0x08u) == 0 ? 1.0f : -1.0f;
  const uint8_t exponent = (value >> 1) & 0x03u;
  const uint8_t fraction = value & 0x01u;
  const float magnitude =
      exponent == 0 ? (float)fraction * 0.5f
                    : ldexpf(1.0f + (float)fraction * 0.5f, (int)exponent - 1);
  return copysignf(magnitude, sign);
}

typedef float (*decode_fn)(uint8_t);

/* storage_codecs.cpp:79-94: brute-force nearest code, ties to the even code. */
static uint8_t nearest_finite_code(float magnitude, uint8_t maximum_code,
                                   decode_fn decode) {
  uint8_t nearest = 0;
  float nearest_distance = fabsf(magnitude - decode(nearest));
  for (uint16_t code = 1; code <= maximum_code; ++code) {
    const uint8_t candidate = (uint8_t)code;
    const float distance = fabsf(magnitude - decode(candidate));
    if (distance < nearest_distance ||
        (distance == nearest_distance && (candidate & 1u) == 0 &&
         (nearest & 1u) != 0)) {
      nearest = candidate;
      nearest_distance = distance;
    }
  }
  return nearest;
}

/* storage_codecs.cpp:96-108 */
uint8_t oracle_float_to_fp8_e4m3fn(float value) {
  const uint8_t sign = signbit(value) ? 0x80u : 0u;
  if (isnan(value)) return sign | 0x7fu;
  if (value == 0.0f) return sign;
  if (isinf(value) || fabsf(value) > 448.0f) return sign | 0x7eu;
  return sign | nearest_finite_code(fabsf(value), 0x7eu,
                                    oracle_fp8_e4m3fn_to_float);
}

/* storage_codecs.cpp:110-125 */
uint8_t oracle_float_to_fp8_e5m2(float value) {
  const uint8_t sign = signbit(value) ? 0x80u : 0u;
  if (isnan(value)) return sign | 0x7eu;
  if (value == 0.0f) return sign;
  if (isinf(value)) return sign | 0x7cu;
  if (fabsf(value) > 57344.0f) return sign | 0x7bu;
  return sign | nearest_finite_code(fabsf(value), 0x7bu,
                                    oracle_fp8_e5m2_to_float);
}

/* storage_codecs.cpp:127-136 */
uint8_t oracle_float_to_fp4_e2m1(float value) {
  const uint8_t sign = signbit(value) ? 0x08u : 0u;
  if (isnan(value) || value == 0.0f) return sign;
  if (isinf(value) || fabsf(value) > 6.0f) return sign | 0x07u;
  return sign | nearest_finite_code(fabsf(value), 0x07u,
                                    oracle_fp4_e2m1_to_float);
}

/* storage_codecs.cpp:148-162 */
static int mx_scale_exponent(const float *values, size_t begin, size_t end,
                             int element_maximum_exponent) {
  float maximum = 0.0f;
  for (size_t index = begin; index < end; ++index) {
    const float value = values[index];
    if (isfinite(value)) {
      maximum = fmaxf(maximum, fabsf(value));
    }
  }
  if (maximum == 0.0f) return 0;
  int exponent = ilogbf(maximum) - element_maximum_exponent;
  if (exponent < -127) exponent = -127;
  if (exponent > 128) exponent = 128;
  return exponent;
}

/* Scalar codecs over arrays (copyValuesToStorage, storage_codecs.cpp:249-313). */
void oracle_encode_bf16(const float *src, uint16_t *dst, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] = oracle_float_to_bf16(src[i]);
}
void oracle_encode_fp8_e4m3fn(const float *src, uint8_t *dst, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] = oracle_float_to_fp8_e4m3fn(src[i]);
}
void oracle_encode_fp8_e5m2(const float *src, uint8_t *dst, size_t n) {
  for (size_t i = 0; i < n; ++i) dst[i] = oracle_float_to_fp8_e5m2(src[i]);
}

/* packFp4, storage_codecs.cpp:138-146: low nibble holds the even index. */
void oracle_encode_fp4_e2m1(const float *src, uint8_t *dst, size_t n) {
  for (size_t i = 0; i < n; i += 2) {
    const uint8_t lower = oracle_float_to_fp4_e2m1(src[i]);
    const uint8_t upper = i + 1 < n ? oracle_float_to_fp4_e2m1(src[i + 1]) : 0u;
    dst[i / 2] = (uint8_t)(lower | (uint8_t)(upper << 4));
  }
}

/* packMxfp8, storage_codecs.cpp:164-179; emax is 8 (E4M3) or 15 (E5M2). */
static void encode_mxfp8(const float *src, uint8_t *dst, uint8_t *scales,
                         size_t n, int emax, uint8_t (*encode)(float)) {
  for (size_t begin = 0, block = 0; begin < n; begin += 32, ++block) {
    const size_t end = begin + 32 < n ? begin + 32 : n;
    const int exponent = mx_scale_exponent(src, begin, end, emax);
    scales[block] = (uint8_t)(exponent + 127);
    const float scale = ldexpf(1.0f, exponent);
    for (size_t i = begin; i < end; ++i) dst[i] = encode(src[i] / scale);
  }
}
void oracle_encode_mxfp8_e4m3fn(const float *src, uint8_t *dst, uint8_t *scales,
                                size_t n) {
  encode_mxfp8(src, dst, scales, n, 8, oracle_float_to_fp8_e4m3fn);
}
void oracle_encode_mxfp8_e5m2(const float *src, uint8_t *dst, uint8_t *scales,
                              size_t n) {
  encode_mxfp8(src, dst, scales, n, 15, oracle_float_to_fp8_e5m2);
}

/* packMxfp4, storage_codecs.cpp:181-198 */
void oracle_encode_mxfp4_e2m1(const float *src, uint8_t *dst, uint8_t *scales,
                              size_t n) {
  for (size_t begin = 0, block = 0; begin < n; begin += 32, ++block) {
    const size_t end = begin + 32 < n ? begin + 32 : n;
    const int exponent = mx_scale_exponent(src, begin, end, 2);
    scales[block] = (uint8_t)(exponent + 127);
    const float scale = ldexpf(1.0f, exponent);
    for (size_t i = begin; i < end; i += 2) {
      const uint8_t lower = oracle_float_to_fp4_e2m1(src[i] / scale);
      const uint8_t upper =
          i + 1 < end ? oracle_float_to_fp4_e2m1(src[i + 1] / scale) : 0u;
      dst[i / 2] = (uint8_t)(lower | (uint8_t)(upper << 4));
    }
  }
}

/* storageValueAt, storage_codecs.cpp:323-374, over whole arrays.
 * format: 1 bf16, 2 fp8_e4m3fn, 3 fp8_e5m2, 4 mxfp8_e4m3fn, 5 mxfp8_e5m2,
 *         6 fp4_e2m1, 7 mxfp4_e2m1 (the HostStorageCodec order). */
int oracle_decode(int format, const uint8_t *values, const uint8_t *scales,
                  float *dst, size_t n) {
  for (size_t i = 0; i < n; ++i) {
    switch (format) {
      case 1: {
        uint16_t v;
        memcpy(&v, values + 2 * i, 2);
        dst[i] = oracle_bf16_to_float(v);
        break;
      }
      case 2: dst[i] = oracle_fp8_e4m3fn_to_float(values[i]); break;
      case 3: dst[i] = oracle_fp8_e5m2_to_float(values[i]); break;
      case 4:
        dst[i] = ldexpf(oracle_fp8_e4m3fn_to_float(values[i]),
                        (int)scales[i / 32] - 127);
        break;
      case 5:
        dst[i] = ldexpf(oracle_fp8_e5m2_to_float(values[i]),
                        (int)scales[i / 32] - 127);
        break;
      case 6:
        dst[i] = oracle_fp4_e2m1_to_float((values[i / 2] >> ((i & 1u) * 4)) &
                                          0x0fu);
        break;
      case 7:
        dst[i] = ldexpf(oracle_fp4_e2m1_to_float((values[i / 2] >>
                                                  ((i & 1u) * 4)) & 0x0fu),
                        (int)scales[i / 32] - 127);
        break;
      default:
        return -1;
    }
  }
  return 0;
}

/*
 * Per-tensor amax-scaled FP8 E4M3 (no counterpart in the reference, whose FP8
 * is unscaled, src/precision.hpp:69-70).  The scale is
 *   sigma = (float)(448.0 / (double)amax)  over the finite elements,
 * 1 when amax is 0; each value is encoded as E4M3(value * sigma) with the
 * reference E4M3 codec above, and decodes as E4M3^-1(code) / sigma.
 */
float oracle_fp8_amax_scale(const float *src, size_t n) {
  float amax = 0.0f;
  for (size_t i = 0; i < n; ++i) {
    if (isfinite(src[i])) amax = fmaxf(amax, fabsf(src[i]));
  }
  if (amax == 0.0f) return 1.0f;
  return (float)(448.0 / (double)amax);
}
void oracle_encode_fp8_e4m3fn_scaled(const float *src, uint8_t *dst, size_t n,
                                     float sigma) {
  for (size_t i = 0; i < n; ++i) {
    dst[i] = oracle_float_to_fp8_e4m3fn(src[i] * sigma);
  }
}

/* ------------------------------------------------------ stochastic rounding
 * The reference's stochastic quantizers (metal_mixed_sqrt.metal:394-405 hash
 * and round-up test; :462-483 E4M3FN, :545-567 E5M2, :740-762 E2M1,
 * :1888-1910 BF16; the MX kernels :972-1010, :1114-1140 and :1233-1259 apply
 * the scalar ones to value / 2^scale).  The neighbouring codes are found with
 * the nearest-code search above, as the kernels find them with their own
 * nearest encoders.  Element index i of the array is the hash's index.
 */
uint32_t oracle_stochastic_hash(uint32_t index, uint32_t seed, uint32_t stream) {
  const uint32_t state = index * 747796405u + seed + stream * 2891336453u;
  const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

static int round_up(float probability, uint32_t index, uint32_t seed, uint32_t stream) {
  const float sample =
      (float)(oracle_stochastic_hash(index, seed, stream) >> 8u) * 0x1.0p-24f;
  return sample < probability;
}

typedef uint8_t (*encode_fn)(float);

/* Shared shape of the FP8 stochastic encoders (:462-483, :545-567). */
static uint8_t fp8_stochastic(float value, float saturation, encode_fn encode,
                              decode_fn decode, uint32_t index, uint32_t seed,
                              uint32_t stream, int inf_is_special) {
  const uint8_t sign = signbit(value) ? 0x80u : 0u;
  const float magnitude = fabsf(value);
  if (isnan(value) || (inf_is_special && isinf(value)) || magnitude == 0.0f ||
      magnitude >= saturation)
    return encode(value);
  const uint8_t nearest = encode(magnitude);
  const float nearest_value = decode(nearest);
  if (nearest_value == magnitude) return sign | nearest;
  const uint8_t lower = nearest_value < magnitude ? nearest : (uint8_t)(nearest - 1u);
  const uint8_t upper = (uint8_t)(lower + 1u);
  const float lower_value = decode(lower);
  const float upper_value = decode(upper);
  const float probability = (magnitude - lower_value) / (upper_value - lower_value);
  return sign | (round_up(probability, index, seed, stream) ? upper : lower);
}

uint8_t oracle_e4m3_stochastic(float value, uint32_t index, uint32_t seed, uint32_t stream) {
  return fp8_stochastic(value, 448.0f, oracle_float_to_fp8_e4m3fn, oracle_fp8_e4m3fn_to_float,
                        index, seed, stream, 0);
}
uint8_t oracle_e5m2_stochastic(float value, uint32_t index, uint32_t seed, uint32_t stream) {
  return fp8_stochastic(value, 57344.0f, oracle_float_to_fp8_e5m2, oracle_fp8_e5m2_to_float,
                        index, seed, stream, 1);
}
uint8_t oracle_e2m1_stochastic(float value, uint32_t index, uint32_t seed, uint32_t stream) {
  const uint8_t sign = signbit(value) ? 0x08u : 0u;
  const float magnitude = fabsf(value);
  if (isnan(value) || magnitude == 0.0f || magnitude >= 6.0f)
    return oracle_float_to_fp4_e2m1(value);
  uint8_t lower = 0;
  while (lower < 0x07u && oracle_fp4_e2m1_to_float((uint8_t)(lower + 1u)) <= magnitude) ++lower;
  const float lower_value = oracle_fp4_e2m1_to_float(lower);
  if (lower_value == magnitude || lower == 0x07u) return sign | lower;
  const uint8_t upper = (uint8_t)(lower + 1u);
  const float upper_value = oracle_fp4_e2m1_to_float(upper);
  const float probability = (magnitude - lower_value) / (upper_value - lower_value);
  return sign | (round_up(probability, index, seed, stream) ? upper : lower);
}
uint16_t oracle_bf16_stochastic(float value, uint32_t index, uint32_t seed, uint32_t stream) {
  if (!isfinite(value) || value == 0.0f) return oracle_float_to_bf16(value);
  uint32_t bits;
  memcpy(&bits, &value, sizeof(bits));
  const uint32_t discarded = bits & 0xffffu;
  uint32_t rounded = bits & 0xffff0000u;
  if ((oracle_stochastic_hash(index, seed, stream) & 0xffffu) < discarded) rounded += 0x00010000u;
  return (uint16_t)(rounded >> 16);
}

/* format: the oracle_decode ids (1 bf16 .. 7 mxfp4_e2m1). */
int oracle_encode_stochastic(int format, const float *src, uint8_t *dst, uint8_t *scales,
                             size_t n, uint32_t seed, uint32_t stream) {
  const int mx = format == 4 || format == 5 || format == 7;
  for (size_t begin = 0, block = 0; begin < n; begin += 32, ++block) {
    const size_t end = begin + 32 < n ? begin + 32 : n;
    int e = 0;
    if (mx) {
      e = mx_scale_exponent(src, begin, end, format == 4 ? 8 : (format == 5 ? 15 : 2));
      scales[block] = (uint8_t)(e + 127);
    }
    const float scale = ldexpf(1.0f, e);
    for (size_t i = begin; i < end; ++i) {
      const float x = mx ? src[i] / scale : src[i];
      const uint32_t idx = (uint32_t)i;
      switch (format) {
        case 1: {
          const uint16_t c = oracle_bf16_stochastic(x, idx, seed, stream);
          memcpy(dst + 2 * i, &c, 2);
          break;
        }
        case 2:
        case 4: dst[i] = oracle_e4m3_stochastic(x, idx, seed, stream); break;
        case 3:
        case 5: dst[i] = oracle_e5m2_stochastic(x, idx, seed, stream); break;
        case 6:
        case 7: {
          const uint8_t c = oracle_e2m1_stochastic(x, idx, seed, stream);
          if ((i & 1u) == 0) dst[i / 2] = c;
          else dst[i / 2] = (uint8_t)(dst[i / 2] | (c << 4));
          break;
        }
        default: return -1;
      }
    }
  }
  return 0;
}

/* ----------------------------------------------------- log-affine spectrum
 * The reference's GPU log-spectrum codec (log_spectrum_metadata,
 * quantize_*_log_s and decode_*_log_s_to_f32, metal_mixed_sqrt.metal:
 * 1689-1817): per block of 8 elements the pair (min log2 v, span), the span
 * one ulp toward zero when min + span exceeds the block maximum; code =
 * encode(M (log2 v - min) / span) (0 for a constant block), M = 448 for E4M3FN
 * and 6 for E2M1; decode exp2(min + span * code / M).  exp2 is the binary64
 * function rounded to float, log2 the binary64 function rounded toward -inf.
 */
/* log2 rounded toward -inf: the float below the binary64 log2 when rounding
 * to nearest went up, so exp2 of a block's largest log never exceeds its
 * largest value (log2 of FLT_MAX rounds up to 128 to nearest). */
static float log2_f32(float x) {
  const double l = log2((double)x);
  float f = (float)l;
  if ((double)f > l) f = nextafterf(f, -INFINITY);
  return f;
}
static float exp2_f32(float y) { return (float)exp2((double)y); }

void oracle_encode_log(int e4m3, const float *src, uint8_t *dst, float *metadata, size_t n,
                       int stochastic, uint32_t seed, uint32_t stream) {
  const float maximum = e4m3 ? 448.0f : 6.0f;
  for (size_t begin = 0, block = 0; begin < n; begin += 8, ++block) {
    const size_t end = begin + 8 < n ? begin + 8 : n;
    float minimum = INFINITY, maximum_log = -INFINITY;
    for (size_t i = begin; i < end; ++i) {
      const float l = log2_f32(src[i]);
      minimum = fminf(minimum, l);
      maximum_log = fmaxf(maximum_log, l);
    }
    float span = maximum_log - minimum;
    if (minimum + span > maximum_log) span = nextafterf(span, 0.0f);
    metadata[2 * block] = minimum;
    metadata[2 * block + 1] = span;
    for (size_t i = begin; i < end; ++i) {
      const float x = span == 0.0f ? 0.0f : maximum * (log2_f32(src[i]) - minimum) / span;
      uint8_t c;
      if (e4m3)
        c = stochastic ? oracle_e4m3_stochastic(x, (uint32_t)i, seed, stream)
                       : oracle_float_to_fp8_e4m3fn(x);
      else
        c = stochastic ? oracle_e2m1_stochastic(x, (uint32_t)i, seed, stream)
                       : oracle_float_to_fp4_e2m1(x);
      if (e4m3) dst[i] = c;
      else if ((i & 1u) == 0) dst[i / 2] = c;
      else dst[i / 2] = (uint8_t)(dst[i / 2] | (c << 4));
    }
  }
}

void oracle_decode_log(int e4m3, const uint8_t *codes, const float *metadata, float *dst,
                       size_t n) {
  const float maximum = e4m3 ? 448.0f : 6.0f;
  for (size_t i = 0; i < n; ++i) {
    const size_t block = i / 8;
    const float d = e4m3 ? oracle_fp8_e4m3fn_to_float(codes[i])
                         : oracle_fp4_e2m1_to_float((codes[i / 2] >> ((i & 1u) * 4)) & 0x0fu);
    dst[i] = exp2_f32(metadata[2 * block] + metadata[2 * block + 1] * d / maximum);
  }
}
