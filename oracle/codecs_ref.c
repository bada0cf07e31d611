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
