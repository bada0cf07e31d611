// Host build of the device codecs (mixed-mlx_amd/csrc/device_codecs.hpp)
// checked against the C oracle (oracle/codecs_ref.c) over float bit patterns.
// Built and run by tests/test_codecs_host.py; prints "ok <count>" or the first
// mismatch.
//
// A second, independent formulation of the minifloat encoders (integer
// round-half-even of the significand: floor(log2), shift, compare the dropped
// bits with half) is pinned against the oracle here too, and with "full" as
// the third argument the library's encoders are compared with it on every
// float bit pattern (2^32, on all host threads) and, for the MX scaled forms,
// on every scale exponent over a dense sample.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
#include <string>
#include <thread>
#include <vector>

#include "device_codecs.hpp"

extern "C" {
uint16_t oracle_float_to_bf16(float);
uint8_t oracle_float_to_fp8_e4m3fn(float);
uint8_t oracle_float_to_fp8_e5m2(float);
uint8_t oracle_float_to_fp4_e2m1(float);
float oracle_fp8_e4m3fn_to_float(uint8_t);
float oracle_fp8_e5m2_to_float(uint8_t);
float oracle_fp4_e2m1_to_float(uint8_t);
float oracle_bf16_to_float(uint16_t);
}

namespace {
using msq::f32_bits;
// Rounds the finite positive value sig * 2^e2 (sig < 2^24) to the code of a
// minifloat with `mant` mantissa bits and exponent bias `bias`, where codes
// count up monotonically with value: code = ((E + bias - 1) << mant) + k with
// E = max(floor(log2 value), 1 - bias) and k = round_half_even(value /
// 2^(E - mant)).  A carry out of k walks into the next binade correctly.
static uint32_t minifloat_round(uint32_t sig, int e2, int mant, int bias) {
  int msb = 31 - __builtin_clz(sig);       // sig != 0
  int log2v = e2 + msb;
  int E = log2v > 1 - bias ? log2v : 1 - bias;
  int shift = (E - mant) - e2;  // value / 2^(E - mant) = sig / 2^shift
  uint32_t k;
  if (shift <= 0) {
    k = sig << (-shift);
  } else if (shift > 25) {
    k = 0;  // value < 2^(E - mant - 1): below half the quantum
  } else {
    k = sig >> shift;
    const uint32_t rem = sig & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1);
    if (rem > half || (rem == half && (k & 1u))) ++k;
  }
  return (static_cast<uint32_t>(E + bias - 1) << mant) + k;
}

// Splits |v| (finite, non-zero) into sig * 2^e2, applying an extra power of
// two `scale_exp` exactly (value / 2^scale_exp).
static void split_abs(uint32_t bits, int scale_exp, uint32_t &sig, int &e2) {
  const uint32_t ef = (bits >> 23) & 0xffu;
  const uint32_t mf = bits & 0x7fffffu;
  if (ef == 0) {
    sig = mf;
    e2 = -149;
  } else {
    sig = mf | 0x800000u;
    e2 = static_cast<int>(ef) - 150;
  }
  e2 -= scale_exp;
}

// E4M3FN of v / 2^scale_exp (scale_exp = 0 for the scalar codec).
static uint8_t generic_e4m3(float v, int scale_exp = 0) {
  const uint32_t bits = f32_bits(v);
  const uint8_t sign = (bits >> 31) ? 0x80u : 0u;
  const uint32_t a = bits & 0x7fffffffu;
  if (a > 0x7f800000u) return sign | 0x7fu;   // NaN
  if (a == 0) return sign;                     // +-0
  if (a == 0x7f800000u) return sign | 0x7eu;  // inf saturates
  uint32_t sig;
  int e2;
  split_abs(bits, scale_exp, sig, e2);
  // |v| / 2^s > 448  <=>  sig * 2^e2 > 7 * 2^6
  const int msb = 31 - __builtin_clz(sig);
  if (e2 + msb > 8) return sign | 0x7eu;
  if (e2 + msb == 8) {
    // value in [256, 512): compare with 448 = 0b111 << 6 exactly
    const int sh = msb - 2;  // keep the top three bits
    const uint32_t top = sig >> sh;
    const bool above = top > 7u || (top == 7u && (sig & ((1u << sh) - 1u)));
    if (above) return sign | 0x7eu;
  }
  return sign | static_cast<uint8_t>(minifloat_round(sig, e2, 3, 7));
}

// E5M2 of v / 2^scale_exp.
static uint8_t generic_e5m2(float v, int scale_exp = 0) {
  const uint32_t bits = f32_bits(v);
  const uint8_t sign = (bits >> 31) ? 0x80u : 0u;
  const uint32_t a = bits & 0x7fffffffu;
  if (a > 0x7f800000u) return sign | 0x7eu;   // NaN
  if (a == 0) return sign;
  if (a == 0x7f800000u) return sign | 0x7cu;  // inf stays inf
  uint32_t sig;
  int e2;
  split_abs(bits, scale_exp, sig, e2);
  // finite |v| / 2^s > 57344 = 0b111 << 13 saturates to 0x7b
  const int msb = 31 - __builtin_clz(sig);
  if (e2 + msb > 15) return sign | 0x7bu;
  if (e2 + msb == 15) {
    const int sh = msb - 2;
    const uint32_t top = sig >> sh;
    const bool above = top > 7u || (top == 7u && (sig & ((1u << sh) - 1u)));
    if (above) return sign | 0x7bu;
  }
  return sign | static_cast<uint8_t>(minifloat_round(sig, e2, 2, 15));
}

// FP4 E2M1 of v / 2^scale_exp (a nibble in the low four bits).
static uint8_t generic_e2m1(float v, int scale_exp = 0) {
  const uint32_t bits = f32_bits(v);
  const uint8_t sign = (bits >> 31) ? 0x8u : 0u;
  const uint32_t a = bits & 0x7fffffffu;
  if (a > 0x7f800000u || a == 0) return sign;  // NaN and zero
  if (a == 0x7f800000u) return sign | 0x7u;
  uint32_t sig;
  int e2;
  split_abs(bits, scale_exp, sig, e2);
  // > 6 = 0b11 << 1 saturates
  const int msb = 31 - __builtin_clz(sig);
  if (e2 + msb > 2) return sign | 0x7u;
  if (e2 + msb == 2) {
    const int sh = msb - 1;  // keep the top two bits
    const uint32_t top = sig >> sh;
    const bool above = top > 3u || (top == 3u && (sig & ((1u << sh) - 1u)));
    if (above) return sign | 0x7u;
  }
  return sign | static_cast<uint8_t>(minifloat_round(sig, e2, 1, 1));
}

}  // namespace

static float f(uint32_t u) { float x; std::memcpy(&x, &u, 4); return x; }
static uint32_t u(float x) { uint32_t v; std::memcpy(&v, &x, 4); return v; }

int main(int argc, char** argv) {
  const uint32_t stride = argc > 1 ? static_cast<uint32_t>(std::atoi(argv[1])) : 9973u;
  const uint32_t mstep = argc > 2 ? static_cast<uint32_t>(std::atoi(argv[2])) : 61u;
  uint64_t checked = 0;
  // Scalar encoders over a strided sweep of all bit patterns plus every
  // pattern near the formats' ranges (|x| in [2^-20, 2^17] step 2^k ulps).
  for (uint64_t b = 0; b <= 0xffffffffull; b += stride) {
    const float x = f(static_cast<uint32_t>(b));
    if (msq::encode_bf16(x) != oracle_float_to_bf16(x)) { printf("bf16 %08x\n", (unsigned)b); return 1; }
    if (msq::encode_e4m3(x) != oracle_float_to_fp8_e4m3fn(x)) { printf("e4m3 %08x\n", (unsigned)b); return 1; }
    if (msq::encode_e5m2(x) != oracle_float_to_fp8_e5m2(x)) { printf("e5m2 %08x\n", (unsigned)b); return 1; }
    if (msq::encode_e2m1(x) != oracle_float_to_fp4_e2m1(x)) { printf("e2m1 %08x\n", (unsigned)b); return 1; }
    if (generic_e4m3(x) != oracle_float_to_fp8_e4m3fn(x)) { printf("generic e4m3 %08x\n", (unsigned)b); return 1; }
    if (generic_e5m2(x) != oracle_float_to_fp8_e5m2(x)) { printf("generic e5m2 %08x\n", (unsigned)b); return 1; }
    if (generic_e2m1(x) != oracle_float_to_fp4_e2m1(x)) { printf("generic e2m1 %08x\n", (unsigned)b); return 1; }
    ++checked;
  }
  for (uint32_t e = 107; e <= 145; ++e) {       // 2^-20 .. 2^18
    for (uint32_t m = 0; m < (1u << 23); m += mstep) {
      for (uint32_t s = 0; s < 2; ++s) {
        const uint32_t b = (s << 31) | (e << 23) | m;
        const float x = f(b);
        if (msq::encode_e4m3(x) != oracle_float_to_fp8_e4m3fn(x)) { printf("e4m3 %08x\n", b); return 1; }
        if (msq::encode_e5m2(x) != oracle_float_to_fp8_e5m2(x)) { printf("e5m2 %08x\n", b); return 1; }
        if (msq::encode_e2m1(x) != oracle_float_to_fp4_e2m1(x)) { printf("e2m1 %08x\n", b); return 1; }
        if (generic_e4m3(x) != oracle_float_to_fp8_e4m3fn(x)) { printf("generic e4m3 %08x\n", b); return 1; }
        if (generic_e5m2(x) != oracle_float_to_fp8_e5m2(x)) { printf("generic e5m2 %08x\n", b); return 1; }
        if (generic_e2m1(x) != oracle_float_to_fp4_e2m1(x)) { printf("generic e2m1 %08x\n", b); return 1; }
        ++checked;
      }
    }
  }
  // Every tie point and its neighbours: midpoints between adjacent codes.
  for (int fmt = 0; fmt < 3; ++fmt) {
    const int maxc = fmt == 0 ? 0x7e : (fmt == 1 ? 0x7b : 0x7);
    for (int c = 0; c < maxc; ++c) {
      float lo, hi;
      if (fmt == 0) { lo = oracle_fp8_e4m3fn_to_float(c); hi = oracle_fp8_e4m3fn_to_float(c + 1); }
      else if (fmt == 1) { lo = oracle_fp8_e5m2_to_float(c); hi = oracle_fp8_e5m2_to_float(c + 1); }
      else { lo = oracle_fp4_e2m1_to_float(c); hi = oracle_fp4_e2m1_to_float(c + 1); }
      const float mid = 0.5f * (lo + hi);
      for (int d = -2; d <= 2; ++d) {
        const float x = f(u(mid) + d);
        uint8_t a, b;
        if (fmt == 0) { a = msq::encode_e4m3(x); b = oracle_float_to_fp8_e4m3fn(x); }
        else if (fmt == 1) { a = msq::encode_e5m2(x); b = oracle_float_to_fp8_e5m2(x); }
        else { a = msq::encode_e2m1(x); b = oracle_float_to_fp4_e2m1(x); }
        if (a != b) { printf("tie fmt %d code %d d %d\n", fmt, c, d); return 1; }
        ++checked;
      }
    }
  }
  // Decoders: every code.
  for (int c = 0; c < 256; ++c) {
    const float a = msq::decode_e4m3(c), b = oracle_fp8_e4m3fn_to_float(c);
    if (u(a) != u(b)) { printf("dec e4m3 %d\n", c); return 1; }
    const float a2 = msq::decode_e5m2(c), b2 = oracle_fp8_e5m2_to_float(c);
    if (u(a2) != u(b2)) { printf("dec e5m2 %d\n", c); return 1; }
  }
  for (int c = 0; c < 16; ++c)
    if (u(msq::decode_e2m1(c)) != u(oracle_fp4_e2m1_to_float(c))) { printf("dec e2m1 %d\n", c); return 1; }
  for (uint32_t c = 0; c < 65536; ++c)
    if (u(msq::decode_bf16(c)) != u(oracle_bf16_to_float(c))) { printf("dec bf16 %u\n", c); return 1; }
  // MX scaled encoders: v / 2^s for every exponent the MX scale can take, over
  // the elements a block with that scale can hold: |v| <= block max <
  // 2^(s + emax + 1) (emax 2 / 8 / 15 for E2M1 / E4M3 / E5M2).
  for (int s = -127; s <= 125; s += 1) {
    const float scale = std::ldexp(1.0f, s);
    for (uint64_t b = 0; b <= 0xffffffffull; b += 1000003ull) {
      const float x = f(static_cast<uint32_t>(b));
      if (!std::isfinite(x)) continue;
      const int lx = x == 0.0f ? -1000 : std::ilogb(x);
      if (lx <= s + 2 && msq::encode_e2m1(x, s) != oracle_float_to_fp4_e2m1(x / scale)) { printf("mx e2m1 %08x s %d\n", (unsigned)b, s); return 1; }
      if (lx <= s + 8 && msq::encode_e4m3(x, s) != oracle_float_to_fp8_e4m3fn(x / scale)) { printf("mx e4m3 %08x s %d\n", (unsigned)b, s); return 1; }
      if (lx <= s + 15 && msq::encode_e5m2(x, s) != oracle_float_to_fp8_e5m2(x / scale)) { printf("mx e5m2 %08x s %d\n", (unsigned)b, s); return 1; }
      if (lx <= s + 2 && generic_e2m1(x, s) != oracle_float_to_fp4_e2m1(x / scale)) { printf("generic mx e2m1 %08x s %d\n", (unsigned)b, s); return 1; }
      if (lx <= s + 8 && generic_e4m3(x, s) != oracle_float_to_fp8_e4m3fn(x / scale)) { printf("generic mx e4m3 %08x s %d\n", (unsigned)b, s); return 1; }
      if (lx <= s + 15 && generic_e5m2(x, s) != oracle_float_to_fp8_e5m2(x / scale)) { printf("generic mx e5m2 %08x s %d\n", (unsigned)b, s); return 1; }
      ++checked;
    }
  }
  // ldexp_exact over every code and scale byte.
  for (int sb = 0; sb < 256; ++sb) {
    for (int c = 0; c < 256; ++c) {
      const float a = msq::ldexp_exact(msq::decode_e4m3(c), sb - 127);
      const float b = std::ldexp(oracle_fp8_e4m3fn_to_float(c), sb - 127);
      if (u(a) != u(b) && !(std::isnan(a) && std::isnan(b))) { printf("ldexp %d %d\n", sb, c); return 1; }
    }
  }
  if (argc > 3 && std::string(argv[3]) == "full") {
    // Library encoders against the generic formulation: every bit pattern,
    // then every MX scale exponent (the block-valid range |v| < 2^(s + emax
    // + 1) and everything beyond it) over a dense sample.
    const unsigned nt = std::max(1u, std::thread::hardware_concurrency());
    std::atomic<uint64_t> bad{0}, first{~0ull}, count{0};
    auto run = [&](unsigned t) {
      uint64_t local = 0;
      for (uint64_t b = t; b <= 0xffffffffull; b += nt) {
        const float x = f(static_cast<uint32_t>(b));
        if (msq::encode_e4m3(x) != generic_e4m3(x) || msq::encode_e5m2(x) != generic_e5m2(x) ||
            msq::encode_e2m1(x) != generic_e2m1(x)) {
          ++bad;
          uint64_t cur = first.load();
          while (b < cur && !first.compare_exchange_weak(cur, b)) {}
        }
        ++local;
      }
      for (int s = -127 + static_cast<int>(t); s <= 125; s += static_cast<int>(nt)) {
        for (uint64_t b = 0; b <= 0xffffffffull; b += 4099ull) {
          const float x = f(static_cast<uint32_t>(b));
          if (msq::encode_e4m3(x, s) != generic_e4m3(x, s) ||
              msq::encode_e5m2(x, s) != generic_e5m2(x, s) ||
              msq::encode_e2m1(x, s) != generic_e2m1(x, s)) {
            ++bad;
            first = (static_cast<uint64_t>(s + 1000) << 32) | b;
          }
          ++local;
        }
      }
      count += local;
    };
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; ++t) pool.emplace_back(run, t);
    for (auto& th : pool) th.join();
    if (bad) { printf("full: %llu mismatches, first %llx\n", (unsigned long long)bad.load(), (unsigned long long)first.load()); return 1; }
    checked += count;
  }
  printf("ok %llu\n", static_cast<unsigned long long>(checked));
  return 0;
}
