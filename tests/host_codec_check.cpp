// Host build of the device codecs (mixed-mlx_amd/csrc/device_codecs.hpp)
// checked against the C oracle (oracle/codecs_ref.c) over float bit patterns.
// Built and run by tests/test_codecs_host.py; prints "ok <count>" or the first
// mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

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
  printf("ok %llu\n", static_cast<unsigned long long>(checked));
  return 0;
}
