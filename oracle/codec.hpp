// Storage codecs for the distribution buffers of the oracle (test infrastructure only).
//
// The oracle is the CPU restatement the HIP kernels are checked against; nothing in the
// product links it. Codecs follow src/shaders/common3d.ts:92-103 (FP16 scaled by 2^15 with a
// +-1.99 clamp); BF16 has no counterpart in the reference and is defined here as round to
// nearest even of the float32 bit pattern with no scaling or clamp.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>

namespace orc {

inline uint32_t f32_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return u;
}
inline float bits_f32(uint32_t u) {
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// IEEE 754 binary32 -> binary16, round to nearest, ties to even, with subnormals, inf and NaN.
// Written from the format definitions: the 10-bit mantissa keeps bits 22..13 of the float
// mantissa; the dropped bits decide the rounding.
inline uint16_t f32_to_f16_rne(float f) {
  const uint32_t x = f32_bits(f);
  const uint16_t sign = uint16_t((x >> 16) & 0x8000u);
  const uint32_t ax = x & 0x7fffffffu;
  if (ax >= 0x7f800000u) {  // inf or NaN
    return uint16_t(sign | 0x7c00u | (ax > 0x7f800000u ? 0x0200u : 0u));
  }
  const int32_t exp = int32_t(ax >> 23) - 127;  // unbiased exponent
  if (exp > 15) return uint16_t(sign | 0x7c00u);  // overflow (before rounding)
  uint32_t mant = (ax & 0x7fffffu) | 0x800000u;   // 24-bit significand with hidden bit
  if (ax < 0x00800000u) mant = ax & 0x7fffffu;    // float subnormal: no hidden bit
  // Half value = mant * 2^(exp - 23); half subnormal unit is 2^-24, normal unit 2^(e-10).
  int32_t shift;  // bits of mant dropped
  uint32_t hexp;  // biased half exponent field for normals
  if (exp < -14) {
    // Result is a half subnormal (or zero): units of 2^-24, value/2^-24 = mant * 2^(exp - 23 + 24).
    shift = 23 - 24 - exp;  // = -1 - exp; >= 14
    hexp = 0;
    if (ax < 0x00800000u) shift = 126;  // float subnormals are far below half's range
  } else {
    shift = 13;
    hexp = uint32_t(exp + 15);
  }
  uint32_t q, rem, half;
  if (shift >= 32) {
    q = 0;
    rem = mant;
    half = 0xffffffffu;  // remainder is always below half a unit
    if (shift == 32) half = 0x80000000u;
  } else {
    q = mant >> shift;
    rem = mant & ((1u << shift) - 1u);
    half = 1u << (shift - 1);
  }
  if (rem > half || (rem == half && (q & 1u))) q += 1;
  if (hexp == 0) {
    // q counts 2^-24 units; q == 1024 rolls over into the smallest normal, which is encoded the same way.
    return uint16_t(sign | q);
  }
  // q holds the 11-bit significand with hidden bit; a carry to 2048 bumps the exponent.
  uint32_t h = (hexp << 10) + (q - 0x400u);
  if (h >= 0x7c00u) return uint16_t(sign | 0x7c00u);
  return uint16_t(sign | h);
}

inline float f16_to_f32(uint16_t h) {
  const uint32_t sign = uint32_t(h & 0x8000u) << 16;
  const uint32_t e = (h >> 10) & 0x1fu;
  const uint32_t m = h & 0x3ffu;
  if (e == 0) {
    // zero or subnormal: m * 2^-24, exact in float
    float v = float(m) * 5.9604644775390625e-08f;
    return sign ? -v : v;
  }
  if (e == 31) return bits_f32(sign | 0x7f800000u | (m << 13));
  return bits_f32(sign | ((e - 15 + 127) << 23) | (m << 13));
}

// binary32 -> bfloat16 (top 16 bits), round to nearest even; NaN stays a quiet NaN.
inline uint16_t f32_to_bf16_rne(float f) {
  uint32_t x = f32_bits(f);
  if ((x & 0x7fffffffu) > 0x7f800000u) return uint16_t((x >> 16) | 0x0040u);
  x += 0x7fffu + ((x >> 16) & 1u);
  return uint16_t(x >> 16);
}
inline float bf16_to_f32(uint16_t b) { return bits_f32(uint32_t(b) << 16); }

// Storage policies: load() gives the compute value, store() encodes it.
struct StoreF64 {
  using elem = double;
  static constexpr int bytes = 8;
  static constexpr bool stochastic = false;
  template <class T> static T load(elem e) { return T(e); }
  template <class T> static elem store(T v) { return double(v); }
};
struct StoreF32 {
  using elem = float;
  static constexpr int bytes = 4;
  static constexpr bool stochastic = false;
  template <class T> static T load(elem e) { return T(e); }
  template <class T> static elem store(T v) { return float(v); }
};
// src/shaders/common3d.ts:95,100: f32(h) * (1/32768) and f16(clamp(v, -1.99, 1.99) * 32768).
struct StoreF16 {
  using elem = uint16_t;
  static constexpr int bytes = 2;
  static constexpr bool stochastic = false;
  template <class T> static T load(elem e) { return T(f16_to_f32(e) * (1.0f / 32768.0f)); }
  template <class T> static elem store(T v) {
    float c = float(v);
    c = std::fmin(std::fmax(c, -1.99f), 1.99f);
    return f32_to_f16_rne(c * 32768.0f);
  }
};
struct StoreBF16 {
  using elem = uint16_t;
  static constexpr int bytes = 2;
  static constexpr bool stochastic = false;
  template <class T> static T load(elem e) { return T(bf16_to_f32(e)); }
  template <class T> static elem store(T v) { return f32_to_bf16_rne(float(v)); }
};

// lowbias32, a 32-bit integer hash (xor-shift 16, multiply 0x7feb352d, xor-shift 15, multiply
// 0x846ca68b, xor-shift 16). The stochastic codec's random word for population i of a node in the step
// with index n is lowbias32(node ^ lowbias32(n)) + i * 0x9E3779B9 (mod 2^32).
inline uint32_t lowbias32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}
inline uint32_t step_random(uint32_t node, uint32_t step, int i) {
  return lowbias32(node ^ lowbias32(step)) + uint32_t(i) * 0x9E3779B9u;
}

// binary32 -> binary16 with stochastic rounding by the random word r, for magnitudes in the binary16
// normal range [2^-14, 65504]: the top 13 bits of r are added to the 13 mantissa bits binary16 drops
// and the sum is truncated (towards zero), so the magnitude rounds up with probability equal to the
// dropped fraction; smaller magnitudes round to nearest even. (The gfx950 conversion the kernels use,
// v_cvt_sr_f16_f32, behaves this way: checked bit for bit on 65,536 random values and words.)
inline uint16_t f32_to_f16_sr(float f, uint32_t r) {
  const uint32_t x = f32_bits(f);
  const uint32_t ax = x & 0x7fffffffu;
  if (ax < 0x38800000u) return f32_to_f16_rne(f);  // below 2^-14
  const uint32_t sum = ax + (r >> 19);
  const uint32_t e = (sum >> 23) - 112u;  // binary16 exponent field (no overflow below 65504 + 2^-13 ulp)
  return uint16_t(((x >> 16) & 0x8000u) | (e << 10) | ((sum >> 13) & 0x3ffu));
}

// LBM_FP16 with stochastic rounding in the time step (store_r); store() rounds to nearest (initial and
// refilled values).
struct StoreF16SR {
  using elem = uint16_t;
  static constexpr int bytes = 2;
  static constexpr bool stochastic = true;
  template <class T> static T load(elem e) { return StoreF16::load<T>(e); }
  template <class T> static elem store(T v) { return StoreF16::store<T>(v); }
  template <class T> static elem store_r(T v, uint32_t r) {
    float c = float(v);
    c = std::fmin(std::fmax(c, -1.99f), 1.99f);
    return f32_to_f16_sr(c * 32768.0f, r);
  }
};

}  // namespace orc
