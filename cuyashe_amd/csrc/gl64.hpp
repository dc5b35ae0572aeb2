// Arithmetic modulo the Goldilocks prime P = 2^64 - 2^32 + 1.
//
// The reference computes its NTT over this prime (src/cuda_functions.cu:18,
// PRIMEP) with s_add/s_sub/s_mul at src/cuda_functions.cu:282-383. The
// pinned definition of its product is the host branch uint128 % P
// (src/cuda_functions.cu:351); the helpers here compute the same canonical
// values in [0, P) with the 2^64 = 2^32 - 1 (mod P) fold instead of a
// 128-bit division, so the device code needs no division at all.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define GL_HD __host__ __device__ __forceinline__
#else
#define GL_HD inline
#endif

namespace gl {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod P

GL_HD uint64_t umulhi(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// a + b mod P for canonical a, b.
GL_HD uint64_t add(uint64_t a, uint64_t b) {
  uint64_t s = a + b;
  // On carry the true sum is s + 2^64; subtracting P leaves s + EPS.
  // Without carry but s >= P, s - P == s + EPS (mod 2^64) as well.
  bool fix = (s < a) | (s >= P);
  return fix ? s + EPS : s;
}

// a - b mod P for canonical a, b.
GL_HD uint64_t sub(uint64_t a, uint64_t b) {
  uint64_t d = a - b;
  // On borrow d holds a - b + 2^64; a - b + P is that minus EPS.
  return (a < b) ? d - EPS : d;
}

GL_HD uint64_t neg(uint64_t a) { return a == 0 ? 0 : P - a; }

// Reduce hi * 2^64 + lo modulo P, canonical result.
// hi = hh * 2^32 + hl; 2^64 = EPS and 2^96 = -1 (mod P).
GL_HD uint64_t reduce128(uint64_t hi, uint64_t lo) {
  uint64_t hh = hi >> 32;
  uint64_t hl = hi & EPS;
  uint64_t t0 = lo - hh;
  if (lo < hh) t0 -= EPS;              // borrow: add P = 2^64 - EPS
  uint64_t t1 = (hl << 32) - hl;       // hl * EPS < 2^64
  uint64_t t2 = t0 + t1;
  if (t2 < t0) t2 += EPS;              // carry: 2^64 = EPS
  if (t2 >= P) t2 -= P;
  return t2;
}

// Full 64x64 -> 128 product from four 32x32 -> 64 partial products
// (v_mad_u64_u32 on CDNA); no partial sum overflows 64 bits.
GL_HD void mul128(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
  uint64_t a0 = a & EPS, a1 = a >> 32, b0 = b & EPS, b1 = b >> 32;
  uint64_t p00 = a0 * b0;
  uint64_t mid = a0 * b1 + (p00 >> 32);         // < 2^64
  uint64_t mid2 = a1 * b0 + (mid & EPS);        // < 2^64
  hi = a1 * b1 + (mid >> 32) + (mid2 >> 32);
  lo = (mid2 << 32) | (p00 & EPS);
}

GL_HD uint64_t mul(uint64_t a, uint64_t b) {
  uint64_t hi, lo;
  mul128(a, b, hi, lo);
  return reduce128(hi, lo);
}

GL_HD uint64_t pow(uint64_t b, uint64_t e) {
  uint64_t r = 1;
  while (e) {
    if (e & 1) r = mul(r, b);
    b = mul(b, b);
    e >>= 1;
  }
  return r;
}

GL_HD uint64_t inv(uint64_t a) { return pow(a, P - 2); }

// Lift a canonical residue to the signed integer in (-P/2, P/2] it
// represents; the Goldilocks convolutions in this library are sized so the
// true (signed) value always lies in that range.
GL_HD int64_t lift(uint64_t v) {
  return v > (P >> 1) ? -(int64_t)(P - v) : (int64_t)v;
}

// Encode a small signed integer |x| < P/2 as a canonical residue.
GL_HD uint64_t from_signed(int64_t x) {
  return x < 0 ? P - (uint64_t)(-x) : (uint64_t)x;
}

}  // namespace gl
