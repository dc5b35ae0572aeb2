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

// 32-bit add/subtract with carry/borrow in and out (v_add_co/v_sub_co chains
// on the device; plain arithmetic in a host build). The helpers below are
// written as such chains of 32-bit operations: on gfx950 they issue faster
// than the equivalent 64-bit adds and compares (a canonical add as a chain
// runs about 13 % faster in a register-resident microbenchmark).
GL_HD uint32_t addc32(uint32_t a, uint32_t b, uint32_t cin, uint32_t* cout) {
#if defined(__clang__)
  unsigned c;
  const uint32_t r = __builtin_addc(a, b, cin, &c);
  *cout = c;
  return r;
#else
  const uint64_t s = (uint64_t)a + b + cin;
  *cout = (uint32_t)(s >> 32);
  return (uint32_t)s;
#endif
}
GL_HD uint32_t subc32(uint32_t a, uint32_t b, uint32_t bin, uint32_t* bout) {
#if defined(__clang__)
  unsigned c;
  const uint32_t r = __builtin_subc(a, b, bin, &c);
  *bout = c;
  return r;
#else
  const uint64_t d = (uint64_t)a - b - bin;
  *bout = (uint32_t)(d >> 63);
  return (uint32_t)d;
#endif
}

// a + b mod P for canonical a, b. With s = a + b mod 2^64 and carry c, and
// s + EPS mod 2^64 with carry c2: on c the true sum is s + 2^64 = s + EPS
// (mod P), and s + EPS < P (a + b <= 2P - 2); without c, s >= P exactly when
// c2, and then s + EPS - 2^64 = s - P. So the result is s + EPS when c or c2.
GL_HD uint64_t add(uint64_t a, uint64_t b) {
  uint32_t k0, c, k1, c2;
  const uint32_t s0 = addc32((uint32_t)a, (uint32_t)b, 0u, &k0);
  const uint32_t s1 = addc32((uint32_t)(a >> 32), (uint32_t)(b >> 32), k0, &c);
  const uint32_t t0 = addc32(s0, 0xFFFFFFFFu, 0u, &k1);
  const uint32_t t1 = addc32(s1, 0u, k1, &c2);
  return (c | c2) ? (((uint64_t)t1 << 32) | t0) : (((uint64_t)s1 << 32) | s0);
}

// a - b mod P for canonical a, b, with the borrow taken from the
// subtraction's own carry chain.
GL_HD uint64_t sub(uint64_t a, uint64_t b) {
  uint32_t br, x;
  const uint32_t d0 = subc32((uint32_t)a, (uint32_t)b, 0u, &br);
  const uint32_t d1 = subc32((uint32_t)(a >> 32), (uint32_t)(b >> 32), br, &br);
  // On borrow d holds a - b + 2^64; a - b + P is that minus EPS.
  const uint32_t r0 = subc32(d0, br ? 0xFFFFFFFFu : 0u, 0u, &x);
  const uint32_t r1 = subc32(d1, 0u, x, &x);
  return ((uint64_t)r1 << 32) | r0;
}
GL_HD uint64_t sub_full(uint64_t a, uint64_t b) { return sub(a, b); }

GL_HD uint64_t neg(uint64_t a) { return a == 0 ? 0 : P - a; }

// Reduce top 2^128 + hi 2^64 + lo modulo P, canonical result, for top < 2^4
// (0 for a 128-bit value; up to 15 for a sum of up to 16 products).
// hi = hh' 2^32 + hl; 2^64 = EPS, 2^96 = -1 and 2^128 = -2^32 (mod P), so
// with hh = top 2^32 + hh' (< 2^36 < P) the value is lo - hh + hl EPS.
GL_HD uint64_t reduce_wide(uint32_t top, uint64_t hi, uint64_t lo) {
  const uint32_t hl = (uint32_t)hi;
  uint32_t b, b2, x, c, c2;
  // t0 = lo - hh (hh = top 2^32 + (hi >> 32) < 2^36), plus P on borrow
  // (subtract EPS; lo - hh + P >= 0, so that step never borrows)
  const uint32_t d0 = subc32((uint32_t)lo, (uint32_t)(hi >> 32), 0u, &b);
  const uint32_t d1 = subc32((uint32_t)(lo >> 32), top, b, &b);
  const uint32_t t00 = subc32(d0, b ? 0xFFFFFFFFu : 0u, 0u, &b2);
  const uint32_t t01 = subc32(d1, 0u, b2, &x);
  // t1 = hl EPS = hl 2^32 - hl <= 2^64 - 2^33 + 1
  const uint32_t u0 = subc32(0u, hl, 0u, &b);
  const uint32_t u1 = subc32(hl, 0u, b, &x);
  // t2 = t0 + t1. On carry the value is t2 + 2^64 = t2 + EPS (mod P), and
  // then t2 + EPS <= P - 2; without carry t2 >= P needs t2 - P = t2 + EPS
  // (mod 2^64), which is when t2 + EPS carries. One conditional addition
  // covers both.
  const uint32_t s0 = addc32(t00, u0, 0u, &c);
  const uint32_t s1 = addc32(t01, u1, c, &c);
  const uint32_t v0 = addc32(s0, 0xFFFFFFFFu, 0u, &c2);
  const uint32_t v1 = addc32(s1, 0u, c2, &c2);
  return (c | c2) ? (((uint64_t)v1 << 32) | v0) : (((uint64_t)s1 << 32) | s0);
}

// Reduce hi * 2^64 + lo modulo P, canonical result.
GL_HD uint64_t reduce128(uint64_t hi, uint64_t lo) { return reduce_wide(0, hi, lo); }

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

// Montgomery reduction with R = 2^64: hi 2^64 + lo (< P 2^64) times 2^-64
// mod P, canonical. P^-1 = 1 + 2^32 (mod 2^64), so the quotient digit is
// m = lo (1 + 2^32) mod 2^64 = (m1 : l0) with m1 = l1 + l0 mod 2^32 and no
// multiply, and floor(m P / 2^64) = m - m1 - e, e the carry of l1 + l0. The
// result hi - floor(m P / 2^64) = (hi + m1 + e) - m lies in (-P, P); the sum
// hi + m1 + e stays below 2^64 (hi < P and m1 + e < 2^32), so it is one
// 64-bit addition, one 64-bit subtraction and +P when negative.
GL_HD uint64_t mont_reduce(uint64_t hi, uint64_t lo) {
  const uint32_t l0 = (uint32_t)lo, l1 = (uint32_t)(lo >> 32);
  uint32_t e, c0, c1, br, br2, x;
  const uint32_t m1 = addc32(l1, l0, 0u, &e);  // m = (m1 : l0)
  const uint32_t a0 = addc32((uint32_t)hi, m1, e, &c0);
  const uint32_t a1 = addc32((uint32_t)(hi >> 32), 0u, c0, &c1);  // no carry out
  const uint32_t r0 = subc32(a0, l0, 0u, &br);
  const uint32_t r1 = subc32(a1, m1, br, &br2);
  // negative: add P = 2^64 - EPS, i.e. subtract EPS modulo 2^64
  const uint32_t s0 = subc32(r0, br2 ? 0xFFFFFFFFu : 0u, 0u, &x);
  const uint32_t s1 = subc32(r1, 0u, x, &x);
  return ((uint64_t)s1 << 32) | s0;
}

// a b 2^-64 mod P: with b = w 2^64 mod P (a constant stored in Montgomery
// form) this is a w mod P, six instructions cheaper than mul().
GL_HD uint64_t mont_mul(uint64_t a, uint64_t b) {
  uint64_t hi, lo;
  mul128(a, b, hi, lo);
  return mont_reduce(hi, lo);
}

// w 2^64 mod P (Montgomery form of a canonical w).
GL_HD uint64_t to_mont(uint64_t w) { return mul(w, EPS); }

// x * 2^e mod P for canonical x, canonical result. Meant for exponents
// known at compile time (the twiddles inside the 16-, 8-, 4- and 2-point
// DFTs): after inlining only one branch remains. 2^96 = -1 and 2^192 = 1, so
// e reduces to (0, 96) with a sign; a negated product negates x first
// (P - x; x = 0 gives P, which the reductions below map to 0).
//  * 0 < e <= 32:  x 2^e = lo + hi 2^64 with hi < 2^32, and 2^64 = EPS.
//  * 64 <= e < 96: with k = 96 - e <= 32, x = xh 2^k + xl,
//      x 2^e = -x 2^-k = -(xh - xl 2^(96-k)) = zh EPS - xh,
//    zh = xl 2^(32-k) < 2^32 (xl 2^(96-k) = zh 2^64): one subtraction of
//    canonical values instead of a 128-bit reduction.
//  * 32 < e < 64:  with s = 64 - e and y = y1 2^32 + y0, y0 2^e = h 2^32 +
//    (y0 >> s) 2^64 (h = y0 2^(32-s) mod 2^32) and, by the case above,
//    y1 2^(e+32) = (y1 2^(32-s) mod 2^32) EPS - (y1 >> s); the two EPS
//    factors are disjoint bit fields of m = bits [s, s+32) of y, so
//    y 2^e = h 2^32 + m EPS - (y1 >> s), all three terms canonical.
GL_HD uint64_t mul_pow2(uint64_t x, int e) {
  e %= 192;
  if (e < 0) e += 192;
  if (e == 0) return x;
  if (e == 96) return neg(x);
  const bool ng = e > 96;
  if (ng) e -= 96;
  if (e >= 64) {
    const int k = 96 - e;
    const uint64_t zh = (uint64_t)(uint32_t)((uint32_t)x << (32 - k));
    const uint64_t a = (zh << 32) - zh;  // zh * EPS < P
    const uint64_t xh = x >> k;           // < 2^(64-k) < P
    return ng ? sub(xh, a) : sub(a, xh);
  }
  const uint64_t y = ng ? P - x : x;
  if (e > 32) {
    // h 2^32 + m EPS - r = (h + m) 2^32 - (m + r); with h + m = cA 2^32 + a,
    // (h + m) 2^32 = (a + cA) 2^32 - cA (2^64 = EPS), so the value is
    // a1 2^32 - B, a1 = a + cA < 2^32 (h has its low 32 - s bits clear),
    // B = m + r + cA < 2^33: one 64-bit subtraction, plus P when negative,
    // and the result is canonical either way.
    const int s = 64 - e;
    const uint32_t h = (uint32_t)y << (32 - s);
    const uint32_t m = (uint32_t)(y >> s);
    const uint32_t r = (uint32_t)(y >> 32) >> s;
    uint32_t cA, cB, c2, br, br2, z;
    const uint32_t a = addc32(h, m, 0u, &cA);
    const uint32_t b0 = addc32(m, r, 0u, &cB);
    const uint32_t a1 = a + cA;
    const uint32_t bl = addc32(b0, cA, 0u, &c2);
    const uint32_t bh = cB + c2;
    const uint32_t v0 = subc32(0u, bl, 0u, &br);
    const uint32_t v1 = subc32(a1, bh, br, &br);
    const uint32_t w0 = subc32(v0, br ? 0xFFFFFFFFu : 0u, 0u, &br2);  // + P = - EPS (mod 2^64)
    const uint32_t w1 = subc32(v1, 0u, br2, &z);
    return ((uint64_t)w1 << 32) | w0;
  }
  return reduce128(y >> (64 - e), y << e);
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
