// Residue arithmetic modulo a small CRT prime p (p < 2^28) on 32-bit lanes.
//
// After an inverse transform, a small-prime residue of the product is a
// canonical Goldilocks element x standing for the signed integer
// c = x - P [x > P/2] (|c| <= P/2 by the prime-size rule); the reconstruction
// wants y = (c mod p) (M/p)^-1 mod p. The reference reduces the coefficient
// with a 32-bit-half long division per prime (bn_mod1_low,
// src/cuda_bn.cu:290-317) and multiplies by (M/p)^-1 in cuPreICRT
// (src/cuda_bn.cu:854-903). Here both steps use precomputed constants per
// prime and 32x32-bit products only:
//   x = xh 2^32 + xl,  c = xh (2^32 mod p) + xl + [x > P/2] (p - P mod p) (mod p)
// with xh (2^32 mod p) by Shoup's method (constant multiplier w, w' =
// floor(w 2^32 / p): a w - floor(a w' / 2^32) p lies in [0, 2p) for any
// a < 2^32), xl by Barrett with floor(2^32 / p) (result in [0, 2p)), and the
// sum u < 5p < 2^31 multiplied by (M/p)^-1 with Shoup again.
#pragma once
#include <stdint.h>

#include "gl64.hpp"

namespace yhk {

// Constants of one small prime (filled on the host, read as one 32-byte row).
struct SmallPrime {
  uint32_t p;
  uint32_t c32;    // 2^32 mod p
  uint32_t c32s;   // floor(c32 2^32 / p)
  uint32_t bm;     // floor(2^32 / p)
  uint32_t kc;     // (p - P mod p) mod p: adds -P (mod p) for x > P/2
  uint32_t inv;    // (M/p)^-1 mod p
  uint32_t invs;   // floor(inv 2^32 / p)
  uint32_t bl;     // floor(2^64 / p) mod 2^32 (floor(2^64 / p) = bm 2^32 + bl)
};
static_assert(sizeof(SmallPrime) == 32, "one 32-byte row per prime");

GL_HD uint32_t mulhi32(uint32_t a, uint32_t b) { return (uint32_t)(((uint64_t)a * b) >> 32); }

// a w mod p in [0, 2p) for a < 2^32, w < p, ws = floor(w 2^32 / p).
GL_HD uint32_t shoup(uint32_t a, uint32_t w, uint32_t ws, uint32_t p) {
  return a * w - mulhi32(a, ws) * p;
}

// y = ((lift(x) + extra) mod p) (M/p)^-1 mod p in [0, p), x a canonical
// Goldilocks element, extra < p.
GL_HD uint32_t finish_small(uint64_t x, const SmallPrime& k, uint32_t extra = 0) {
  const uint32_t xh = (uint32_t)(x >> 32), xl = (uint32_t)x;
  const uint32_t t = shoup(xh, k.c32, k.c32s, k.p);       // [0, 2p)
  const uint32_t r = xl - mulhi32(xl, k.bm) * k.p;        // [0, 2p)
  const uint32_t s = x > (gl::P >> 1) ? k.kc : 0u;        // [0, p)
  const uint32_t y = shoup(t + r + s + extra, k.inv, k.invs, k.p);  // sum < 6p < 2^31; y in [0, 2p)
  const uint32_t y2 = y - k.p;
  return y2 < y ? y2 : y;  // y >= p: y - p; else y - p wraps above y
}

// x mod p for x < 2^63 (a sum of 32-bit words times residues mod p), p < 2^28.
// Barrett with floor(2^64/p) = bm 2^32 + bl, dropping the x0 bl term: the
// quotient estimate x1 bm + floor((x1 bl + x0 bm) / 2^32) is at most two
// below floor(x/p), so r = x - q p lies in [0, 3p) and only the low 32-bit
// words of x and q p are needed; two conditional subtractions finish.
GL_HD uint32_t mod_small(uint64_t x, const SmallPrime& k) {
  const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
  const uint64_t m = (uint64_t)x1 * k.bl + (uint64_t)x0 * k.bm;  // < 2^63 + 2^64/p
  const uint32_t q = x1 * k.bm + (uint32_t)(m >> 32);
  uint32_t r = x0 - q * k.p;
  uint32_t r2 = r - k.p;
  r = r2 < r ? r2 : r;
  r2 = r - k.p;
  return r2 < r ? r2 : r;
}

// Host: the constants of prime p with (M/p)^-1 mod p = inv.
inline SmallPrime make_small_prime(uint32_t p, uint32_t inv) {
  SmallPrime k{};
  k.p = p;
  k.c32 = (uint32_t)((1ull << 32) % p);
  k.c32s = (uint32_t)(((uint64_t)k.c32 << 32) / p);
  k.bm = (uint32_t)((1ull << 32) / p);
  k.kc = (uint32_t)((p - gl::P % p) % p);
  k.inv = inv;
  k.invs = (uint32_t)(((uint64_t)inv << 32) / p);
  k.bl = (uint32_t)(~0ull / p);  // floor((2^64 - 1) / p) = floor(2^64 / p) for p not a power of two
  return k;
}

}  // namespace yhk
