// Host check of the Goldilocks and small-prime helpers the kernels use
// (cuyashe_amd/csrc/gl64.hpp, modp.hpp): every result against 128-bit
// arithmetic, for every twiddle exponent of the shift-only DFTs, the edge
// values of the field and a spread of CRT primes. Prints "ok" and exits 0.
#include <stdint.h>
#include <stdio.h>

#include "gl64.hpp"
#include "modp.hpp"

static const unsigned __int128 PP = gl::P;

static uint64_t ref_mul(uint64_t a, uint64_t b) { return (uint64_t)((unsigned __int128)a * b % PP); }

static uint64_t ref_pow2(uint64_t x, int e) {
  e %= 192;
  if (e < 0) e += 192;
  unsigned __int128 r = x % PP;
  for (int i = 0; i < e; ++i) r = (r * 2) % PP;
  return (uint64_t)r;
}

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static uint64_t next() {  // xorshift64*
  rng_state ^= rng_state >> 12;
  rng_state ^= rng_state << 25;
  rng_state ^= rng_state >> 27;
  return rng_state * 0x2545F4914F6CDD1Dull;
}

int main() {
  const uint64_t P = gl::P;
  uint64_t edges[] = {0, 1, 2, P - 1, P - 2, P >> 1, (P >> 1) + 1, 0xFFFFFFFFull, 0x100000000ull,
                      0xFFFFFFFF00000000ull, 0xFFFFFFFEFFFFFFFFull, 0x8000000000000000ull, 0x7FFFFFFFFFFFFFFFull,
                      0x00000000FFFFFFFEull, 0x0000000100000001ull, 0xFFFFFFFF00000000ull - 1};
  const int NE = sizeof(edges) / sizeof(edges[0]);
  long checks = 0;
  for (int e = -192; e < 400; ++e) {
    for (int i = 0; i < NE + 2000; ++i) {
      uint64_t x = i < NE ? edges[i] : next() % P;
      uint64_t got = gl::mul_pow2(x, e), want = ref_pow2(x, e);
      if (got != want) {
        printf("mul_pow2(%llu, %d) = %llu, want %llu\n", (unsigned long long)x, e, (unsigned long long)got,
               (unsigned long long)want);
        return 1;
      }
      ++checks;
    }
  }
  for (int i = 0; i < NE + 200000; ++i) {
    uint64_t a = i < NE ? edges[i] : next() % P;
    for (int j = 0; j < (i < NE ? NE : 1); ++j) {
      uint64_t b = i < NE ? edges[j] : next() % P;
      if (gl::mul(a, b) != ref_mul(a, b) || gl::add(a, b) != (uint64_t)(((unsigned __int128)a + b) % PP) ||
          gl::sub(a, b) != (uint64_t)(((unsigned __int128)a + PP - b) % PP) ||
          gl::sub_full(a, b) != (uint64_t)(((unsigned __int128)a + PP - b) % PP)) {
        printf("mul/add/sub mismatch at %llu %llu\n", (unsigned long long)a, (unsigned long long)b);
        return 1;
      }
      ++checks;
    }
  }
  // reduce128 on arbitrary 128-bit inputs
  for (int i = 0; i < 200000; ++i) {
    uint64_t hi = i < NE ? edges[i] : next(), lo = next();
    unsigned __int128 v = ((unsigned __int128)hi << 64) | lo;
    if (gl::reduce128(hi, lo) != (uint64_t)(v % PP)) {
      printf("reduce128 mismatch\n");
      return 1;
    }
    ++checks;
  }
  // reduce_wide: top 2^128 + hi 2^64 + lo for top < 16
  {
    const unsigned __int128 e = ((unsigned __int128)1 << 64) % PP;  // 2^64 mod P
    const uint64_t two128 = (uint64_t)(e * e % PP);
    for (int i = 0; i < 300000; ++i) {
      const uint32_t top = (uint32_t)(next() % 16);
      const uint64_t hi = i < NE ? edges[i] : next(), lo = i < NE ? ~edges[i] : next();
      const unsigned __int128 v = ((unsigned __int128)hi << 64) | lo;
      const uint64_t want = (uint64_t)(((unsigned __int128)top * two128 + v % PP) % PP);
      if (gl::reduce_wide(top, hi, lo) != want) {
        printf("reduce_wide mismatch\n");
        return 1;
      }
      ++checks;
    }
  }
  // mont_mul: a * to_mont(w) = a w mod P; mont_reduce(hi, lo) = (hi 2^64 + lo) 2^-64 mod P
  {
    const unsigned __int128 e = ((unsigned __int128)1 << 64) % PP;
    // 2^-64 mod P by exponentiation: (2^64)^(P-2)
    unsigned __int128 inv64 = 1, base = e;
    for (uint64_t k = P - 2; k; k >>= 1) {
      if (k & 1) inv64 = inv64 * base % PP;
      base = base * base % PP;
    }
    for (int i = 0; i < 400000; ++i) {
      const uint64_t a = i < NE * NE ? edges[i % NE] : next() % P;
      const uint64_t w = i < NE * NE ? edges[i / NE] : next() % P;
      if (gl::mont_mul(a, gl::to_mont(w)) != ref_mul(a, w)) {
        printf("mont_mul mismatch %llu %llu\n", (unsigned long long)a, (unsigned long long)w);
        return 1;
      }
      // any 128-bit value below P 2^64
      const uint64_t hi = (i < NE ? edges[i] : next()) % P, lo = next();
      const unsigned __int128 v = ((unsigned __int128)hi << 64) | lo;
      if (gl::mont_reduce(hi, lo) != (uint64_t)(v % PP * inv64 % PP)) {
        printf("mont_reduce mismatch\n");
        return 1;
      }
      checks += 2;
    }
  }
  // finish_small: (lift(x) mod p) * inv mod p, lift(x) = x - P [x > P/2]
  const uint32_t primes[] = {3, 5, 7, 65537, 262139, 524287, 94906249, 94906247, 134217689, 268435399};
  for (uint32_t p : primes) {
    for (int rep = 0; rep < 4; ++rep) {
      const uint32_t inv = rep == 0 ? 1 : rep == 1 ? p - 1 : (uint32_t)(next() % (p - 1)) + 1;
      const yhk::SmallPrime k = yhk::make_small_prime(p, inv);
      for (int i = 0; i < NE + 50000; ++i) {
        uint64_t x = i < NE ? edges[i] : next() % P;
        if (i == NE) x = P >> 1;
        if (i == NE + 1) x = (P >> 1) + 1;
        __int128 c = x > (P >> 1) ? (__int128)x - (__int128)P : (__int128)x;
        __int128 cm = c % p;
        if (cm < 0) cm += p;
        const uint32_t want = (uint32_t)((cm * inv) % p);
        const uint32_t got = yhk::finish_small(x, k);
        if (got != want) {
          printf("finish_small(%llu) mod %u = %u, want %u\n", (unsigned long long)x, p, got, want);
          return 1;
        }
        ++checks;
      }
    }
  }
  // mod_small: x mod p for x < 2^63 (the CRT sums), every prime size the
  // modulus sets use and both ends of the input range
  // the last five primes have a large low word of floor(2^64/p); with them
  // the quotient estimate can fall two short, and the inputs listed below
  // (found by search) need both conditional subtractions
  const uint32_t mprimes[] = {3,         5,         7,         257,       65537,     262139,   524287, 16777213,
                              94906249,  134217689, 268435399, 91393193,  204562447, 195244051, 19347019, 252698819};
  const uint64_t two_short[] = {9222041850514571444ull, 9222877272036056722ull, 9222919397028618387ull,
                                9223056466532287148ull, 9223270872530781452ull, 9222649857621698113ull};
  const uint64_t top = (1ull << 63) - 1;
  for (uint64_t x : two_short)
    for (uint32_t p : mprimes)
      if (yhk::mod_small(x, yhk::make_small_prime(p, 1)) != x % p) {
        printf("mod_small(%llu) mod %u wrong\n", (unsigned long long)x, p);
        return 1;
      }
  for (uint32_t p : mprimes) {
    const yhk::SmallPrime k = yhk::make_small_prime(p, 1);
    if ((((uint64_t)k.bm << 32) | k.bl) != (uint64_t)(((unsigned __int128)1 << 64) / p)) {
      printf("make_small_prime(%u): floor(2^64/p) words wrong\n", p);
      return 1;
    }
    const uint64_t qtop = top / p;
    const uint64_t special[] = {0, 1, p - 1, p, p + 1, 2ull * p - 1, top, top - 1, qtop * p, qtop * p - 1,
                                (qtop - 1) * p + (p - 1), 0xFFFFFFFFull, 0x100000000ull, (1ull << 62), top - p};
    const int NS = sizeof(special) / sizeof(special[0]);
    for (int i = 0; i < NS + 400000; ++i) {
      const uint64_t x = i < NS ? special[i] : (i & 1 ? next() & top : next() >> (next() % 63 + 1));
      const uint32_t got = yhk::mod_small(x, k), want = (uint32_t)(x % p);
      if (got != want) {
        printf("mod_small(%llu) mod %u = %u, want %u\n", (unsigned long long)x, p, got, want);
        return 1;
      }
      ++checks;
    }
  }
  printf("ok %ld\n", checks);
  return 0;
}
