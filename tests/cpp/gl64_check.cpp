// Host check of the Goldilocks helpers the kernels use (cuyashe_amd/csrc/gl64.hpp):
// every result against 128-bit arithmetic, for every twiddle exponent of the
// shift-only DFTs and the edge values of the field. Prints "ok" and exits 0.
#include <stdint.h>
#include <stdio.h>

#include "gl64.hpp"

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
          gl::sub(a, b) != (uint64_t)(((unsigned __int128)a + PP - b) % PP)) {
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
  printf("ok %ld\n", checks);
  return 0;
}
