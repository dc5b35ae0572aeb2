// Host-side code of the library exercised under AddressSanitizer and
// UndefinedBehaviorSanitizer (tests/test_sanitizers.py builds it with g++
// -fsanitize=address,undefined): the wire-format decoder on corrupted and
// random input (it parses untrusted blobs), the bignum (bigint.hpp) against
// its own identities, Z_q arithmetic and the polynomial inverse (host/zq.h),
// and the ring planner (ring_params.hpp) on every supported shape. Prints OK
// and exits 0 when every check holds; a sanitizer report aborts the run.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../cuyashe_amd/csrc/ring_params.hpp"
#include "../../cuyashe_amd/csrc/wire.hpp"
#include "../../cuyashe_amd/host/zq.h"

static int failures = 0;
#define CHECK(c)                                                        \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c);          \
      ++failures;                                                       \
    }                                                                   \
  } while (0)

static std::mt19937_64 rng(20241204);

static ZZ rand_zz(int bits, bool sign) {
  std::vector<uint64_t> w((bits + 63) / 64);
  for (auto& x : w) x = rng();
  if (bits % 64) w.back() &= (1ull << (bits % 64)) - 1;
  ZZ r = ZZ::from_limbs(w.data(), (int)w.size());
  return sign && (rng() & 1) ? -r : r;
}

static void wire_checks() {
  for (int L : {1, 2, 4, 8}) {
    const int n = 8, count = 2;
    yhw::Params p;
    p.n = n;
    p.L = L;
    for (int l = 0; l < L; ++l) p.q[l] = rng() | 1;
    p.q[L - 1] |= 1ull << 63;
    p.t = 1024;
    p.w_bits = 32;
    std::vector<uint64_t> src((size_t)count * L * n);
    for (auto& x : src) x = rng();
    for (int i = 0; i < count; ++i)
      for (int j = 0; j < n; ++j) src[((size_t)i * L + (L - 1)) * n + j] = p.q[L - 1] - 1;  // below q
    size_t len = 0;
    CHECK(yhw::encode(p, 1, false, src.data(), count, nullptr, 0, &len).empty());
    std::vector<uint8_t> blob(len);
    CHECK(yhw::encode(p, 1, false, src.data(), count, blob.data(), blob.size(), &len).empty());
    std::vector<uint64_t> dst(src.size());
    int got = 0;
    CHECK(yhw::decode(p, 1, false, blob.data(), blob.size(), dst.data(), count, &got).empty() && got == count &&
          dst == src);
    // every single-bit flip is refused (checksum, header or range)
    for (size_t bit = 0; bit < 8 * blob.size(); ++bit) {
      std::vector<uint8_t> b = blob;
      b[bit / 8] ^= (uint8_t)(1u << (bit % 8));
      CHECK(!yhw::decode(p, 1, false, b.data(), b.size(), dst.data(), count, &got).empty());
    }
    // every truncation, and copies of the input at every shorter length
    for (size_t k = 0; k < blob.size(); ++k) {
      std::vector<uint8_t> b(blob.begin(), blob.begin() + k);
      yhw::Header h;
      (void)yhw::peek(b.empty() ? nullptr : b.data(), b.size(), &h);
      CHECK(!yhw::decode(p, 1, false, b.empty() ? nullptr : b.data(), b.size(), dst.data(), count, &got).empty());
    }
  }
  // random buffers, some with a valid magic and version
  yhw::Params any;
  any.n = 8;
  any.L = 2;
  any.q[0] = any.q[1] = ~0ull;
  for (int it = 0; it < 3000; ++it) {
    std::vector<uint8_t> b(rng() % 400);
    for (auto& x : b) x = (uint8_t)rng();
    if (b.size() >= 8 && (it & 1)) {
      yhw::put_u32(b.data(), yhw::MAGIC);
      yhw::put_u16(b.data() + 4, yhw::VERSION);
    }
    std::vector<uint64_t> dst(64);
    int got = 0;
    yhw::Header h;
    (void)yhw::peek(b.empty() ? nullptr : b.data(), b.size(), &h);
    (void)yhw::decode(any, 0, false, b.empty() ? nullptr : b.data(), b.size(), dst.data(), 2, &got);
  }
}

static void bigint_checks() {
  for (int it = 0; it < 4000; ++it) {
    const ZZ a = rand_zz(1 + (int)(rng() % 700), true);
    ZZ b = rand_zz(1 + (int)(rng() % 400), true);
    if (b.is_zero()) b = ZZ(7);
    ZZ q, r;
    ZZ::divmod_floor(a, b, q, r);
    CHECK(q * b + r == a);
    CHECK(b.is_negative() ? (r <= ZZ(0) && r > b) : (r >= ZZ(0) && r < b));  // sign of the divisor
    CHECK((a * b) / b == a);
    CHECK((a + b) - b == a);
    CHECK(ZZ::from_string(a.to_string()) == a);
    const int s = (int)(rng() % 200);
    CHECK(((a.abs() << s) >> s) == a.abs());
    if (!a.is_negative()) {
      std::vector<uint64_t> w(a.num_limbs() + 1);
      a.to_limbs(w.data(), (int)w.size());
      CHECK(ZZ::from_limbs(w.data(), (int)w.size()) == a);
    }
  }
  const ZZ p = ZZ::pow2(127) - ZZ(1);  // prime
  for (int it = 0; it < 200; ++it) {
    ZZ x = rand_zz(126, false) + ZZ(1);
    CHECK((ZZ::inv_mod(x, p) * x) % p == ZZ(1));
    CHECK(ZZ::pow_mod(x, p - ZZ(1), p) == ZZ(1));
  }
}

static void zq_checks() {
  for (const char* qs : {"170141183460469231731687303715884105727",
                         "57896044618658097711785492504343953926634992332820282019728792003956564819949",
                         "77287149995192912462927307869", "2305843009213693951"}) {
    const ZZ q = ZZ::from_string(qs);
    yh::Zq Z(q);
    for (int it = 0; it < 500; ++it) {
      const ZZ a = rand_zz(q.bits() + 8, false) % q, b = rand_zz(q.bits(), false) % q;
      CHECK(Z.to(Z.mul(Z.from(a), Z.from(b))) == (a * b) % q);
      CHECK(Z.to(Z.sub(Z.from(a), Z.from(b))) == (a - b) % q);
    }
    const int n = 16;
    std::vector<ZZ> f(n), inv;
    for (auto& v : f) v = rand_zz(q.bits(), false) % q;
    if (yh::poly_inverse_negacyclic(Z, f, n, inv)) {
      // f * f^-1 = 1 mod (x^n + 1, q)
      std::vector<ZZ> prod(n, ZZ(0));
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          const ZZ t = f[i] * inv[j];
          if (i + j < n) prod[i + j] += t;
          else prod[i + j - n] -= t;
        }
      for (int k = 0; k < n; ++k) CHECK(prod[k] % q == ZZ(k == 0 ? 1 : 0));
    }
  }
}

static void planner_checks() {
  const char* qs[] = {"170141183460469231731687303715884105727",
                      "57896044618658097711785492504343953926634992332820282019728792003956564819949",
                      "6703903964971298549787012499102923063739682910296196688861780721860882015036773488400937149083451"
                      "713845015929093243025426876941405973284973216824503041861",
                      "2305843009213693951", "655615111"};
  for (int logn = 1; logn <= 14; ++logn)
    for (const char* q : qs)
      for (int wb : {0, 16, 32}) {
        yh::RingParams P(1 << logn, yh::BigInt::from_string(q), 1024, wb);
        CHECK(!P.primes.empty() && P.primes[0] == 0xFFFFFFFF00000001ull);
      }
}

int main() {
  wire_checks();
  bigint_checks();
  zq_checks();
  planner_checks();
  std::printf("%s (%d failures)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
