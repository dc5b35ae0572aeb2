// Host-side construction of one ring's constant tables.
//
// Reference counterparts:
//  * prime choice      src/polynomial.h:898-955 (gen_crt_primes: primes taken
//                      in descending order until M > 2*degree*q^2) over the
//                      bucket at src/polynomial.cpp:22;
//  * ICRT constants    src/polynomial.h:939-943 (M/p_i and (M/p_i)^-1 mod p_i),
//                      sent to the device by src/cuda_bn.cu:1047-1147;
//  * twiddles          src/cuda_functions.cu:878-918 (omega = 7^((P-1)/N)).
//
// Differences that keep every result identical while fixing the reference's
// exactness holes (its 29-bit primes overflow the Goldilocks convolution for
// n > 64):
//  * The prime size is the largest that keeps the Goldilocks convolution exact:
//    for a ring product (residues of both operands centred in (-p/2, p/2]),
//    n*((p-1)/2)^2 <= (P-1)/2; for the keyswitch, whose digits are transformed
//    once, centred but not reduced, and shared by every modulus,
//    l_w*n*2^(w_bits-1)*((p-1)/2) <= (P-1)/2.
//  * The prime count makes M >= 4*B, B the largest |coefficient| the operation
//    can produce, so the ICRT quotient k = round(sum y_i/p_i) can be found
//    with 64-bit fixed point and no correction step.
//  * Modulus 0 is the Goldilocks prime P itself: the transform computes the
//    convolution mod P exactly whatever the operand sizes, so the residue mod
//    P needs no size bound and carries 64 bits of M for the price of one
//    small prime (the reference's CRT uses small primes only; the
//    reconstructed integers are the same).
#pragma once
#include <stdint.h>

#include <stdexcept>
#include <vector>

#include "bigint.hpp"
#include "gl64.hpp"

namespace yh {

inline bool is_prime_u64(uint64_t x) {
  if (x < 2) return false;
  for (uint64_t p : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull}) {
    if (x % p == 0) return x == p;
  }
  uint64_t d = x - 1;
  int s = 0;
  while ((d & 1) == 0) { d >>= 1; ++s; }
  auto mulm = [x](uint64_t a, uint64_t b) { return (uint64_t)((unsigned __int128)a * b % x); };
  auto powm = [&](uint64_t b, uint64_t e) {
    uint64_t r = 1;
    b %= x;
    while (e) { if (e & 1) r = mulm(r, b); b = mulm(b, b); e >>= 1; }
    return r;
  };
  for (uint64_t a : {2ull, 3ull, 5ull, 7ull, 11ull, 13ull, 17ull, 19ull, 23ull, 29ull, 31ull, 37ull}) {
    uint64_t v = powm(a, d);
    if (v == 1 || v == x - 1) continue;
    bool comp = true;
    for (int r = 1; r < s; ++r) {
      v = mulm(v, v);
      if (v == x - 1) { comp = false; break; }
    }
    if (comp) return false;
  }
  return true;
}

inline uint64_t inv_mod_u64(uint64_t a, uint64_t m) {
  int64_t t = 0, nt = 1;
  int64_t r = (int64_t)m, nr = (int64_t)(a % m);
  while (nr) {
    int64_t q = r / nr;
    int64_t tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  if (r != 1) throw std::domain_error("inverse does not exist");
  if (t < 0) t += (int64_t)m;
  return (uint64_t)t;
}

// Largest prime size allowed by the exactness bound of the Goldilocks
// convolution: K * n * ((p-1)/2)^2 <= (P-1)/2.
// Primes are also kept below 2^28 so that the CRT word sums of the device
// (products of a 32-bit word and a residue, eight per partial sum) fit 64 bits.
inline uint64_t max_prime_for(int n, int K) {
  unsigned __int128 lim = (unsigned __int128)((gl::P - 1) / 2);
  uint64_t lo = 3, hi = (1ull << 28) - 1;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo + 1) / 2;
    unsigned __int128 h = (mid - 1) / 2;
    unsigned __int128 v = (unsigned __int128)K * (unsigned __int128)n * h * h;
    if (v <= lim) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Largest prime size for the keyswitch set. The keyswitch transforms every
// digit once, centred (d - 2^(w_bits-1), |.| <= 2^(w_bits-1)) but otherwise
// unreduced, and shares that transform between all moduli; only the key
// residue is reduced mod p and centred. The l_w-term multiply-accumulate
// stays exact in the Goldilocks field when
// l_w * n * 2^(w_bits-1) * ((p-1)/2) <= (P-1)/2; the shift of the digits is
// a fixed correction per key (added after the inverse transform).
inline uint64_t max_prime_ks(int n, int lw, int w_bits) {
  const unsigned __int128 lim = (unsigned __int128)((gl::P - 1) / 2);
  const unsigned __int128 per = (unsigned __int128)(lw > 1 ? lw : 1) * (unsigned __int128)n *
                                (unsigned __int128)(1ull << (w_bits - 1));
  const unsigned __int128 h = lim / per;  // largest (p-1)/2
  const unsigned __int128 cap = (1ull << 28) - 1;
  const unsigned __int128 pm = 2 * h + 1;
  if (pm < 3) throw std::invalid_argument("keyswitch digits too wide for exact transforms");
  return (uint64_t)(pm < cap ? pm : cap);
}

// Descending primes <= pmax until their product reaches `need`.
inline std::vector<uint64_t> take_primes(uint64_t pmax, const BigInt& need) {
  std::vector<uint64_t> out;
  BigInt M(1);
  uint64_t c = pmax | 1;
  if (c > pmax) c -= 2;
  while (M < need) {
    while (!is_prime_u64(c)) c -= 2;
    out.push_back(c);
    M *= BigInt(c);
    c -= 2;
    if (out.size() > 255) throw std::runtime_error("too many CRT primes");
  }
  return out;
}

// One reconstruction set: a prefix of the ring primes.
struct IcrtSet {
  int R = 0;
  BigInt M;
  std::vector<uint64_t> Mi_limbs;   // [R][LW]
  std::vector<uint64_t> invMi;      // [R]  (M/p_i)^-1 mod p_i
  std::vector<uint64_t> M_limbs;    // [LW]
  // Reconstruction mod q without the full integer: c = sum y_i M_i - k M, so
  // c mod q = (sum y_i (M_i mod q) + k (q - M mod q)) mod q.
  // Row R: 2^32 (M_0 mod q) mod q, the constant of the high 32 bits of the
  // 64-bit y of modulus 0 (the device splits that y into two halves).
  std::vector<uint64_t> Mq_limbs;   // [R + 1][LQ]: M_i mod q
  std::vector<uint64_t> negMq_limbs;  // [LQ]: q - (M mod q), in (0, q]
};

// Scale-and-round tables of the product set: t M_i = q alpha_i + beta_i with
// beta_i in [0, q), t M = q alpha_M + beta_M. For c = sum y_i M_i - k M,
//   t c / q = sum y_i alpha_i - k alpha_M + S / q,  S = sum y_i beta_i - k beta_M,
// so round(t c / q) mod q follows from two short accumulations.
struct ScaleTables {
  // Row R belongs to the high half y_0h of the 64-bit y of modulus 0:
  // y_0h 2^32 t M_0 = q y_0h (2^32 alpha_0 + f) + y_0h beta_0', with
  // 2^32 beta_0 = q f + beta_0', so row R of alpha is (2^32 alpha_0 + f) mod q
  // and row R of beta is beta_0'.
  std::vector<uint64_t> alpha_limbs;  // [R + 1][LQ]: alpha_i mod q
  std::vector<uint64_t> beta_limbs;   // [R + 1][LQ]: beta_i
  std::vector<uint64_t> negbetaM;     // [LQ]: q - beta_M, in (0, q]
  std::vector<uint64_t> negalpha1;    // [LQ]: q - ((alpha_M + 1) mod q), in (0, q]
};

struct RingParams {
  int n = 0, logn = 0;
  BigInt q;
  int LQ = 0;       // limbs of q
  int LW = 0;       // working limbs of the reconstruction (2*LQ + 1)
  uint64_t t = 0;   // plaintext modulus
  int w_bits = 0;   // keyswitch base 2^w_bits (0: no keyswitch)
  int lw = 0;       // number of keyswitch digits
  uint64_t pmax = 0;                // largest small prime of the product set
  uint64_t pmax_ks = 0;             // ... of the keyswitch set
  // Product set (index 0): R moduli, P then small primes descending, sized
  // for one product per convolution. Keyswitch set (index 1): P then smaller
  // primes sized for the l_w-term multiply-accumulate.
  std::vector<uint64_t> primes, primes_ks;
  std::vector<uint64_t> barrett, barrett_ks;  // floor(2^64 / p_i) (1 for P, unused)
  std::vector<uint64_t> pow32, pow32_ks;      // [R][2*LW]: 2^(32k) mod p_i
  IcrtSet full, ks;
  ScaleTables scale;
  std::vector<uint64_t> q_limbs;    // [LQ]
  std::vector<uint64_t> mu_limbs;   // [LW+2-LQ]: floor(2^(64*(LW+1)) / q)
  // One-word division of the reconstruction epilogues: for V < 2^(64 LQ + 35)
  // the quotient V / q is below 2^62 when bits(q) >= 64 LQ - 26; then
  // floor(floor(V / 2^s) qmu / 2^64), s = bits(q) - 1 (bit q_shift of limb
  // LQ - 1), qmu = floor(2^(s+64) / q), is the quotient or falls short by <= 2.
  bool fast_div = false;
  int q_shift = 0;
  uint64_t qmu = 0;

  static int ilog2(int v) { int l = 0; while ((1 << l) < v) ++l; return l; }

  // P followed by descending primes <= pm until the product reaches `need`.
  // The reconstruction epilogues sum y_i * (32-bit word) over the small
  // primes in one 64-bit column, exact while sum_i p_i < 2^32; a set that
  // would exceed it (very short rings with wide q) takes smaller primes.
  static std::vector<uint64_t> moduli_for(uint64_t& pm, const BigInt& need) {
    const BigInt Pb(gl::P);
    BigInt need_small = (need + Pb - BigInt(1)) / Pb;
    for (;;) {
      std::vector<uint64_t> out(1, gl::P);
      if (need_small > BigInt(1)) {
        std::vector<uint64_t> sp = take_primes(pm, need_small);
        out.insert(out.end(), sp.begin(), sp.end());
      }
      uint64_t psum = 0;
      for (size_t i = 1; i < out.size(); ++i) psum += out[i];
      if (psum < (1ull << 32)) return out;
      pm = pm / 8 * 7;
      if (pm < 1024) throw std::runtime_error("no CRT prime set fits the reconstruction bounds");
    }
  }

  void tables_for(const std::vector<uint64_t>& pr, std::vector<uint64_t>& bar, std::vector<uint64_t>& pw) {
    const int R = (int)pr.size();
    bar.resize(R);
    pw.assign((size_t)R * 2 * LW, 0);
    for (int i = 0; i < R; ++i) {
      uint64_t p = pr[i];
      bar[i] = p == gl::P ? 1 : (uint64_t)((((unsigned __int128)1) << 64) / p);
      uint64_t v = 1;
      for (int k = 0; k < 2 * LW; ++k) {
        pw[(size_t)i * 2 * LW + k] = v;
        v = (uint64_t)(((unsigned __int128)v << 32) % p);
      }
    }
  }

  void build_set(IcrtSet& s, const std::vector<uint64_t>& primes) {
    const int R = (int)primes.size();
    uint64_t psum = 0;  // bound of the folded epilogues (moduli_for)
    for (int i = 1; i < R; ++i) psum += primes[i];
    if (psum >= (1ull << 32)) throw std::logic_error("small CRT primes exceed the folded-epilogue bound");
    s.R = R;
    s.M = BigInt(1);
    for (int i = 0; i < R; ++i) s.M *= BigInt(primes[i]);
    if (s.M.bits() + 2 > 64 * LW) throw std::runtime_error("CRT modulus exceeds working width");
    s.Mi_limbs.assign((size_t)R * LW, 0);
    s.invMi.assign(R, 0);
    s.M_limbs.assign(LW, 0);
    s.M.to_limbs(s.M_limbs.data(), LW);
    s.Mq_limbs.assign((size_t)(R + 1) * LQ, 0);
    for (int i = 0; i < R; ++i) {
      BigInt Mi = s.M / BigInt(primes[i]);
      Mi.to_limbs(&s.Mi_limbs[(size_t)i * LW], LW);
      const uint64_t r = Mi.mod_u64(primes[i]);
      s.invMi[i] = primes[i] == gl::P ? gl::inv(r) : inv_mod_u64(r, primes[i]);
      (Mi % q).to_limbs(&s.Mq_limbs[(size_t)i * LQ], LQ);
    }
    s.negMq_limbs.assign(LQ, 0);
    (q - s.M % q).to_limbs(s.negMq_limbs.data(), LQ);
    ((BigInt::pow2(32) * ((s.M / BigInt(primes[0])) % q)) % q).to_limbs(&s.Mq_limbs[(size_t)R * LQ], LQ);
  }

  void build_scale(const IcrtSet& s) {
    const int R = s.R;
    scale.alpha_limbs.assign((size_t)(R + 1) * LQ, 0);
    scale.beta_limbs.assign((size_t)(R + 1) * LQ, 0);
    scale.negbetaM.assign(LQ, 0);
    scale.negalpha1.assign(LQ, 0);
    if (t == 0) return;
    const BigInt tb(t);
    for (int i = 0; i < R; ++i) {
      const BigInt tMi = tb * (s.M / BigInt(primes[i]));
      ((tMi / q) % q).to_limbs(&scale.alpha_limbs[(size_t)i * LQ], LQ);
      (tMi % q).to_limbs(&scale.beta_limbs[(size_t)i * LQ], LQ);
    }
    {
      const BigInt tM0 = tb * (s.M / BigInt(primes[0]));
      const BigInt a0 = tM0 / q, b0 = tM0 % q;
      const BigInt hb = BigInt::pow2(32) * b0;
      ((BigInt::pow2(32) * a0 + hb / q) % q).to_limbs(&scale.alpha_limbs[(size_t)R * LQ], LQ);
      (hb % q).to_limbs(&scale.beta_limbs[(size_t)R * LQ], LQ);
    }
    const BigInt tM = tb * s.M;
    (q - tM % q).to_limbs(scale.negbetaM.data(), LQ);
    (q - (tM / q + BigInt(1)) % q).to_limbs(scale.negalpha1.data(), LQ);
  }

  // n: ring degree (power of two, x^n + 1); q: modulus (odd, > 1);
  // t: plaintext modulus (0 allowed when unused); w_bits: keyswitch digit
  // width in bits (0: ring has no keyswitch).
  RingParams(int n_, const BigInt& q_, uint64_t t_, int w_bits_) {
    if (n_ < 2 || (n_ & (n_ - 1))) throw std::invalid_argument("n must be a power of two >= 2");
    if (n_ > 16384) throw std::invalid_argument("n > 16384 is not supported");
    if (q_ <= BigInt(1)) throw std::invalid_argument("q must be > 1");
    if (w_bits_ < 0 || w_bits_ > 32) throw std::invalid_argument("w_bits must be in [0, 32]");
    n = n_;
    logn = ilog2(n);
    q = q_;
    t = t_;
    w_bits = w_bits_;
    LQ = (q.bits() + 63) / 64;
    if (LQ > 8) throw std::invalid_argument("q wider than 512 bits is not supported");
    LW = 2 * LQ + 1;
    lw = w_bits ? (q.bits() + w_bits - 1) / w_bits : 0;
    BigInt qm1 = q - BigInt(1);
    // Ring products: |c_j| <= n (q-1)^2 for operands in [0, q).
    BigInt B_mul = BigInt(n) * qm1 * qm1;
    BigInt need = BigInt(4) * B_mul;
    if (need < BigInt(4)) need = BigInt(4);
    pmax = max_prime_for(n, 1);
    primes = moduli_for(pmax, need);
    tables_for(primes, barrett, pow32);
    build_set(full, primes);
    build_scale(full);
    if (w_bits) {
      // Keyswitch: |c_j| <= l_w n (w-1)(q-1), l_w products summed per convolution.
      BigInt B_ks = BigInt(lw) * BigInt(n) * (BigInt::pow2(w_bits) - BigInt(1)) * qm1;
      pmax_ks = max_prime_ks(n, lw, w_bits);
      primes_ks = moduli_for(pmax_ks, BigInt(4) * B_ks);
    } else {
      pmax_ks = pmax;
      primes_ks = primes;
    }
    tables_for(primes_ks, barrett_ks, pow32_ks);
    build_set(ks, primes_ks);
    q_limbs.assign(LQ, 0);
    q.to_limbs(q_limbs.data(), LQ);
    int LMU = LW + 2 - LQ;
    BigInt mu = BigInt::pow2(64 * (LW + 1)) / q;
    mu_limbs.assign(LMU, 0);
    mu.to_limbs(mu_limbs.data(), LMU);
    const int b = q.bits();
    q_shift = (b - 1) & 63;
    fast_div = b >= 64 * LQ - 26 && q > BigInt::pow2(b - 1);
    if (fast_div) {
      BigInt m = BigInt::pow2(b - 1 + 64) / q;  // < 2^64 because q > 2^(b-1)
      qmu = m.to_u64();
    }
  }
};

// Tables of the radix-16 transforms of length N (layout documented in
// ntt16.hpp): inter-pass twiddles w_L^(o*k) (k = 1..15, o < s, per pass with
// s = 2^b0 > 1), their inverses, the negacyclic twist psi^j and the untwist
// psi^-j * N^-1, with w_N = 7^((P-1)/N) and psi = 7^((P-1)/(2N)).
// fwd/inv in Montgomery form (applied with gl::mont_mul); untwist_pw and
// n_inv_pw carry 2^128 (Montgomery form and the 2^-64 of a mont_mul
// pointwise product), untwist_m 2^64.
// Negacyclic transforms with the twist folded into the first forward pass
// and the untwist into the last inverse pass (s = N/16, g < s, c < 16, the
// pass touches elements j = g + c s): psi^j = psi^g 2^(78 c) because
// psi^s = 7^((P-1)/32) = 2^78, so the per-thread factor psi^g joins the
// pass's twiddles and the per-slot factor is a shift:
//   fwd0[c s + g] = w_N^(brev4(c) g) psi^g                  (Montgomery form)
//   inv0_m[c s + g] = w_N^(-brev4(c) g) psi^-g N^-1         (Montgomery form)
//   inv0_pw = inv0_m carrying 2^128 (after a mont_mul pointwise product).
struct Ntt16Tables {
  std::vector<uint64_t> fwd, inv, twist, untwist_pw, untwist_m, fwd0, inv0_pw, inv0_m;
  uint64_t n_inv_pw = 0;
};

inline Ntt16Tables ntt16_tables(int N) {
  Ntt16Tables t;
  int logN = 0;
  while ((1 << logN) < N) ++logN;
  const uint64_t w = gl::pow(7, (gl::P - 1) / (uint64_t)N);
  for (int p = 0; p < logN / 4; ++p) {
    const int b0 = logN - 4 - 4 * p;
    if (b0 <= 0) continue;
    const int s = 1 << b0;
    const uint64_t wL = gl::pow(w, (uint64_t)(N / (16 * s)));
    const uint64_t wLi = gl::inv(wL);
    for (int k = 1; k < 16; ++k) {
      const uint64_t step = gl::pow(wL, (uint64_t)k), stepi = gl::pow(wLi, (uint64_t)k);
      uint64_t a = 1, ai = 1;
      for (int o = 0; o < s; ++o) {
        t.fwd.push_back(gl::to_mont(a));
        t.inv.push_back(gl::to_mont(ai));
        a = gl::mul(a, step);
        ai = gl::mul(ai, stepi);
      }
    }
  }
  if (t.fwd.empty()) { t.fwd.push_back(gl::to_mont(1)); t.inv.push_back(gl::to_mont(1)); }
  const uint64_t psi = gl::pow(7, (gl::P - 1) / (2 * (uint64_t)N));
  const uint64_t psii = gl::inv(psi);
  const uint64_t n_inv = gl::inv((uint64_t)N);
  t.n_inv_pw = gl::to_mont(gl::to_mont(n_inv));
  t.twist.resize(N);
  t.untwist_pw.resize(N);
  t.untwist_m.resize(N);
  uint64_t a = 1, ai = n_inv;
  for (int j = 0; j < N; ++j) {
    t.twist[j] = a;
    t.untwist_m[j] = gl::to_mont(ai);
    t.untwist_pw[j] = gl::to_mont(t.untwist_m[j]);
    a = gl::mul(a, psi);
    ai = gl::mul(ai, psii);
  }
  const int s = N / 16;
  const uint64_t wi = gl::inv(w);
  t.fwd0.resize(N);
  t.inv0_m.resize(N);
  t.inv0_pw.resize(N);
  for (int c = 0; c < 16; ++c) {
    const int br = ((c & 1) << 3) | ((c & 2) << 1) | ((c & 4) >> 1) | ((c & 8) >> 3);
    for (int g = 0; g < s; ++g) {
      const uint64_t f = gl::mul(gl::pow(w, (uint64_t)br * g), t.twist[g]);
      const uint64_t v = gl::mul(gl::mul(gl::pow(wi, (uint64_t)br * g), gl::pow(psii, (uint64_t)g)), n_inv);
      t.fwd0[(size_t)c * s + g] = gl::to_mont(f);
      t.inv0_m[(size_t)c * s + g] = gl::to_mont(v);
      t.inv0_pw[(size_t)c * s + g] = gl::to_mont(gl::to_mont(v));
    }
  }
  return t;
}

}  // namespace yh
