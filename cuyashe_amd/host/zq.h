// Fixed-width arithmetic in Z_q (q up to 512 bits) for host-side key
// generation: the polynomial inverse f^-1 mod (x^n + 1, q) that
// Yashe::generate_keys needs (src/yashe.cpp:37-55, Polynomial::InvMod at
// src/polynomial.h:1338-1367) is O(n^2) Z_q operations, too many for a
// general-purpose bignum. Values are kept as L little-endian 64-bit limbs;
// products are reduced with a Barrett reciprocal mu = floor(2^(128 L) / q).
#pragma once
#include <stdint.h>

#include <array>
#include <stdexcept>
#include <vector>

#include "zz.h"

namespace yh {

class Zq {
 public:
  static constexpr int MAXL = 8;
  using Elt = std::array<uint64_t, MAXL>;

  explicit Zq(const ZZ& q) : q_(q) {
    L_ = (q.bits() + 63) / 64;
    if (L_ > MAXL) throw std::invalid_argument("q too wide");
    qv_.fill(0);
    q.to_limbs(qv_.data(), L_);
    ZZ mu = ZZ::pow2(128 * L_) / q;
    mu_.fill(0);
    mu.to_limbs(mu_.data(), L_ + 2);
  }
  int limbs() const { return L_; }
  const ZZ& modulus() const { return q_; }

  Elt from(const ZZ& a) const {
    ZZ r = a % q_;
    Elt e{};
    r.to_limbs(e.data(), L_);
    return e;
  }
  ZZ to(const Elt& e) const { return ZZ::from_limbs(e.data(), L_); }
  Elt zero() const { return Elt{}; }
  Elt one() const { Elt e{}; e[0] = 1; return e; }
  bool is_zero(const Elt& a) const {
    for (int l = 0; l < L_; ++l) if (a[l]) return false;
    return true;
  }

  Elt sub(const Elt& a, const Elt& b) const {
    Elt r{};
    uint64_t br = 0;
    for (int l = 0; l < L_; ++l) {
      unsigned __int128 d = (unsigned __int128)a[l] - b[l] - br;
      r[l] = (uint64_t)d;
      br = (uint64_t)(d >> 64) ? 1 : 0;
    }
    if (br) {
      uint64_t c = 0;
      for (int l = 0; l < L_; ++l) {
        unsigned __int128 s = (unsigned __int128)r[l] + qv_[l] + c;
        r[l] = (uint64_t)s;
        c = (uint64_t)(s >> 64);
      }
    }
    return r;
  }

  Elt mul(const Elt& a, const Elt& b) const {
    uint64_t x[2 * MAXL] = {0};
    for (int i = 0; i < L_; ++i) {
      unsigned __int128 c = 0;
      for (int k = 0; k < L_; ++k) {
        unsigned __int128 t = (unsigned __int128)a[i] * b[k] + x[i + k] + c;
        x[i + k] = (uint64_t)t;
        c = t >> 64;
      }
      x[i + L_] = (uint64_t)c;
    }
    // qhat = floor(x * mu / 2^(128 L)); only the top limbs matter
    uint64_t xm[3 * MAXL + 4] = {0};
    for (int i = 0; i < 2 * L_; ++i) {
      unsigned __int128 c = 0;
      for (int k = 0; k < L_ + 2; ++k) {
        unsigned __int128 t = (unsigned __int128)x[i] * mu_[k] + xm[i + k] + c;
        xm[i + k] = (uint64_t)t;
        c = t >> 64;
      }
      xm[i + L_ + 2] = (uint64_t)c;
    }
    const uint64_t* qh = xm + 2 * L_;
    uint64_t qq[MAXL + 1] = {0};
    for (int i = 0; i <= L_; ++i) {
      unsigned __int128 c = 0;
      for (int k = 0; k < L_ && i + k <= L_; ++k) {
        unsigned __int128 t = (unsigned __int128)qh[i] * qv_[k] + qq[i + k] + c;
        qq[i + k] = (uint64_t)t;
        c = t >> 64;
      }
      if (i + L_ <= L_) qq[i + L_] += (uint64_t)c;
    }
    uint64_t r[MAXL + 1];
    uint64_t br = 0;
    for (int l = 0; l <= L_; ++l) {
      unsigned __int128 d = (unsigned __int128)x[l] - qq[l] - br;
      r[l] = (uint64_t)d;
      br = (uint64_t)(d >> 64) ? 1 : 0;
    }
    while (geq_q(r)) {
      br = 0;
      for (int l = 0; l <= L_; ++l) {
        uint64_t ql = l < L_ ? qv_[l] : 0;
        unsigned __int128 d = (unsigned __int128)r[l] - ql - br;
        r[l] = (uint64_t)d;
        br = (uint64_t)(d >> 64) ? 1 : 0;
      }
    }
    Elt e{};
    for (int l = 0; l < L_; ++l) e[l] = r[l];
    return e;
  }

  // a^-1 for prime q (Fermat); throws on zero.
  Elt inv(const Elt& a) const {
    if (is_zero(a)) throw std::domain_error("zero has no inverse");
    ZZ e = q_ - ZZ(2);
    Elt r = one(), b = a;
    for (int i = 0; i < e.bits(); ++i) {
      if (e.bit(i)) r = mul(r, b);
      b = mul(b, b);
    }
    return r;
  }

 private:
  bool geq_q(const uint64_t* r) const {
    if (r[L_]) return true;
    for (int l = L_ - 1; l >= 0; --l)
      if (r[l] != qv_[l]) return r[l] > qv_[l];
    return true;
  }
  ZZ q_;
  int L_;
  Elt qv_{}, mu_{};
};

// f^-1 in Z_q[x]/(x^n + 1) (q prime) by the extended Euclidean algorithm on
// (x^n + 1, f). Returns false if f is not invertible.
inline bool poly_inverse_negacyclic(const Zq& Z, const std::vector<ZZ>& f, int n, std::vector<ZZ>& out) {
  using E = Zq::Elt;
  std::vector<E> r0(n + 1, Z.zero()), r1(n + 1, Z.zero()), s0(n + 1, Z.zero()), s1(n + 1, Z.zero());
  r0[0] = Z.one();
  r0[n] = Z.one();
  int d0 = n, d1 = -1;
  for (int j = 0; j < n && j < (int)f.size(); ++j) {
    r1[j] = Z.from(f[j]);
    if (!Z.is_zero(r1[j])) d1 = j;
  }
  s1[0] = Z.one();
  int ds0 = -1, ds1 = 0;
  while (d1 >= 0) {
    if (d1 == 0) {
      E inv = Z.inv(r1[0]);
      out.assign(n, ZZ(0));
      for (int j = 0; j <= ds1 && j < n; ++j) out[j] = Z.to(Z.mul(s1[j], inv));
      return true;
    }
    E inv = Z.inv(r1[d1]);
    while (d0 >= d1) {
      int k = d0 - d1;
      E c = Z.mul(r0[d0], inv);
      for (int j = 0; j <= d1; ++j) r0[j + k] = Z.sub(r0[j + k], Z.mul(c, r1[j]));
      for (int j = 0; j <= ds1; ++j) s0[j + k] = Z.sub(s0[j + k], Z.mul(c, s1[j]));
      if (ds1 + k > ds0) ds0 = ds1 + k;
      while (d0 >= 0 && Z.is_zero(r0[d0])) --d0;
      while (ds0 >= 0 && Z.is_zero(s0[ds0])) --ds0;
    }
    std::swap(r0, r1);
    std::swap(s0, s1);
    std::swap(d0, d1);
    std::swap(ds0, ds1);
  }
  return false;
}

}  // namespace yh
