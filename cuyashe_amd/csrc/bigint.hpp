// Host-side arbitrary-precision signed integer.
//
// The reference keeps every host coefficient as an NTL ZZ (src/polynomial.h:7,
// std::vector<ZZ> coefs at :1431) and derives its CRT tables with NTL
// (src/polynomial.h:898-955). NTL is not part of this build, so this header
// supplies the subset of ZZ arithmetic the ring set-up and the host API need:
// sign-magnitude storage in 64-bit limbs, floor division with a remainder that
// takes the divisor's sign (NTL's convention for / and %), shifts, bit masks
// and modular exponentiation.
#pragma once
#include <stdint.h>
#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

namespace yh {

class BigInt {
 public:
  BigInt() {}
  BigInt(int64_t v) { set_i64(v); }           // NOLINT: implicit like ZZ
  BigInt(int v) { set_i64(v); }               // NOLINT
  BigInt(uint64_t v) { if (v) mag_.push_back(v); }  // NOLINT
  BigInt(unsigned v) { if (v) mag_.push_back(v); }  // NOLINT
  BigInt(long long v) { set_i64((int64_t)v); }  // NOLINT
  BigInt(unsigned long long v) { if (v) mag_.push_back((uint64_t)v); }  // NOLINT

  static BigInt from_limbs(const uint64_t* limbs, int n, bool negative = false) {
    BigInt r;
    r.mag_.assign(limbs, limbs + n);
    r.neg_ = negative;
    r.trim();
    return r;
  }
  static BigInt from_string(const std::string& s) {
    BigInt r;
    size_t i = 0;
    bool negative = false;
    while (i < s.size() && (s[i] == ' ' || s[i] == '+' || s[i] == '-')) {
      if (s[i] == '-') negative = !negative;
      ++i;
    }
    if (i + 1 < s.size() && s[i] == '0' && (s[i + 1] == 'x' || s[i + 1] == 'X')) {
      for (i += 2; i < s.size(); ++i) {
        char c = s[i];
        int d;
        if (c >= '0' && c <= '9') d = c - '0';
        else if (c >= 'a' && c <= 'f') d = c - 'a' + 10;
        else if (c >= 'A' && c <= 'F') d = c - 'A' + 10;
        else break;
        r.mul_small_add(16, (uint64_t)d);
      }
    } else {
      for (; i < s.size(); ++i) {
        char c = s[i];
        if (c < '0' || c > '9') break;  // trailing 'L' etc. as NTL accepts
        r.mul_small_add(10, (uint64_t)(c - '0'));
      }
    }
    r.neg_ = negative;
    r.trim();
    return r;
  }

  std::string to_string() const {
    if (is_zero()) return "0";
    BigInt t = abs();
    std::string out;
    while (!t.is_zero()) {
      uint64_t rem = t.div_small_inplace(10000000000000000000ull);
      std::string chunk = std::to_string(rem);
      if (!t.is_zero()) chunk = std::string(19 - chunk.size(), '0') + chunk;
      out = chunk + out;
    }
    return neg_ ? "-" + out : out;
  }

  bool is_zero() const { return mag_.empty(); }
  bool is_negative() const { return neg_ && !mag_.empty(); }
  int sign() const { return is_zero() ? 0 : (neg_ ? -1 : 1); }
  int num_limbs() const { return (int)mag_.size(); }
  uint64_t limb(int i) const { return i < (int)mag_.size() ? mag_[i] : 0; }
  const std::vector<uint64_t>& limbs() const { return mag_; }
  // Writes |x| into n limbs; throws if it does not fit.
  void to_limbs(uint64_t* out, int n) const {
    if ((int)mag_.size() > n) throw std::overflow_error("BigInt does not fit in limbs");
    for (int i = 0; i < n; ++i) out[i] = limb(i);
  }
  uint64_t to_u64() const { return limb(0); }
  int64_t to_i64() const { return neg_ ? -(int64_t)limb(0) : (int64_t)limb(0); }

  int bits() const {
    if (mag_.empty()) return 0;
    return 64 * ((int)mag_.size() - 1) + (64 - __builtin_clzll(mag_.back()));
  }
  bool bit(int i) const {
    int w = i / 64;
    if (w >= (int)mag_.size()) return false;
    return (mag_[w] >> (i % 64)) & 1;
  }

  BigInt abs() const { BigInt r = *this; r.neg_ = false; return r; }
  BigInt operator-() const { BigInt r = *this; if (!r.is_zero()) r.neg_ = !r.neg_; return r; }

  static int cmp_abs(const BigInt& a, const BigInt& b) {
    if (a.mag_.size() != b.mag_.size()) return a.mag_.size() < b.mag_.size() ? -1 : 1;
    for (int i = (int)a.mag_.size() - 1; i >= 0; --i)
      if (a.mag_[i] != b.mag_[i]) return a.mag_[i] < b.mag_[i] ? -1 : 1;
    return 0;
  }
  static int cmp(const BigInt& a, const BigInt& b) {
    int sa = a.sign(), sb = b.sign();
    if (sa != sb) return sa < sb ? -1 : 1;
    int c = cmp_abs(a, b);
    return sa >= 0 ? c : -c;
  }
  friend bool operator==(const BigInt& a, const BigInt& b) { return cmp(a, b) == 0; }
  friend bool operator!=(const BigInt& a, const BigInt& b) { return cmp(a, b) != 0; }
  friend bool operator<(const BigInt& a, const BigInt& b) { return cmp(a, b) < 0; }
  friend bool operator<=(const BigInt& a, const BigInt& b) { return cmp(a, b) <= 0; }
  friend bool operator>(const BigInt& a, const BigInt& b) { return cmp(a, b) > 0; }
  friend bool operator>=(const BigInt& a, const BigInt& b) { return cmp(a, b) >= 0; }

  friend BigInt operator+(const BigInt& a, const BigInt& b) {
    if (a.neg_ == b.neg_) { BigInt r = add_abs(a, b); r.neg_ = a.neg_; r.trim(); return r; }
    int c = cmp_abs(a, b);
    if (c == 0) return BigInt();
    BigInt r = c > 0 ? sub_abs(a, b) : sub_abs(b, a);
    r.neg_ = c > 0 ? a.neg_ : b.neg_;
    r.trim();
    return r;
  }
  friend BigInt operator-(const BigInt& a, const BigInt& b) { return a + (-b); }
  friend BigInt operator*(const BigInt& a, const BigInt& b) {
    BigInt r;
    if (a.is_zero() || b.is_zero()) return r;
    r.mag_.assign(a.mag_.size() + b.mag_.size(), 0);
    for (size_t i = 0; i < a.mag_.size(); ++i) {
      unsigned __int128 carry = 0;
      for (size_t j = 0; j < b.mag_.size(); ++j) {
        unsigned __int128 cur = (unsigned __int128)a.mag_[i] * b.mag_[j] + r.mag_[i + j] + carry;
        r.mag_[i + j] = (uint64_t)cur;
        carry = cur >> 64;
      }
      size_t k = i + b.mag_.size();
      while (carry) {
        unsigned __int128 cur = (unsigned __int128)r.mag_[k] + carry;
        r.mag_[k] = (uint64_t)cur;
        carry = cur >> 64;
        ++k;
      }
    }
    r.neg_ = a.neg_ != b.neg_;
    r.trim();
    return r;
  }
  // Floor division (quotient rounded toward -inf), like NTL's operator/.
  friend BigInt operator/(const BigInt& a, const BigInt& b) { BigInt q, r; divmod_floor(a, b, q, r); return q; }
  // Remainder with the sign of the divisor, like NTL's operator%.
  friend BigInt operator%(const BigInt& a, const BigInt& b) { BigInt q, r; divmod_floor(a, b, q, r); return r; }
  BigInt& operator+=(const BigInt& b) { *this = *this + b; return *this; }
  BigInt& operator-=(const BigInt& b) { *this = *this - b; return *this; }
  BigInt& operator*=(const BigInt& b) { *this = *this * b; return *this; }
  BigInt& operator/=(const BigInt& b) { *this = *this / b; return *this; }
  BigInt& operator%=(const BigInt& b) { *this = *this % b; return *this; }

  // Shifts act on the magnitude; the sign is kept (NTL LeftShift/RightShift
  // on the non-negative values the reference shifts).
  friend BigInt operator<<(const BigInt& a, int s) {
    if (a.is_zero() || s == 0) return a;
    if (s < 0) return a >> (-s);
    BigInt r;
    int w = s / 64, b = s % 64;
    r.mag_.assign(a.mag_.size() + w + 1, 0);
    for (size_t i = 0; i < a.mag_.size(); ++i) {
      r.mag_[i + w] |= a.mag_[i] << b;
      if (b) r.mag_[i + w + 1] |= a.mag_[i] >> (64 - b);
    }
    r.neg_ = a.neg_;
    r.trim();
    return r;
  }
  friend BigInt operator>>(const BigInt& a, int s) {
    if (a.is_zero() || s == 0) return a;
    if (s < 0) return a << (-s);
    int w = s / 64, b = s % 64;
    if (w >= (int)a.mag_.size()) return BigInt();
    BigInt r;
    r.mag_.assign(a.mag_.size() - w, 0);
    for (size_t i = w; i < a.mag_.size(); ++i) {
      r.mag_[i - w] |= a.mag_[i] >> b;
      if (b && i + 1 < a.mag_.size()) r.mag_[i - w] |= a.mag_[i + 1] << (64 - b);
    }
    r.neg_ = a.neg_;
    r.trim();
    return r;
  }
  // Bitwise AND of magnitudes (operands are non-negative where used).
  friend BigInt operator&(const BigInt& a, const BigInt& b) {
    BigInt r;
    size_t n = std::min(a.mag_.size(), b.mag_.size());
    r.mag_.assign(n, 0);
    for (size_t i = 0; i < n; ++i) r.mag_[i] = a.mag_[i] & b.mag_[i];
    r.trim();
    return r;
  }

  static BigInt pow2(int e) { return BigInt(1) << e; }
  static BigInt pow(BigInt b, uint64_t e) {
    BigInt r(1);
    while (e) { if (e & 1) r *= b; b *= b; e >>= 1; }
    return r;
  }
  static BigInt pow_mod(BigInt b, BigInt e, const BigInt& m) {
    BigInt r(1);
    b = b % m;
    int nb = e.bits();
    for (int i = nb - 1; i >= 0; --i) {
      r = (r * r) % m;
      if (e.bit(i)) r = (r * b) % m;
    }
    return r % m;
  }
  // Inverse of a modulo m (m > 0); throws if gcd(a, m) != 1.
  static BigInt inv_mod(const BigInt& a, const BigInt& m) {
    BigInt r0 = m, r1 = a % m, s0 = 0, s1 = 1;
    while (!r1.is_zero()) {
      BigInt q, r;
      divmod_floor(r0, r1, q, r);
      r0 = r1; r1 = r;
      BigInt s = s0 - q * s1;
      s0 = s1; s1 = s;
    }
    if (r0 != BigInt(1)) throw std::domain_error("inv_mod: not invertible");
    return s0 % m;
  }

  // Floor division: a = q*b + r with 0 <= r < |b| when b > 0 and
  // b < r <= 0 when b < 0.
  static void divmod_floor(const BigInt& a, const BigInt& b, BigInt& q, BigInt& r) {
    if (b.is_zero()) throw std::domain_error("division by zero");
    BigInt qa, ra;
    divmod_abs(a, b, qa, ra);  // |a| = qa*|b| + ra
    bool neg_q = a.is_negative() != b.is_negative();
    if (!neg_q) {
      q = qa;
      r = ra;
      r.neg_ = b.is_negative();
      r.trim();
    } else if (ra.is_zero()) {
      q = -qa;
      r = BigInt();
    } else {
      q = -(qa + BigInt(1));
      r = b.abs() - ra;
      r.neg_ = b.is_negative();
      r.trim();
    }
  }

  // |a| = q*|b| + r, 0 <= r < |b|, all non-negative (Knuth D, 32-bit digits).
  static void divmod_abs(const BigInt& a, const BigInt& b, BigInt& q, BigInt& r) {
    if (cmp_abs(a, b) < 0) { q = BigInt(); r = a.abs(); return; }
    std::vector<uint32_t> u = to_u32(a.mag_), v = to_u32(b.mag_);
    while (!v.empty() && v.back() == 0) v.pop_back();
    while (!u.empty() && u.back() == 0) u.pop_back();
    size_t n = v.size(), m = u.size() - n;
    std::vector<uint32_t> qd(m + 1, 0);
    if (n == 1) {
      uint64_t rem = 0;
      for (int i = (int)u.size() - 1; i >= 0; --i) {
        uint64_t cur = (rem << 32) | u[i];
        qd[i] = (uint32_t)(cur / v[0]);
        rem = cur % v[0];
      }
      q = from_u32(qd);
      r = BigInt((uint64_t)rem);
      return;
    }
    int s = __builtin_clz(v[n - 1]);
    std::vector<uint32_t> vn(n), un(u.size() + 1);
    for (size_t i = n - 1; i > 0; --i) vn[i] = (v[i] << s) | (s ? (uint32_t)((uint64_t)v[i - 1] >> (32 - s)) : 0);
    vn[0] = v[0] << s;
    un[u.size()] = s ? (uint32_t)((uint64_t)u[u.size() - 1] >> (32 - s)) : 0;
    for (size_t i = u.size() - 1; i > 0; --i) un[i] = (u[i] << s) | (s ? (uint32_t)((uint64_t)u[i - 1] >> (32 - s)) : 0);
    un[0] = u[0] << s;
    const uint64_t B = 1ull << 32;
    for (int j = (int)m; j >= 0; --j) {
      uint64_t num = ((uint64_t)un[j + n] << 32) | un[j + n - 1];
      uint64_t qhat = num / vn[n - 1];
      uint64_t rhat = num % vn[n - 1];
      while (qhat >= B || qhat * vn[n - 2] > ((rhat << 32) | un[j + n - 2])) {
        --qhat;
        rhat += vn[n - 1];
        if (rhat >= B) break;
      }
      int64_t borrow = 0;
      uint64_t carry = 0;
      for (size_t i = 0; i < n; ++i) {
        uint64_t p = qhat * vn[i] + carry;
        carry = p >> 32;
        int64_t t = (int64_t)un[i + j] - (int64_t)(p & 0xFFFFFFFFull) - borrow;
        un[i + j] = (uint32_t)t;
        borrow = t < 0 ? 1 : 0;
      }
      int64_t t = (int64_t)un[j + n] - (int64_t)carry - borrow;
      un[j + n] = (uint32_t)t;
      if (t < 0) {
        --qhat;
        uint64_t c = 0;
        for (size_t i = 0; i < n; ++i) {
          uint64_t sum = (uint64_t)un[i + j] + vn[i] + c;
          un[i + j] = (uint32_t)sum;
          c = sum >> 32;
        }
        un[j + n] = (uint32_t)((uint64_t)un[j + n] + c);
      }
      qd[j] = (uint32_t)qhat;
    }
    std::vector<uint32_t> rd(n);
    for (size_t i = 0; i < n; ++i)
      rd[i] = (un[i] >> s) | (s ? (uint32_t)((uint64_t)un[i + 1] << (32 - s)) : 0);
    q = from_u32(qd);
    r = from_u32(rd);
  }

  uint64_t mod_u64(uint64_t m) const {  // |a| mod m
    unsigned __int128 rem = 0;
    for (int i = (int)mag_.size() - 1; i >= 0; --i) rem = ((rem << 64) | mag_[i]) % m;
    return (uint64_t)rem;
  }
  // a mod m in [0, m) for a signed value.
  uint64_t mod_u64_nonneg(uint64_t m) const {
    uint64_t r = mod_u64(m);
    return (is_negative() && r) ? m - r : r;
  }

 private:
  bool neg_ = false;
  std::vector<uint64_t> mag_;

  void set_i64(int64_t v) {
    mag_.clear();
    neg_ = v < 0;
    uint64_t m = v < 0 ? (uint64_t)(-(v + 1)) + 1 : (uint64_t)v;
    if (m) mag_.push_back(m);
  }
  void trim() {
    while (!mag_.empty() && mag_.back() == 0) mag_.pop_back();
    if (mag_.empty()) neg_ = false;
  }
  void mul_small_add(uint64_t m, uint64_t a) {
    unsigned __int128 carry = a;
    for (auto& w : mag_) {
      unsigned __int128 cur = (unsigned __int128)w * m + carry;
      w = (uint64_t)cur;
      carry = cur >> 64;
    }
    if (carry) mag_.push_back((uint64_t)carry);
  }
  uint64_t div_small_inplace(uint64_t d) {
    unsigned __int128 rem = 0;
    for (int i = (int)mag_.size() - 1; i >= 0; --i) {
      unsigned __int128 cur = (rem << 64) | mag_[i];
      mag_[i] = (uint64_t)(cur / d);
      rem = cur % d;
    }
    trim();
    return (uint64_t)rem;
  }
  static BigInt add_abs(const BigInt& a, const BigInt& b) {
    BigInt r;
    size_t n = std::max(a.mag_.size(), b.mag_.size());
    r.mag_.assign(n + 1, 0);
    uint64_t carry = 0;
    for (size_t i = 0; i < n; ++i) {
      unsigned __int128 s = (unsigned __int128)a.limb((int)i) + b.limb((int)i) + carry;
      r.mag_[i] = (uint64_t)s;
      carry = (uint64_t)(s >> 64);
    }
    r.mag_[n] = carry;
    r.trim();
    return r;
  }
  static BigInt sub_abs(const BigInt& a, const BigInt& b) {  // |a| >= |b|
    BigInt r;
    r.mag_.assign(a.mag_.size(), 0);
    uint64_t borrow = 0;
    for (size_t i = 0; i < a.mag_.size(); ++i) {
      uint64_t bi = b.limb((int)i);
      uint64_t d = a.mag_[i] - bi - borrow;
      borrow = (a.mag_[i] < bi) || (a.mag_[i] - bi < borrow) ? 1 : 0;
      r.mag_[i] = d;
    }
    r.trim();
    return r;
  }
  static std::vector<uint32_t> to_u32(const std::vector<uint64_t>& m) {
    std::vector<uint32_t> r(2 * m.size());
    for (size_t i = 0; i < m.size(); ++i) { r[2 * i] = (uint32_t)m[i]; r[2 * i + 1] = (uint32_t)(m[i] >> 32); }
    return r;
  }
  static BigInt from_u32(const std::vector<uint32_t>& d) {
    BigInt r;
    r.mag_.assign((d.size() + 1) / 2, 0);
    for (size_t i = 0; i < d.size(); ++i) r.mag_[i / 2] |= (uint64_t)d[i] << (32 * (i % 2));
    r.trim();
    return r;
  }
};

}  // namespace yh
