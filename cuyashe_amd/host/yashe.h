// The YASHE scheme of the drop-in API (src/yashe.h:10-42): statics for the
// parameters and keys, key generation, encryption and decryption.
//
// Parameters follow the reference's names. w is the keyswitch base (a power
// of two, e.g. 2^32); the digit count is ceil(bits(q) / log2(w)) whatever lwq
// a caller computed (the reference's callers use a formula that treats w as a
// bit count, src/benchmark_yashe.cpp:149). Key generation computes
// h = t * g * f^-1 (the reference skips the inverse, src/yashe.cpp:48-50).
#pragma once
#include <vector>

#include "ciphertext.h"
#include "distribution.h"

class Integer {
 public:
  Integer(ZZ a = ZZ(0)) : value_(a) {}  // NOLINT
  Integer(uint64_t a) : value_(ZZ(a)) {}  // NOLINT
  Integer(int a) : value_(ZZ((int64_t)a)) {}  // NOLINT
  ZZ get_value() const { return value_; }
  void set_value(const ZZ& a) { value_ = a; }
  ZZ operator*(const ZZ& a) const { return value_ * a; }
  // Integer::operator*(Polynomial&) (src/integer.cpp:167-222): the scalar
  // product followed by % the polynomial's modulus; on a ring element the
  // product runs on the device (yh_ring_scalar) and is already reduced.
  Polynomial operator*(const Polynomial& p) const {
    Polynomial r = p * value_;
    const ZZ m = p.get_mod();
    if (!r.get_crt_computed() && m > ZZ(1)) r %= m;
    return r;
  }
  Polynomial operator+(const Polynomial& p) const { return p + value_; }

 private:
  ZZ value_;
};

class Yashe {
 public:
  static int d;
  static Polynomial phi;
  static ZZ q;
  static Integer t;
  static Integer delta;
  static ZZ w;
  static std::vector<Polynomial> gamma;
  static Polynomial h;
  static Polynomial f;
  static Polynomial ff;
  static int lwq;
  static ZZ WDMasking;
  // keyswitch key resident on the device (pre-transformed gamma)
  static yh_evk* evk;

  explicit Yashe(uint64_t seed = 1234, double sigma_err = 8.0, int bound = 48)
      : xkey(NARROW, 0, 0, seed), xerr(DISCRETE_GAUSSIAN, sigma_err, bound, seed ^ 0x5DEECE66Dull) {}
  void generate_keys();
  Ciphertext encrypt(const Polynomial& m);
  // Every coefficient of m = [round(t/q * [f c]_q)]_t (the reference decodes
  // coefficient 0 only, src/yashe.cpp:172).
  Polynomial decrypt(const Ciphertext& c);

  // the device ring of the current parameters
  static yh_ring* ring();
  static int w_bits();

 private:
  Distribution xkey, xerr;
};
