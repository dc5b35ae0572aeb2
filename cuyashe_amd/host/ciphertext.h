// Ciphertext of the drop-in API (src/ciphertext.h:6-170): a ring element
// mod (x^n + 1, q) with the reference's level / aftermul bookkeeping. The
// homomorphic operations run on the GPU through the C ABI.
#pragma once
#include "polynomial.h"

class Ciphertext : public Polynomial {
 public:
  Ciphertext() : Polynomial() {}
  explicit Ciphertext(ZZ p) : Polynomial(p) {}
  Ciphertext(const Polynomial& p) : Polynomial(p) {}  // NOLINT: as the reference
  Ciphertext& operator=(const Polynomial& p) {
    Polynomial::operator=(p);
    level = 0;
    aftermul = false;
    return *this;
  }
  // c1 + c2 mod q on the device (common_addition, src/common.h:24-80, with the
  // reduction mod q that keeps ciphertexts in [0, q)).
  Ciphertext operator+(Ciphertext& b);
  Ciphertext operator+(Polynomial& b);
  Ciphertext& operator+=(Ciphertext& b);
  Ciphertext& operator+=(Polynomial& b);
  // YASHE multiply [round(t/q * c1*c2)]_q (src/ciphertext.cpp:38-73). As in the
  // reference the result is left before relinearisation (aftermul = true) and
  // decrypts with f^2; operands that are themselves aftermul are keyswitched
  // first (convert()).
  Ciphertext operator*(Ciphertext& b);
  // Keyswitch (relinearisation) with Yashe::gamma (src/ciphertext.cpp:75-107).
  void convert();
  int level = 0;
  bool aftermul = false;

 private:
  void keyswitch();
};
