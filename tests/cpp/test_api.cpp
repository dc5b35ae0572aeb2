// Drop-in API checks: the reference's own test programs (src/test_mul.cu and
// the PolySuite / YasheSuite cases of src/test.cu) written against the
// reference's Polynomial / Ciphertext / Yashe interface, running on the GPU.
// Expected values are the reference's literal KAT data and closed forms on
// ZZ. Prints one line per case; exit status 0 iff all pass.
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

#include "../../cuyashe_amd/host/yashe.h"

static int failures = 0;
static void run(const char* name, const std::function<bool()>& f) {
  std::fprintf(stderr, "[case] %s\n", name);
  bool ok = false;
  try {
    ok = f();
  } catch (const std::exception& e) {
    std::printf("ERROR %s: %s\n", name, e.what());
  }
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
  std::fflush(stdout);
  if (!ok) ++failures;
}

static std::vector<ZZ> negacyclic_mod(const std::vector<ZZ>& a, const std::vector<ZZ>& b, int n, const ZZ& q) {
  std::vector<ZZ> out(n, ZZ(0));
  for (size_t i = 0; i < a.size(); ++i)
    for (size_t j = 0; j < b.size(); ++j) {
      size_t k = i + j;
      if ((int)k >= n) out[k - n] -= a[i] * b[j];
      else out[k] += a[i] * b[j];
    }
  for (auto& v : out) v = v % q;
  return out;
}

static void setup_ring(const ZZ& q, int degree, Polynomial& phi) {
  Polynomial::global_mod = q;
  Polynomial::BuildNthCyclotomic(&phi, degree);
  phi.set_mod(q);
  Polynomial::global_phi = &phi;
  Polynomial::gen_crt_primes(q, degree);
}

int main() {
  // ------------------------------------------------ src/test_mul.cu
  {
    Polynomial phi;
    setup_ring(to_ZZ("1171313591017775093490277364417"), 8, phi);
    srand(36251);
    Polynomial a, b, c;
    Polynomial::random(&a, 8);
    Polynomial::random(&b, 8);
    run("test_mul first (cyclic wrap mod x^16 - 1)", [&] {
      c = a * b;
      const char* expected[] = {"52549502886738203", "729378972269732879", "3315814916818960945",
                                "5373031442410099707", "5487719755647497370", "6002790757075884044",
                                "7641866224560971944", "8444884424660384993", "10787872446794447944",
                                "10671875472706824289", "7566104911804466568", "5469026184047188285",
                                "4332423149998847987", "2715876362046054824", "2047760673349179087",
                                "1502584925726431011"};
      if (c.get_coeffs().size() != 16) return false;
      for (int i = 0; i < 16; ++i)
        if (c.get_coeff(i) != to_ZZ(expected[i])) return false;
      return true;
    });
    run("test_mul second (exact product)", [&] {
      a.set_coeffs(8);
      const long av[] = {1304768405, 1912295543, 584306160, 476779113, 2057832156, 1274012458, 1818971124, 1943566066};
      for (int i = 0; i < 8; ++i) a.set_coeff(i, to_ZZ(av[i]));
      const char* bv[] = {"13860721405712617825882419437", "911979968073012785910226533",
                          "60134818568832314735148266017", "30228208549061091796489949751",
                          "42496336825615196107672318132", "2538357487121433195951809389",
                          "19401757472799351763805908821", "68415359680988770693366249506"};
      b.set_coeffs(8);
      for (int i = 0; i < 8; ++i) b.set_coeff(i, to_ZZ(bv[i]));
      c = a * b;
      const char* expected[] = {"18085031360681010249051172126355487985", "27695718415243509623513993478082759156",
                                "88304891436672768443904444999435867224", "161577734953354918933760345835324042047",
                                "177348142234080397165460901588496065374", "150446511525560453344956466169095450596",
                                "219533428887126844535620298106253625669", "315527818352347626764175594181751346765",
                                "380494624177053395698844189209855453911", "280450476250800780083801312965924383119",
                                "211828603909665970332138017649830954350", "252717044998320508082871060193586798902",
                                "127386722625350085506673181613116824226", "162154161142687163774421077602378932930",
                                "132969771469154360046681933849550863396"};
      if (c.get_coeffs().size() != 15) return false;
      for (int i = 0; i < 15; ++i)
        if (c.get_coeff(i) != to_ZZ(expected[i])) return false;
      return true;
    });
  }
  // ------------------------------------------------ src/test.cu PolySuite (n = 8)
  {
    Polynomial phi;
    const ZZ q = to_ZZ("77287149995192912462927307869");
    const int degree = 8;
    setup_ring(q, degree, phi);
    run("justCRT (residues on the device, back to coefficients)", [&] {
      srand(36251);
      for (int count = 0; count < 10; ++count) {
        Polynomial a;
        Polynomial::random(&a, degree - 1);
        uint64_t* res = a.get_device_crt_residues();
        yh_ring* r = cuyashe::runtime().ring(a.get_crt_spacing(), q);
        yh_ring_info info;
        cuyashe::check(yh_ring_get_info(r, &info), "info");
        int sp = a.get_crt_spacing();
        cuyashe::DeviceBuffer out((size_t)info.work_limbs * sp * 8);
        cuyashe::check(yh_icrt(r, res, sp, sp, out.get(), info.work_limbs), "yh_icrt");
        std::vector<uint64_t> h((size_t)info.work_limbs * sp);
        out.download(h.data(), h.size() * 8);
        Polynomial b;
        b.from_device(h.data(), info.work_limbs, sp);
        if (!(a == b)) return false;
      }
      return true;
    });
    run("multiplyByPolynomial (a*b; reduce(); %= q)", [&] {
      srand(36251);
      for (int count = 0; count < 10; ++count) {
        Polynomial a, b;
        Polynomial::random(&a, degree - 1);
        Polynomial::random(&b, degree - 1);
        Polynomial c = a * b;
        c.reduce();
        c %= q;
        auto want = negacyclic_mod(a.get_coeffs(), b.get_coeffs(), degree, q);
        Polynomial w;
        w.set_coeffs(want);
        w.normalize();
        if (!(c == w)) return false;
      }
      return true;
    });
    run("multiplyAndAddByPolynomial (a*b + a; reduce(); %= q)", [&] {
      srand(36251);
      for (int count = 0; count < 10; ++count) {
        Polynomial a, b;
        Polynomial::random(&a, degree - 1);
        Polynomial::random(&b, degree - 1);
        Polynomial c = a * b + a;
        c.reduce();
        c %= q;
        auto want = negacyclic_mod(a.get_coeffs(), b.get_coeffs(), degree, q);
        for (int i = 0; i < degree; ++i) want[i] = (want[i] + a.get_coeff(i)) % q;
        Polynomial w;
        w.set_coeffs(want);
        w.normalize();
        if (!(c == w)) return false;
      }
      return true;
    });
    run("severalMultiplications (c = c*a; reduce())", [&] {
      srand(36251);
      Polynomial a, b;
      Polynomial::random(&a, degree - 1);
      Polynomial::random(&b, degree - 1);
      Polynomial c = a * b;
      c.reduce();
      auto want = negacyclic_mod(a.get_coeffs(), b.get_coeffs(), degree, q);
      for (int it = 0; it < 10; ++it) {
        Polynomial w;
        w.set_coeffs(want);
        w.normalize();
        if (!(c == w)) return false;
        c = c * a;
        c.reduce();
        want = negacyclic_mod(want, a.get_coeffs(), degree, q);
      }
      return true;
    });
    run("phiReduce (degree 2n-2 input, reduce(), %= q)", [&] {
      srand(36251);
      for (int count = 0; count < 10; ++count) {
        Polynomial a;
        Polynomial::random(&a, 2 * degree - 2);
        std::vector<ZZ> want(degree);
        for (int i = 0; i < degree; ++i) want[i] = (a.get_coeff(i) - a.get_coeff(i + degree)) % q;
        a.reduce();
        a %= q;
        Polynomial w;
        w.set_coeffs(want);
        w.normalize();
        if (!(a == w)) return false;
      }
      return true;
    });
    run("randomPolynomialOperations (scalar *= and += on the device vs Z_q arithmetic)", [&] {
      // src/test.cu:630-790 with random_ab = 3: c = a + b, then random
      // c *= z or c += z (z = rand() % q), reduce(), compared with the same
      // operations mod q; scalar operators on a reduced ring element run on
      // the device and leave the device copy valid
      srand(36251);
      Polynomial a, b;
      Polynomial::random(&a, degree - 1);
      Polynomial::random(&b, degree - 1);
      Polynomial c = a + b;
      std::vector<ZZ> want(degree);
      for (int i = 0; i < degree; ++i) want[i] = (a.get_coeff(i) + b.get_coeff(i)) % q;
      int on_device = 0;
      for (int count = 0; count < 40; ++count) {
        const int op = rand() % 3;
        ZZ z = ZZ((int64_t)rand()) * ZZ((int64_t)rand()) * ZZ((int64_t)rand()) % q;
        if (op == 0) {
          c *= z;
          for (auto& v : want) v = (v * z) % q;
        } else if (op == 1) {
          c += z;
          want[0] = (want[0] + z) % q;
        } else {
          c -= z;
          want[0] = (want[0] - z) % q;
        }
        on_device += c.get_crt_computed() ? 1 : 0;
        c.reduce();
        c.normalize();
        Polynomial w;
        w.set_coeffs(want);
        w.normalize();
        if (!(c == w)) return false;
      }
      // every operation after the first reduce() ran on the device
      if (on_device < 39) return false;
      // Integer::operator* (src/integer.cpp:167-222): product then % mod
      Integer t((uint64_t)35951);
      Polynomial tc = t * c;
      for (int i = 0; i < degree; ++i)
        if (tc.get_coeff(i) != (c.get_coeff(i) * ZZ((int64_t)35951)) % q) return false;
      // a polynomial without a valid ring image keeps exact integer arithmetic
      Polynomial big;
      big.set_coeff(0, q * ZZ((int64_t)3));
      big *= ZZ((int64_t)5);
      return big.get_coeff(0) == q * ZZ((int64_t)15) && !big.get_crt_computed();
    });
    run("modularInversion (a * InvMod(a, phi) == 1)", [&] {
      srand(36251);
      Polynomial a;
      Polynomial::random(&a, degree - 1);
      Polynomial ai = Polynomial::InvMod(a, phi);
      Polynomial r = a * ai;
      r.reduce();
      r %= q;
      return r.is_one();
    });
  }
  // ------------------------------------------------ src/test.cu YasheSuite (n = 4, t = 35951)
  {
    Polynomial phi;
    const ZZ q = to_ZZ("77287149995192912462927307869");
    setup_ring(q, 4, phi);
    Yashe cipher(36251);
    Yashe::d = 4;
    Yashe::phi = phi;
    Yashe::q = q;
    Yashe::t = Integer((uint64_t)35951);
    Yashe::w = ZZ::pow2(32);
    cipher.generate_keys();
    srand(36251);
    run("encryptDecrypt", [&] {
      for (int i = 0; i < 10; ++i) {
        Polynomial a;
        a.set_coeff(0, ZZ((int64_t)rand()) % ZZ((int64_t)35951));
        Ciphertext c = cipher.encrypt(a);
        if (!(cipher.decrypt(c) == a)) return false;
      }
      return true;
    });
    run("encryptandAdd", [&] {
      for (int i = 0; i < 10; ++i) {
        Polynomial a;
        a.set_coeff(0, ZZ((int64_t)rand()) % ZZ((int64_t)35951));
        Ciphertext c = cipher.encrypt(a);
        c = c + c;
        Polynomial want = (a + a) % ZZ((int64_t)35951);
        if (!(cipher.decrypt(c) == want)) return false;
      }
      return true;
    });
    run("encryptandMul (decrypts with f^2 before keyswitch, f after)", [&] {
      for (int i = 0; i < 10; ++i) {
        Polynomial a;
        a.set_coeff(0, ZZ((int64_t)rand()) % ZZ((int64_t)35951));
        Ciphertext c1 = cipher.encrypt(a), c2 = cipher.encrypt(a);
        Ciphertext c = c1 * c2;
        Polynomial want = (a * a) % ZZ((int64_t)35951);
        if (!(cipher.decrypt(c) == want)) return false;
        c.convert();
        if (!(cipher.decrypt(c) == want)) return false;
      }
      return true;
    });
  }
  // ------------------------------------------------ C2: n = 4096, q = 2^127 - 1, t = 1024, w = 2^32
  {
    Polynomial phi;
    const ZZ q = ZZ::pow2(127) - ZZ(1);
    const int n = 4096;
    setup_ring(q, n, phi);
    Yashe cipher(20240613);
    Yashe::d = n;
    Yashe::phi = phi;
    Yashe::q = q;
    Yashe::t = Integer((uint64_t)1024);
    Yashe::w = ZZ::pow2(32);
    cipher.generate_keys();
    run("C2 full-polynomial mul + keyswitch decrypts to m1*m2 mod (x^n+1, t)", [&] {
      Distribution msg(UNIFORMLY, 0, 0, 7);
      std::vector<ZZ> m1 = msg.sample_mod(n, ZZ((int64_t)1024)), m2 = msg.sample_mod(n, ZZ((int64_t)1024));
      Polynomial p1, p2;
      p1.set_coeffs(m1);
      p2.set_coeffs(m2);
      Ciphertext c1 = cipher.encrypt(p1), c2 = cipher.encrypt(p2);
      Ciphertext c = c1 * c2;
      c.convert();
      Polynomial got = cipher.decrypt(c);
      auto want = negacyclic_mod(m1, m2, n, ZZ((int64_t)1024));
      Polynomial w;
      w.set_coeffs(want);
      w.normalize();
      if (!(got == w)) return false;
      Ciphertext s = c1 + c2;
      Polynomial gs = cipher.decrypt(s);
      for (int j = 0; j < n; ++j)
        if (gs.get_coeff(j) != (m1[j] + m2[j]) % ZZ((int64_t)1024)) return false;
      return true;
    });
  }
  // ------------------------------------------------ operands of degree 16383
  {
    const ZZ q = to_ZZ("170141183460469231731687303715884105727");
    Polynomial phi;
    setup_ring(q, 16384, phi);
    run("operator* of two degree-16383 polynomials (N = 32768, exact over Z)", [&] {
      srand(4242);
      Polynomial a, b;
      Polynomial::random(&a, 16383);
      Polynomial::random(&b, 16383);
      a.set_coeff(16383, q - ZZ(1));
      b.set_coeff(16383, q - ZZ(1));
      const Polynomial c = a * b;
      if (c.deg() != 32766) return false;
      // evaluation at r modulo m is a ring map Z[x] -> Z_m, so an exact
      // product satisfies c(r) = a(r) b(r) (mod m) for every r and m
      auto eval = [](const Polynomial& p, const ZZ& r, const ZZ& m) {
        ZZ v(0);
        for (int k = p.deg(); k >= 0; --k) v = (v * r + p.get_coeff(k)) % m;
        return v;
      };
      for (const ZZ& m : {ZZ::pow2(61) - ZZ(1), to_ZZ("1000000000000000009")})
        for (const ZZ& r : {ZZ(3), to_ZZ("1234567891011"), m - ZZ(2)})
          if (eval(c, r, m) != (eval(a, r, m) * eval(b, r, m)) % m) return false;
      return true;
    });
  }
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
