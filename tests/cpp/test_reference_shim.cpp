// The reference's device-layer interface bound to the gfx950 library
// through cuyashe_amd/host/reference_shim.cpp, exercised the way the
// reference's own product path uses it: coefficients as bn_t on the device
// (Polynomial::update_device_data, src/polynomial.cpp:265-340), callCRT,
// callPolynomialMul at N = 2^ceil(log2(deg a + deg b)) (common_multiplication,
// src/common.h:136-226), callICRT. The two known answers of src/test_mul.cu
// (:46-50 with its cyclic wrap, :84-88 exact) must come out bit for bit;
// callNTT, callPolynomialAddSub and callPolynomialOPInteger are checked
// against closed forms. Prints one line per case; exit status 0 iff all pass.
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <vector>

#include "../../cuyashe_amd/host/reference_shim.h"
#include "../../cuyashe_amd/host/zz.h"

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

static void check(int st, const char* what) {
  if (st != YH_OK) throw std::runtime_error(std::string(what) + ": " + yh_last_error());
}

// Device coefficients in the reference's layout: a device array of bn_t
// whose dp point into one device word buffer (alloc words each).
struct DeviceBn {
  int n = 0, alloc = 0;
  void* words = nullptr;
  void* headers = nullptr;
  DeviceBn(const std::vector<ZZ>& coefs, int n_, int alloc_) : n(n_), alloc(alloc_) {
    yh_ctx* ctx = reference_shim::context();
    std::vector<uint64_t> w((size_t)n * alloc, 0);
    std::vector<bn_t> h(n);
    check(yh_malloc(ctx, w.size() * 8, &words), "yh_malloc");
    check(yh_malloc(ctx, h.size() * sizeof(bn_t), &headers), "yh_malloc");
    for (int j = 0; j < n; ++j) {
      const ZZ c = j < (int)coefs.size() ? coefs[j] : ZZ(0);
      c.to_limbs(&w[(size_t)j * alloc], alloc);
      h[j].alloc = alloc;
      h[j].used = c.num_limbs();
      h[j].sign = 0;
      h[j].dp = (uint64_t*)words + (size_t)j * alloc;
    }
    check(yh_memcpy_htod(ctx, words, w.data(), w.size() * 8), "yh_memcpy_htod");
    check(yh_memcpy_htod(ctx, headers, h.data(), h.size() * sizeof(bn_t)), "yh_memcpy_htod");
  }
  ~DeviceBn() {
    yh_free(reference_shim::context(), words);
    yh_free(reference_shim::context(), headers);
  }
  bn_t* get() { return (bn_t*)headers; }
  // coefficient values read back through the headers (used words only)
  std::vector<ZZ> values() {
    yh_ctx* ctx = reference_shim::context();
    std::vector<bn_t> h(n);
    std::vector<uint64_t> w((size_t)n * alloc);
    check(yh_memcpy_dtoh(ctx, h.data(), headers, h.size() * sizeof(bn_t)), "yh_memcpy_dtoh");
    check(yh_memcpy_dtoh(ctx, w.data(), words, w.size() * 8), "yh_memcpy_dtoh");
    std::vector<ZZ> out(n);
    for (int j = 0; j < n; ++j) {
      if (h[j].sign != 0 || h[j].used < 0 || h[j].used > alloc) throw std::runtime_error("bad bn_t header");
      out[j] = ZZ::from_limbs(&w[(size_t)j * alloc], h[j].used);
    }
    return out;
  }
};

static void* dev_alloc(size_t bytes) {
  void* p = nullptr;
  check(yh_malloc(reference_shim::context(), bytes, &p), "yh_malloc");
  return p;
}

// c = a * b through callCRT -> callPolynomialMul -> callICRT, N per the
// reference's rule; returns N coefficients.
static std::vector<ZZ> product(const std::vector<ZZ>& a, const std::vector<ZZ>& b) {
  int da = (int)a.size() - 1, db = (int)b.size() - 1;
  int N = 1;
  while (N < da + db) N <<= 1;
  CUDAFunctions::init(N);
  const int NPolis = (int)reference_shim::crt_primes().size();
  const int alloc = CUDAFunctions::std_bn_t_alloc;
  DeviceBn A(a, N, alloc), B(b, N, alloc), C({}, N, alloc);
  yh_ctx* ctx = reference_shim::context();
  cuyasheint_t* ra = (cuyasheint_t*)dev_alloc((size_t)N * NPolis * 8);
  cuyasheint_t* rb = (cuyasheint_t*)dev_alloc((size_t)N * NPolis * 8);
  callCRT(A.get(), (int)a.size(), ra, N, NPolis, nullptr);
  callCRT(B.get(), (int)b.size(), rb, N, NPolis, nullptr);
  cuyasheint_t* rc = CUDAFunctions::callPolynomialMul(nullptr, ra, false, N, rb, false, N, N, NPolis);
  callICRT(C.get(), rc, N, NPolis, nullptr);
  std::vector<ZZ> out = C.values();
  yh_free(ctx, ra);
  yh_free(ctx, rb);
  yh_free(ctx, rc);
  return out;
}

static std::vector<ZZ> zz(const std::vector<const char*>& v) {
  std::vector<ZZ> out;
  for (auto s : v) out.push_back(to_ZZ(s));
  return out;
}

int main() {
  const ZZ q = to_ZZ("1171313591017775093490277364417");  // src/test_mul.cu:21
  std::vector<uint64_t> ql(q.num_limbs());
  q.to_limbs(ql.data(), (int)ql.size());
  reference_shim::gen_crt_primes(ql.data(), (int)ql.size(), 16);
  CUDAFunctions::write_crt_primes();

  run("KAT1 of src/test_mul.cu (degree-8 operands of srand(36251), product mod x^16 - 1)", [&] {
    // the operands Polynomial::random draws after srand(36251): glibc rand() mod q
    srand(36251);
    std::vector<ZZ> a(9), b(9);
    for (auto& v : a) v = ZZ((int64_t)rand()) % q;
    for (auto& v : b) v = ZZ((int64_t)rand()) % q;
    std::vector<ZZ> want = zz({"52549502886738203", "729378972269732879", "3315814916818960945",
                               "5373031442410099707", "5487719755647497370", "6002790757075884044",
                               "7641866224560971944", "8444884424660384993", "10787872446794447944",
                               "10671875472706824289", "7566104911804466568", "5469026184047188285",
                               "4332423149998847987", "2715876362046054824", "2047760673349179087",
                               "1502584925726431011"});
    return product(a, b) == want;
  });

  run("KAT2 of src/test_mul.cu (exact product of degree-7 operands)", [&] {
    std::vector<ZZ> a = zz({"1304768405", "1912295543", "584306160", "476779113", "2057832156", "1274012458",
                            "1818971124", "1943566066"});
    std::vector<ZZ> b = zz({"13860721405712617825882419437", "911979968073012785910226533",
                            "60134818568832314735148266017", "30228208549061091796489949751",
                            "42496336825615196107672318132", "2538357487121433195951809389",
                            "19401757472799351763805908821", "68415359680988770693366249506"});
    std::vector<ZZ> want = zz({"18085031360681010249051172126355487985", "27695718415243509623513993478082759156",
                               "88304891436672768443904444999435867224", "161577734953354918933760345835324042047",
                               "177348142234080397165460901588496065374", "150446511525560453344956466169095450596",
                               "219533428887126844535620298106253625669", "315527818352347626764175594181751346765",
                               "380494624177053395698844189209855453911", "280450476250800780083801312965924383119",
                               "211828603909665970332138017649830954350", "252717044998320508082871060193586798902",
                               "127386722625350085506673181613116824226", "162154161142687163774421077602378932930",
                               "132969771469154360046681933849550863396"});
    std::vector<ZZ> got = product(a, b);
    want.push_back(ZZ(0));  // N = 16 coefficients, the last one zero
    return got == want;
  });

  run("callNTT forward/inverse round trip and the DFT of a delta (src/cuda_functions.cu:447-458)", [&] {
    const int N = 32, NPolis = 2;
    yh_ctx* ctx = reference_shim::context();
    std::vector<uint64_t> h((size_t)N * NPolis, 0);
    h[1] = 1;           // row 0: x -> X_k = w^k
    h[N + 0] = 12345;   // row 1: a constant -> X_k = 12345
    cuyasheint_t* dI = (cuyasheint_t*)dev_alloc(h.size() * 8);
    cuyasheint_t* dO = (cuyasheint_t*)dev_alloc(h.size() * 8);
    check(yh_memcpy_htod(ctx, dI, h.data(), h.size() * 8), "htod");
    CUDAFunctions::callNTT(N, NPolis, dI, dO, FORWARD);  // log2 32 = 5: result in dataO
    std::vector<uint64_t> f(h.size());
    check(yh_memcpy_dtoh(ctx, f.data(), dO, f.size() * 8), "dtoh");
    const uint64_t P = 0xFFFFFFFF00000001ull;
    const ZZ w = ZZ::pow_mod(ZZ(7), ZZ((P - 1) / N), ZZ(P));
    bool good = true;
    ZZ wk(1);
    for (int k = 0; k < N; ++k) {
      good &= ZZ(f[k]) == wk && f[N + k] == 12345;
      wk = (wk * w) % ZZ(P);
    }
    CUDAFunctions::callNTT(N, NPolis, dO, dI, INVERSE);
    std::vector<uint64_t> back(h.size());
    check(yh_memcpy_dtoh(ctx, back.data(), dI, back.size() * 8), "dtoh");
    yh_free(ctx, dI);
    yh_free(ctx, dO);
    return good && back == h;
  });

  run("callPolynomialAddSub and callPolynomialOPInteger on residues, back through callICRT", [&] {
    const int N = 16;
    const int NPolis = (int)reference_shim::crt_primes().size();
    const std::vector<uint64_t>& pr = reference_shim::crt_primes();
    CUDAFunctions::init(N);
    yh_ctx* ctx = reference_shim::context();
    std::vector<ZZ> a(N), b(N);
    for (int j = 0; j < N; ++j) {
      a[j] = (ZZ((int64_t)(j + 1)) * to_ZZ("1000000000000000000007")) % q;
      b[j] = (ZZ((int64_t)(3 * j + 2)) * to_ZZ("99999999999999999989")) % q;
    }
    const int alloc = CUDAFunctions::std_bn_t_alloc;
    DeviceBn A(a, N, alloc), B(b, N, alloc), C({}, N, alloc);
    cuyasheint_t* ra = (cuyasheint_t*)dev_alloc((size_t)N * NPolis * 8);
    cuyasheint_t* rb = (cuyasheint_t*)dev_alloc((size_t)N * NPolis * 8);
    cuyasheint_t* rs = (cuyasheint_t*)dev_alloc((size_t)NPolis * 8);
    callCRT(A.get(), N, ra, N, NPolis, nullptr);
    callCRT(B.get(), N, rb, N, NPolis, nullptr);
    CUDAFunctions::callPolynomialAddSubInPlace(nullptr, ra, rb, N * NPolis, ADD);
    callICRT(C.get(), ra, N, NPolis, nullptr);
    std::vector<ZZ> sum = C.values();
    bool good = true;
    for (int j = 0; j < N; ++j) good &= sum[j] == a[j] + b[j];
    // the scalar s as Integer's residues (Integer::crt, src/integer.cpp:62-82)
    const ZZ s = to_ZZ("123456789012345678901234567");
    std::vector<uint64_t> sres(NPolis);
    for (int i = 0; i < NPolis; ++i) sres[i] = (s % ZZ(pr[i])).to_u64();
    check(yh_memcpy_htod(ctx, rs, sres.data(), sres.size() * 8), "htod");
    cuyasheint_t* rm = CUDAFunctions::callPolynomialOPInteger(MUL, nullptr, rb, rs, N, NPolis);
    callICRT(C.get(), rm, N, NPolis, nullptr);
    std::vector<ZZ> prod = C.values();
    for (int j = 0; j < N; ++j) good &= prod[j] == b[j] * s;  // below M: exact
    cuyasheint_t* rp = CUDAFunctions::callPolynomialOPInteger(ADD, nullptr, rb, rs, N, NPolis);
    callICRT(C.get(), rp, N, NPolis, nullptr);
    std::vector<ZZ> plus = C.values();
    good &= plus[0] == b[0] + s;
    for (int j = 1; j < N; ++j) good &= plus[j] == b[j];
    for (void* p : {(void*)ra, (void*)rb, (void*)rs, (void*)rm, (void*)rp}) yh_free(ctx, p);
    return good;
  });

  reference_shim::release();
  std::printf("%s: %d failure(s)\n", failures ? "FAILED" : "OK", failures);
  return failures ? 1 : 0;
}
