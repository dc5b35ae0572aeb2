// Polynomial over the gfx950 C ABI. See polynomial.h for what is kept from
// the reference (src/polynomial.h, src/polynomial.cpp, src/common.h) and
// what changed.
#include "polynomial.h"

#include <algorithm>
#include <cstdlib>
#include <sstream>

#include "zq.h"

namespace cuyashe {

void check(int status, const char* what) {
  if (status != YH_OK) throw std::runtime_error(std::string(what) + ": " + yh_last_error());
}

DeviceBuffer::DeviceBuffer(size_t bytes) : bytes_(bytes) {
  check(yh_malloc(runtime().ctx(), bytes ? bytes : 8, &ptr_), "yh_malloc");
}
DeviceBuffer::~DeviceBuffer() {
  if (ptr_) yh_free(runtime().ctx(), ptr_);
}
void DeviceBuffer::upload(const void* host, size_t bytes) {
  check(yh_memcpy_htod(runtime().ctx(), ptr_, host, bytes), "yh_memcpy_htod");
}
void DeviceBuffer::download(void* host, size_t bytes) const {
  check(yh_memcpy_dtoh(runtime().ctx(), host, ptr_, bytes), "yh_memcpy_dtoh");
}

Runtime& Runtime::get() {
  static Runtime rt;
  return rt;
}
yh_ctx* Runtime::ctx() {
  if (!ctx_) check(yh_ctx_create(0, &ctx_), "yh_ctx_create");
  return ctx_;
}
yh_ring* Runtime::ring(int n, const ZZ& q, uint64_t t, int w_bits) {
  auto key = std::make_tuple(n, q.to_string(), t, w_bits);
  auto it = rings_.find(key);
  if (it != rings_.end()) return it->second;
  std::vector<uint64_t> ql(q.num_limbs());
  q.to_limbs(ql.data(), (int)ql.size());
  yh_ring* r = nullptr;
  check(yh_ring_create(ctx(), n, ql.data(), (int)ql.size(), t, w_bits, &r), "yh_ring_create");
  rings_[key] = r;
  return r;
}
void Runtime::sync() { check(yh_ctx_sync(ctx()), "yh_ctx_sync"); }
Runtime::~Runtime() {
  for (auto& kv : rings_) yh_ring_destroy(kv.second);
  if (ctx_) yh_ctx_destroy(ctx_);
}

std::vector<uint64_t> pack(const std::vector<ZZ>& coefs, int L, int n) {
  std::vector<uint64_t> out((size_t)L * n, 0);
  for (int j = 0; j < n && j < (int)coefs.size(); ++j) {
    const ZZ& c = coefs[j];
    if (c.is_negative()) throw std::invalid_argument("negative coefficient in a device operand");
    if (c.num_limbs() > L) throw std::invalid_argument("coefficient wider than the device operand");
    for (int l = 0; l < L; ++l) out[(size_t)l * n + j] = c.limb(l);
  }
  return out;
}

std::vector<ZZ> unpack(const uint64_t* limbs, int L, int n) {
  std::vector<ZZ> out(n);
  std::vector<uint64_t> w(L);
  for (int j = 0; j < n; ++j) {
    for (int l = 0; l < L; ++l) w[l] = limbs[(size_t)l * n + j];
    out[j] = ZZ::from_limbs(w.data(), L);
  }
  return out;
}

}  // namespace cuyashe

using cuyashe::runtime;

ZZ Polynomial::CRTProduct = ZZ(1);
std::vector<uint64_t> Polynomial::CRTPrimes;
std::vector<ZZ> Polynomial::CRTMpi;
std::vector<uint64_t> Polynomial::CRTInvMpi;
Polynomial* Polynomial::global_phi = nullptr;
ZZ Polynomial::global_mod = ZZ(0);
bool Polynomial::phi_set = false;
int Polynomial::min_ntt_length = 1;

static int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

Polynomial::Polynomial() : mod_(global_mod), phi_(global_phi) {}
Polynomial::Polynomial(ZZ p) : mod_(p), phi_(global_phi) {}
Polynomial::Polynomial(ZZ p, Polynomial P) : mod_(p), phi_(global_phi) { (void)P; }
Polynomial::Polynomial(int spacing) : mod_(global_mod), phi_(global_phi) { (void)spacing; }
Polynomial::Polynomial(ZZ p, Polynomial P, int spacing) : mod_(p), phi_(global_phi) { (void)P; (void)spacing; }
Polynomial::Polynomial(ZZ p, int spacing) : mod_(p), phi_(global_phi) { (void)spacing; }

int Polynomial::ring_degree() { return global_phi ? global_phi->deg() : 0; }

Polynomial Polynomial::get_phi() const { return phi_ ? *phi_ : (global_phi ? *global_phi : Polynomial()); }
void Polynomial::set_phi(const Polynomial& p) {
  phi_ = &p;
  phi_set = true;
}

int Polynomial::deg() const {
  int d = (int)coefs_.size() - 1;
  while (d >= 0 && coefs_[d].is_zero()) --d;
  return d;
}

ZZ Polynomial::get_coeff(int index) const {
  if (index < 0 || index >= (int)coefs_.size()) return ZZ(0);
  return coefs_[index];
}

void Polynomial::set_coeff(int index, const ZZ& value) {
  if (index < 0) throw std::invalid_argument("negative index");
  if (index >= (int)coefs_.size()) {
    if (value.is_zero()) return;
    coefs_.resize(index + 1);
  }
  coefs_[index] = value;
  touched();
}

void Polynomial::set_coeffs(const std::vector<uint64_t>& values) {
  coefs_.assign(values.size(), ZZ(0));
  for (size_t i = 0; i < values.size(); ++i) coefs_[i] = ZZ(values[i]);
  touched();
}
void Polynomial::set_coeffs(const std::vector<ZZ>& values) {
  coefs_ = values;
  touched();
}
void Polynomial::set_coeffs() {
  coefs_.clear();
  if (phi_set && global_phi) coefs_.resize(global_phi->deg() + 1);
  touched();
}
void Polynomial::set_coeffs(int size) {
  coefs_.assign(size > 0 ? size : 0, ZZ(0));
  touched();
}

void Polynomial::normalize() {
  while (!coefs_.empty() && coefs_.back().is_zero()) coefs_.pop_back();
}

std::string Polynomial::to_string() const {
  std::stringstream ss;
  for (int i = 0; i <= deg(); ++i) ss << coefs_[i].to_string() << ", ";
  return ss.str();
}

bool Polynomial::operator==(const Polynomial& b) const {
  int d = deg();
  if (d != b.deg()) return false;
  for (int i = 0; i <= d; ++i)
    if (coefs_[i] != b.coefs_[i]) return false;
  return true;
}

// ---------------------------------------------------------------- host ops
// Addition and subtraction are exact integer operations, like the
// reference's residue add without reduction (src/common.h:24-133).

Polynomial Polynomial::operator+(const Polynomial& b) const {
  Polynomial c(*this);
  c += b;
  return c;
}
Polynomial& Polynomial::operator+=(const Polynomial& b) {
  if (b.coefs_.size() > coefs_.size()) coefs_.resize(b.coefs_.size());
  for (size_t i = 0; i < b.coefs_.size(); ++i) coefs_[i] += b.coefs_[i];
  touched();
  return *this;
}
Polynomial Polynomial::operator-(const Polynomial& b) const {
  Polynomial c(*this);
  c -= b;
  return c;
}
Polynomial& Polynomial::operator-=(const Polynomial& b) {
  if (b.coefs_.size() > coefs_.size()) coefs_.resize(b.coefs_.size());
  for (size_t i = 0; i < b.coefs_.size(); ++i) coefs_[i] -= b.coefs_[i];
  touched();
  return *this;
}
Polynomial Polynomial::operator+(const ZZ& b) const {
  Polynomial c(*this);
  c += b;
  return c;
}
// Scalar operators. On a ring element (a modulus q > 1 from the polynomial
// or global_mod, every coefficient in [0, q), degree below the ring degree)
// they run on the device through yh_ring_scalar and give the result in
// Z_q[x]: the reference runs them as polynomialOPInteger on CRT residues
// (src/polynomial.h:459-573, src/cuda_functions.cu:477-548) and its tests
// compare them with NTL's ZZ_p arithmetic after reduce()
// (src/test.cu:630-790). Other polynomials (no modulus, negative or wider
// coefficients) keep the exact integer operation on the host, the
// reference's host-resident branch of the same operators.
bool Polynomial::device_scalar(const ZZ& b, int op) {
  const ZZ q = mod_.is_zero() ? global_mod : mod_;
  if (q <= ZZ(1)) return false;
  const int d = deg();
  const int n = ring_degree() ? ring_degree() : next_pow2(std::max(d + 1, 2));
  if (n < 2 || n > 16384 || (n & (n - 1)) || d >= n) return false;
  for (int i = 0; i <= d; ++i)
    if (coefs_[i].is_negative() || coefs_[i] >= q) return false;
  yh_ring* r = runtime().ring(n, q);
  yh_ring_info info;
  cuyashe::check(yh_ring_get_info(r, &info), "yh_ring_get_info");
  const int L = info.q_limbs;
  std::vector<ZZ> c(coefs_.begin(), coefs_.begin() + std::max(d + 1, 0));
  c.resize(n);
  auto host = cuyashe::pack(c, L, n);
  cuyashe::DeviceBuffer in(host.size() * 8);
  in.upload(host.data(), host.size() * 8);
  const ZZ s = b % q;  // non-negative (remainder takes the divisor's sign)
  std::vector<uint64_t> sl(std::max(1, s.num_limbs()), 0);
  s.to_limbs(sl.data(), (int)sl.size());
  auto out = std::make_shared<cuyashe::DeviceBuffer>(host.size() * 8);
  cuyashe::check(yh_ring_scalar(r, in.get(), sl.data(), (int)sl.size(), out->get(), 1, op), "yh_ring_scalar");
  adopt_device(out, L, n);
  return true;
}

Polynomial& Polynomial::operator+=(const ZZ& b) {
  if (device_scalar(b, YH_ADD)) return *this;
  set_coeff(0, get_coeff(0) + b);
  return *this;
}
Polynomial Polynomial::operator-(const ZZ& b) const {
  Polynomial c(*this);
  c -= b;
  return c;
}
Polynomial& Polynomial::operator-=(const ZZ& b) {
  if (device_scalar(b, YH_SUB)) return *this;
  set_coeff(0, get_coeff(0) - b);
  return *this;
}
Polynomial Polynomial::operator*(const ZZ& b) const {
  Polynomial c(*this);
  c *= b;
  return c;
}
Polynomial& Polynomial::operator*=(const ZZ& b) {
  if (device_scalar(b, YH_MUL)) return *this;
  for (auto& v : coefs_) v *= b;
  touched();
  return *this;
}
// % and / by a scalar act on each coefficient (host NTL in the reference,
// src/polynomial.h:593-632); % leaves values in [0, b) for b > 0.
Polynomial Polynomial::operator%(const ZZ& b) const {
  Polynomial c(*this);
  c %= b;
  return c;
}
Polynomial& Polynomial::operator%=(const ZZ& b) {
  for (auto& v : coefs_) v = v % b;
  touched();
  return *this;
}
Polynomial Polynomial::operator/(const ZZ& b) const {
  Polynomial c(*this);
  c /= b;
  return c;
}
Polynomial& Polynomial::operator/=(const ZZ& b) {
  for (auto& v : coefs_) v = v / b;
  touched();
  return *this;
}

void Polynomial::MulMod(const ZZ& b, const ZZ& mod) {
  for (auto& v : coefs_) v = (v * b) % mod;
  touched();
}
void Polynomial::CPUAddition(const Polynomial* b) { *this += *b; }
void Polynomial::CPUMulAdd(const Polynomial* b, const ZZ& M) {
  *this += *b;
  *this %= M;
}
void Polynomial::CPUSubtraction(const Polynomial* b) { *this -= *b; }

// ---------------------------------------------------------------- device ops

void Polynomial::update_device_data() {
  if (device_valid_) return;
  const int n = std::max(2, ring_degree() ? ring_degree() : next_pow2(std::max(deg() + 1, 1)));
  const ZZ q = mod_.is_zero() ? global_mod : mod_;
  const int L = std::max(1, (q.bits() + 63) / 64);
  std::vector<ZZ> c = coefs_;
  c.resize(n);
  auto host = cuyashe::pack(c, L, n);
  dev_ = std::make_shared<cuyashe::DeviceBuffer>(host.size() * 8);
  dev_->upload(host.data(), host.size() * 8);
  dev_L_ = L;
  dev_n_ = n;
  device_valid_ = true;
}

void Polynomial::from_device(const uint64_t* limbs, int L, int n) {
  coefs_ = cuyashe::unpack(limbs, L, n);
  normalize();
}

void Polynomial::adopt_device(std::shared_ptr<cuyashe::DeviceBuffer> buf, int L, int n) {
  std::vector<uint64_t> h((size_t)L * n);
  buf->download(h.data(), h.size() * 8);
  from_device(h.data(), L, n);
  res_.reset();
  dev_ = std::move(buf);
  dev_L_ = L;
  dev_n_ = n;
  device_valid_ = true;
}

const uint64_t* Polynomial::device_ring_image() {
  update_device_data();
  return dev_->get();
}

int Polynomial::get_crt_spacing() const {
  int d = deg();
  return next_pow2(std::max(d + 1, 1));
}

uint64_t* Polynomial::get_device_crt_residues() {
  const ZZ q = mod_.is_zero() ? global_mod : mod_;
  if (q <= ZZ(1)) throw std::invalid_argument("no modulus for the CRT");
  const int spacing = std::max(2, get_crt_spacing());
  yh_ring* r = runtime().ring(std::min(spacing, 16384), q);
  yh_ring_info info;
  cuyashe::check(yh_ring_get_info(r, &info), "yh_ring_get_info");
  const int ncoef = deg() + 1;
  int L = 1;
  for (int i = 0; i < ncoef; ++i) L = std::max(L, coefs_[i].num_limbs());
  auto host = cuyashe::pack(std::vector<ZZ>(coefs_.begin(), coefs_.begin() + std::max(ncoef, 0)), L,
                            std::max(ncoef, 1));
  cuyashe::DeviceBuffer in(host.size() * 8);
  in.upload(host.data(), host.size() * 8);
  res_ = std::make_shared<cuyashe::DeviceBuffer>((size_t)info.R * spacing * 8);
  cuyashe::check(yh_crt(r, in.get(), L, std::max(ncoef, 0), spacing, res_->get()), "yh_crt");
  runtime().sync();
  res_spacing_ = spacing;
  return res_->get();
}

Polynomial Polynomial::operator*(const Polynomial& b) const {
  const int da = deg(), db = b.deg();
  Polynomial c(mod_);
  if (da < 0 || db < 0) return c;
  const ZZ qmod = mod_.is_zero() ? global_mod : mod_;
  ZZ maxc(0);
  for (int i = 0; i <= da; ++i) {
    if (coefs_[i].is_negative()) throw std::invalid_argument("GPU product needs non-negative coefficients");
    if (coefs_[i] > maxc) maxc = coefs_[i];
  }
  for (int i = 0; i <= db; ++i) {
    if (b.coefs_[i].is_negative()) throw std::invalid_argument("GPU product needs non-negative coefficients");
    if (b.coefs_[i] > maxc) maxc = b.coefs_[i];
  }
  // The product ring: q bounds the operands (the reference's rule
  // M > 2*degree*q^2 assumes coefficients < q); wider operands use a
  // modulus just above them.
  ZZ q = (qmod > maxc) ? qmod : ZZ::pow2(maxc.bits()) + ZZ(1);
  const int n = next_pow2(std::max(std::max(da, db) + 1, 2));
  yh_ring* r = runtime().ring(n, q);
  yh_ring_info info;
  cuyashe::check(yh_ring_get_info(r, &info), "yh_ring_get_info");
  const int L = info.q_limbs;
  auto ha = cuyashe::pack(std::vector<ZZ>(coefs_.begin(), coefs_.begin() + da + 1), L, da + 1);
  auto hb = cuyashe::pack(std::vector<ZZ>(b.coefs_.begin(), b.coefs_.begin() + db + 1), L, db + 1);
  cuyashe::DeviceBuffer A(ha.size() * 8), B(hb.size() * 8);
  A.upload(ha.data(), ha.size() * 8);
  B.upload(hb.data(), hb.size() * 8);
  const int N = yh_poly_mul_plain_length(da, db, min_ntt_length);
  const int LO = info.work_limbs;
  cuyashe::DeviceBuffer out((size_t)LO * N * 8);
  int Nout = 0;
  cuyashe::check(yh_poly_mul_plain(r, A.get(), da, B.get(), db, min_ntt_length, out.get(), LO, &Nout),
                 "yh_poly_mul_plain");
  if (Nout != N) throw std::runtime_error("product length mismatch");
  std::vector<uint64_t> h((size_t)LO * Nout);
  out.download(h.data(), h.size() * 8);
  c.coefs_ = cuyashe::unpack(h.data(), LO, Nout);
  c.normalize();
  return c;
}

Polynomial& Polynomial::operator*=(const Polynomial& b) {
  *this = *this * b;
  return *this;
}

void Polynomial::reduce() {
  if (!global_phi || !check_special_rem_format(global_phi))
    throw std::runtime_error("Reduce: I don't know how to compute this!");
  const int n = global_phi->deg();
  const ZZ q = global_mod;
  if (q <= ZZ(1)) throw std::runtime_error("reduce needs global_mod");
  std::vector<ZZ> c = coefs_;
  // more than 2n coefficients: fold the tail on the host first
  for (int i = (int)c.size() - 1; i >= 2 * n; --i) {
    int k = i % (2 * n);
    c[k] += ((i / n) % 2 == (k / n) % 2) ? c[i] : -c[i];
    c.pop_back();
  }
  c.resize(2 * n);
  bool any_neg = false;
  int L = 1;
  for (auto& v : c) {
    if (v.is_negative()) any_neg = true;
    L = std::max(L, v.num_limbs());
  }
  if (any_neg) {
    for (auto& v : c) v = v % q;
    L = (q.bits() + 63) / 64;
  }
  yh_ring* r = runtime().ring(n, q);
  yh_ring_info info;
  cuyashe::check(yh_ring_get_info(r, &info), "yh_ring_get_info");
  if (L > info.work_limbs) {
    for (auto& v : c) v = v % q;
    L = info.q_limbs;
  }
  auto host = cuyashe::pack(c, L, 2 * n);
  cuyashe::DeviceBuffer in(host.size() * 8), out((size_t)info.q_limbs * n * 8);
  in.upload(host.data(), host.size() * 8);
  cuyashe::check(yh_ring_reduce(r, in.get(), L, out.get(), 1), "yh_ring_reduce");
  std::vector<uint64_t> h((size_t)info.q_limbs * n);
  out.download(h.data(), h.size() * 8);
  coefs_ = cuyashe::unpack(h.data(), info.q_limbs, n);
  normalize();
  touched();
}

// ---------------------------------------------------------------- statics

void Polynomial::gen_crt_primes(ZZ q, uint64_t degree) {
  int n = next_pow2(std::max<int>((int)degree, 2));
  if (n > 16384) n = 16384;
  std::vector<uint64_t> ql(q.num_limbs());
  q.to_limbs(ql.data(), (int)ql.size());
  yh_ring_info info;
  std::vector<uint64_t> primes(512);
  cuyashe::check(yh_ring_plan(n, ql.data(), (int)ql.size(), 0, 0, &info, primes.data(), 512), "yh_ring_plan");
  primes.resize(info.R);
  CRTPrimes = primes;
  CRTProduct = ZZ(1);
  for (auto p : primes) CRTProduct *= ZZ(p);
  CRTMpi.clear();
  CRTInvMpi.clear();
  for (auto p : primes) {
    ZZ Mi = CRTProduct / ZZ(p);
    CRTMpi.push_back(Mi);
    CRTInvMpi.push_back(ZZ::inv_mod(Mi % ZZ(p), ZZ(p)).to_u64());
  }
}

bool Polynomial::check_special_rem_format(const Polynomial* p) {
  int d = p->deg();
  if (d < 1) return false;
  if (p->get_coeff(0) != ZZ(1) || p->get_coeff(d) != ZZ(1)) return false;
  for (int i = 1; i < d; ++i)
    if (!p->get_coeff(i).is_zero()) return false;
  return true;
}

void Polynomial::DivRem(Polynomial a, Polynomial b, Polynomial& quot, Polynomial& rem) {
  if (a.deg() <= 0) return;
  if (!check_special_rem_format(&b)) throw std::runtime_error("DivRem: I don't know how to div this!");
  const int half = b.deg() - 1;
  rem.set_coeffs(half + 1);
  for (int i = 0; i <= half; ++i) rem.set_coeff(i, a.get_coeff(i) - a.get_coeff(i + half + 1));
  quot.set_coeffs(0);
}

Polynomial Polynomial::operator/(const Polynomial& b) const {
  Polynomial quot, rem;
  DivRem(*this, b, quot, rem);
  return quot;
}
Polynomial Polynomial::operator%(const Polynomial& b) const {
  Polynomial quot, rem;
  DivRem(*this, b, quot, rem);
  return rem;
}

void Polynomial::BuildNthCyclotomic(Polynomial* phi, unsigned int n) {
  if (n == 0 || (n & (n - 1))) throw std::invalid_argument("only x^n + 1 (n a power of two) is supported");
  phi->set_coeffs();
  phi->coefs_.clear();
  phi->set_coeff(0, ZZ(1));
  phi->set_coeff((int)n, ZZ(1));
}

Polynomial Polynomial::InvMod(Polynomial a, Polynomial b) {
  if (!check_special_rem_format(&b)) throw std::invalid_argument("InvMod supports x^n + 1 only");
  const ZZ q = a.mod_.is_zero() ? global_mod : a.mod_;
  yh::Zq Z(q);
  std::vector<ZZ> out;
  if (!yh::poly_inverse_negacyclic(Z, a.coefs_, b.deg(), out)) throw std::domain_error("not invertible");
  Polynomial r(q);
  r.set_coeffs(out);
  r.normalize();
  return r;
}

void Polynomial::random(Polynomial* a, int degree) {
  const ZZ m = a->mod_ > ZZ(0) ? a->mod_ : global_mod;
  a->set_coeffs(degree + 1);
  for (int i = 0; i <= degree; ++i) a->set_coeff(i, ZZ((int64_t)rand()) % m);
  a->normalize();
}
