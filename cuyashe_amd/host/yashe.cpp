// YASHE over the gfx950 C ABI: the scheme layer of src/yashe.cpp and
// src/ciphertext.cpp, with every ring operation on the device.
#include "yashe.h"

#include <cstring>

using cuyashe::DeviceBuffer;
using cuyashe::runtime;

int Yashe::d = 0;
Polynomial Yashe::phi;
ZZ Yashe::q = ZZ(0);
Integer Yashe::t = Integer(0);
Integer Yashe::delta = Integer(0);
ZZ Yashe::w = ZZ(0);
std::vector<Polynomial> Yashe::gamma;
Polynomial Yashe::h;
Polynomial Yashe::f;
Polynomial Yashe::ff;
int Yashe::lwq = 0;
ZZ Yashe::WDMasking = ZZ(0);
yh_evk* Yashe::evk = nullptr;

int Yashe::w_bits() {
  int b = w.bits() - 1;
  if (b < 1 || ZZ::pow2(b) != w) throw std::invalid_argument("Yashe::w must be a power of two");
  return b;
}

yh_ring* Yashe::ring() { return runtime().ring(phi.deg(), q, t.get_value().to_u64(), w_bits()); }

namespace {

struct RingShape {
  yh_ring* r;
  int n, L;
};

RingShape shape() {
  RingShape s;
  s.r = Yashe::ring();
  yh_ring_info info;
  cuyashe::check(yh_ring_get_info(s.r, &info), "yh_ring_get_info");
  s.n = info.n;
  s.L = info.q_limbs;
  return s;
}

std::shared_ptr<DeviceBuffer> upload_ring(Polynomial& p, const RingShape& s) {
  if (auto d = p.device_buffer(s.L, s.n)) return d;  // already resident
  std::vector<ZZ> c = p.get_coeffs();
  for (auto& v : c) v = v % Yashe::q;
  c.resize(s.n);
  auto host = cuyashe::pack(c, s.L, s.n);
  auto buf = std::make_shared<DeviceBuffer>(host.size() * 8);
  buf->upload(host.data(), host.size() * 8);
  return buf;
}

std::shared_ptr<DeviceBuffer> ring_mul(const std::shared_ptr<DeviceBuffer>& a, const std::shared_ptr<DeviceBuffer>& b,
                                       const RingShape& s) {
  auto out = std::make_shared<DeviceBuffer>((size_t)s.L * s.n * 8);
  cuyashe::check(yh_ring_mul(s.r, a->get(), b->get(), out->get(), 1), "yh_ring_mul");
  return out;
}

Polynomial host_poly(const std::vector<ZZ>& c) {
  Polynomial p(Yashe::q);
  p.set_coeffs(c);
  p.normalize();
  return p;
}

}  // namespace

void Yashe::generate_keys() {
  RingShape s = shape();
  const int n = s.n;
  delta = Integer(q / t.get_value());
  // f = t f' + 1 with f' ternary, resampled until invertible mod (x^n + 1, q)
  Polynomial fInv;
  for (int attempt = 0;; ++attempt) {
    std::vector<ZZ> fc = xkey.sample_mod(n, q);
    for (auto& v : fc) v = (v * t.get_value()) % q;
    fc[0] = (fc[0] + ZZ(1)) % q;
    f = host_poly(fc);
    try {
      fInv = Polynomial::InvMod(f, phi);
      break;
    } catch (const std::domain_error&) {
      if (attempt > 16) throw;
    }
  }
  std::vector<ZZ> gc = xkey.sample_mod(n, q);
  auto d_f = upload_ring(f, s);
  auto d_finv = upload_ring(fInv, s);
  Polynomial tg = host_poly(gc) * t.get_value();
  tg %= q;
  auto d_tg = upload_ring(tg, s);
  h.adopt_device(ring_mul(d_tg, d_finv, s), s.L, n);
  ff.adopt_device(ring_mul(d_f, d_f, s), s.L, n);
  // gamma_k = f w^k + e_k + h s_k mod q (src/yashe.cpp:70-93)
  yh_ring_info info;
  cuyashe::check(yh_ring_get_info(s.r, &info), "yh_ring_get_info");
  lwq = info.lw;
  gamma.assign(lwq, Polynomial(q));
  auto d_h = upload_ring(h, s);
  std::vector<uint64_t> all((size_t)lwq * s.L * n);
  for (int k = 0; k < lwq; ++k) {
    std::vector<ZZ> e = xerr.sample_mod(n, q), sk = xerr.sample_mod(n, q);
    Polynomial skp = host_poly(sk);
    auto d_s = upload_ring(skp, s);
    Polynomial hs;
    hs.adopt_device(ring_mul(d_h, d_s, s), s.L, n);
    ZZ wk = ZZ::pow(w, (uint64_t)k) % q;
    std::vector<ZZ> g(n);
    for (int j = 0; j < n; ++j) g[j] = (f.get_coeff(j) * wk + e[j] + hs.get_coeff(j)) % q;
    gamma[k] = host_poly(g);
    auto packed = cuyashe::pack(g, s.L, n);
    std::memcpy(&all[(size_t)k * s.L * n], packed.data(), packed.size() * 8);
  }
  DeviceBuffer d_gamma(all.size() * 8);
  d_gamma.upload(all.data(), all.size() * 8);
  if (evk) yh_evk_destroy(evk);
  cuyashe::check(yh_evk_create(s.r, d_gamma.get(), &evk), "yh_evk_create");
  runtime().sync();
  WDMasking = w - ZZ(1);
}

Ciphertext Yashe::encrypt(const Polynomial& m) {
  RingShape s = shape();
  const int n = s.n;
  // c = h s + e + delta m mod (x^n + 1, q) (src/yashe.cpp:104-142)
  std::vector<ZZ> sv = xerr.sample_mod(n, q), ev = xerr.sample_mod(n, q);
  Polynomial sp = host_poly(sv);
  auto d_h = upload_ring(h, s);
  auto d_s = upload_ring(sp, s);
  auto hs = ring_mul(d_h, d_s, s);
  Polynomial ep = host_poly(ev);
  auto d_e = upload_ring(ep, s);
  // delta * m on the device (yh_ring_scalar; the reference's Integer
  // product, src/integer.cpp:167-222)
  std::vector<ZZ> mc(n);
  for (int j = 0; j < n; ++j) mc[j] = m.get_coeff(j) % q;
  Polynomial mp = host_poly(mc);
  auto d_m = upload_ring(mp, s);
  const ZZ dv = delta.get_value() % q;
  std::vector<uint64_t> dl(std::max(1, dv.num_limbs()), 0);
  dv.to_limbs(dl.data(), (int)dl.size());
  cuyashe::check(yh_ring_scalar(s.r, d_m->get(), dl.data(), (int)dl.size(), d_m->get(), 1, YH_MUL), "yh_ring_scalar");
  auto out = std::make_shared<DeviceBuffer>((size_t)s.L * n * 8);
  cuyashe::check(yh_ring_addsub(s.r, hs->get(), d_e->get(), out->get(), 1, YH_ADD), "yh_ring_addsub");
  cuyashe::check(yh_ring_addsub(s.r, out->get(), d_m->get(), out->get(), 1, YH_ADD), "yh_ring_addsub");
  Ciphertext c(q);
  c.adopt_device(out, s.L, n);
  return c;
}

Polynomial Yashe::decrypt(const Ciphertext& c) {
  RingShape s = shape();
  const int n = s.n;
  Polynomial key = c.aftermul ? ff : f;
  auto d_key = upload_ring(key, s);
  Ciphertext cc = c;
  auto d_c = upload_ring(cc, s);
  DeviceBuffer m((size_t)n * 8);
  cuyashe::check(yh_yashe_decrypt(s.r, d_key->get(), d_c->get(), m.get(), 1), "yh_yashe_decrypt");
  std::vector<uint64_t> hm(n);
  m.download(hm.data(), n * 8);
  Polynomial out(t.get_value());
  std::vector<ZZ> mc(n);
  for (int j = 0; j < n; ++j) mc[j] = ZZ(hm[j]);
  out.set_coeffs(mc);
  out.normalize();
  return out;
}

// ---------------------------------------------------------------- Ciphertext

namespace {
std::shared_ptr<DeviceBuffer> device_of(Polynomial& p, const RingShape& s) { return upload_ring(p, s); }
}  // namespace

Ciphertext Ciphertext::operator+(Ciphertext& b) {
  RingShape s = shape();
  auto da = device_of(*this, s), db = device_of(b, s);
  auto out = std::make_shared<DeviceBuffer>((size_t)s.L * s.n * 8);
  cuyashe::check(yh_ring_addsub(s.r, da->get(), db->get(), out->get(), 1, YH_ADD), "yh_ring_addsub");
  Ciphertext c(Yashe::q);
  c.adopt_device(out, s.L, s.n);
  c.level = std::max(level, b.level);
  c.aftermul = aftermul;
  return c;
}
Ciphertext Ciphertext::operator+(Polynomial& b) {
  Ciphertext bb(b);
  return *this + bb;
}
Ciphertext& Ciphertext::operator+=(Ciphertext& b) {
  int lv = level;
  bool am = aftermul;
  *this = *this + b;
  level = std::max(lv, b.level);
  aftermul = am;
  return *this;
}
Ciphertext& Ciphertext::operator+=(Polynomial& b) {
  Ciphertext bb(b);
  return *this += bb;
}

Ciphertext Ciphertext::operator*(Ciphertext& b) {
  Ciphertext c1(*this), c2(b);
  c1.level = level; c1.aftermul = aftermul;
  c2.level = b.level; c2.aftermul = b.aftermul;
  if (c1.aftermul) c1.convert();
  if (c2.aftermul) c2.convert();
  RingShape s = shape();
  auto d1 = device_of(c1, s), d2 = device_of(c2, s);
  auto out = std::make_shared<DeviceBuffer>((size_t)s.L * s.n * 8);
  cuyashe::check(yh_yashe_mul(s.r, d1->get(), d2->get(), out->get(), 1), "yh_yashe_mul");
  Ciphertext p(Yashe::q);
  p.adopt_device(out, s.L, s.n);
  p.aftermul = true;
  p.level = std::max(level, b.level) + 1;
  return p;
}

void Ciphertext::convert() {
  keyswitch();
  aftermul = false;
}

void Ciphertext::keyswitch() {
  if (!Yashe::evk) throw std::runtime_error("keyswitch: no keys (call Yashe::generate_keys)");
  RingShape s = shape();
  auto dc = device_of(*this, s);
  auto out = std::make_shared<DeviceBuffer>((size_t)s.L * s.n * 8);
  cuyashe::check(yh_keyswitch(s.r, Yashe::evk, dc->get(), out->get(), 1), "yh_keyswitch");
  int lv = level;
  adopt_device(out, s.L, s.n);
  level = lv;
}
