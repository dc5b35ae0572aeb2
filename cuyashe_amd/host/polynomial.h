// Host data model of the drop-in API: Polynomial, mirroring the reference's
// class (src/polynomial.h:31-1441) over the gfx950 C ABI (include/yashe_hip.h).
//
// Kept from the reference: method names, operators, argument meaning, the
// static ring parameters (global_mod, global_phi, CRTPrimes, CRTProduct...),
// the NTT-length rule of products and the host/device residency flags.
// Changed: coefficients stay exact host ZZ values with a cached device copy
// (limb-major u64[L][n]) instead of bn_t headers and per-object streams; the
// device work runs through one process-wide context (cuyashe::runtime()).
// Errors throw std::runtime_error with the library's message (the reference
// asserts).
#pragma once
#include <stdint.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/yashe_hip.h"
#include "zz.h"

namespace cuyashe {

void check(int status, const char* what);

// Device buffer owned by the runtime context.
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t bytes);
  ~DeviceBuffer();
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : ptr_(o.ptr_), bytes_(o.bytes_) { o.ptr_ = nullptr; o.bytes_ = 0; }
  uint64_t* get() const { return (uint64_t*)ptr_; }
  size_t bytes() const { return bytes_; }
  void upload(const void* host, size_t bytes);
  void download(void* host, size_t bytes) const;

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
};

// Process-wide device context and ring cache (the reference keeps these in
// statics: CUDAFunctions::N, Polynomial::CRTPrimes, ...).
class Runtime {
 public:
  static Runtime& get();
  yh_ctx* ctx();
  // Ring x^n + 1 over q, with plaintext modulus t and keyswitch base 2^w_bits.
  yh_ring* ring(int n, const ZZ& q, uint64_t t = 0, int w_bits = 0);
  void sync();
  ~Runtime();

 private:
  Runtime() = default;
  yh_ctx* ctx_ = nullptr;
  std::map<std::tuple<int, std::string, uint64_t, int>, yh_ring*> rings_;
};

inline Runtime& runtime() { return Runtime::get(); }

// Packs coefficients (non-negative, < 2^(64 L)) limb-major [L][n].
std::vector<uint64_t> pack(const std::vector<ZZ>& coefs, int L, int n);
std::vector<ZZ> unpack(const uint64_t* limbs, int L, int n);

}  // namespace cuyashe

class Polynomial {
 public:
  // ---- ring parameters (src/polynomial.h:33-41)
  static ZZ CRTProduct;
  static std::vector<uint64_t> CRTPrimes;
  static std::vector<ZZ> CRTMpi;
  static std::vector<uint64_t> CRTInvMpi;
  static Polynomial* global_phi;
  static ZZ global_mod;
  static bool phi_set;
  // NTT length floor of products (CUDAFunctions::N, src/cuda_functions.h:21)
  static int min_ntt_length;

  // ---- constructors (src/polynomial.h:44-229)
  Polynomial();
  explicit Polynomial(ZZ p);
  Polynomial(ZZ p, Polynomial P);
  explicit Polynomial(int spacing);
  Polynomial(ZZ p, Polynomial P, int spacing);
  Polynomial(ZZ p, int spacing);
  Polynomial(const Polynomial& b) = default;
  Polynomial& operator=(const Polynomial& b) = default;
  void copy(const Polynomial& b) { *this = b; }

  // ---- arithmetic (src/polynomial.h:289-696)
  Polynomial operator+(const Polynomial& b) const;
  Polynomial& operator+=(const Polynomial& b);
  Polynomial operator-(const Polynomial& b) const;
  Polynomial& operator-=(const Polynomial& b);
  // Plain product in Z[x]/(x^N - 1), N = max(min_ntt_length,
  // 2^ceil(log2(deg a + deg b))) (src/common.h:136-226), on the GPU.
  Polynomial operator*(const Polynomial& b) const;
  Polynomial& operator*=(const Polynomial& b);
  Polynomial operator+(const ZZ& b) const;
  Polynomial& operator+=(const ZZ& b);
  Polynomial operator-(const ZZ& b) const;
  Polynomial& operator-=(const ZZ& b);
  Polynomial operator*(const ZZ& b) const;
  Polynomial& operator*=(const ZZ& b);
  Polynomial operator%(const ZZ& b) const;
  Polynomial& operator%=(const ZZ& b);
  Polynomial operator/(const ZZ& b) const;
  Polynomial& operator/=(const ZZ& b);
  Polynomial operator+(uint64_t b) const { return *this + ZZ(b); }
  Polynomial operator-(uint64_t b) const { return *this - ZZ(b); }
  Polynomial operator*(uint64_t b) const { return *this * ZZ(b); }
  Polynomial operator%(uint64_t b) const { return *this % ZZ(b); }
  Polynomial& operator+=(uint64_t b) { return *this += ZZ(b); }
  Polynomial& operator-=(uint64_t b) { return *this -= ZZ(b); }
  Polynomial& operator*=(uint64_t b) { return *this *= ZZ(b); }
  Polynomial& operator%=(uint64_t b) { return *this %= ZZ(b); }
  Polynomial operator/(const Polynomial& b) const;  // DivRem by x^n + 1 forms only
  Polynomial operator%(const Polynomial& b) const;
  bool operator==(const Polynomial& b) const;
  bool operator!=(const Polynomial& b) const { return !(*this == b); }

  // ---- host helpers (src/polynomial.h:698-730)
  void MulMod(const ZZ& b, const ZZ& mod);
  void CPUAddition(const Polynomial* b);
  void CPUMulAdd(const Polynomial* b, const ZZ& M);
  void CPUSubtraction(const Polynomial* b);

  // ---- coefficients (src/polynomial.h:731-900)
  void normalize();
  ZZ get_mod() const { return mod_; }
  void set_mod(const ZZ& v) { mod_ = v; }
  Polynomial get_phi() const;
  void set_phi(const Polynomial& p);
  ZZ get_coeff(int index) const;
  void set_coeff(int index, const ZZ& value);
  void set_coeff(int index, int value) { set_coeff(index, ZZ((int64_t)value)); }
  std::vector<ZZ> get_coeffs() const { return coefs_; }
  void set_coeffs(const std::vector<uint64_t>& values);
  void set_coeffs(const std::vector<ZZ>& values);
  void set_coeffs();
  void set_coeffs(int size);
  int deg() const;
  ZZ lead_coeff() const { return deg() >= 0 ? coefs_[deg()] : ZZ(0); }
  bool is_zero() const { return deg() < 0; }
  bool is_one() const { return deg() == 0 && coefs_[0] == ZZ(1); }
  std::string to_string() const;

  // ---- residency (src/polynomial.h:959-1042, src/polynomial.cpp:252-417)
  // Device copy of the coefficients (limb-major, q's limb count), made on
  // demand; flags as in the reference.
  void update_device_data();
  void update_host_data() {}
  bool get_host_updated() const { return true; }
  bool get_crt_computed() const { return device_valid_; }
  bool get_icrt_computed() const { return true; }
  void set_crt_computed(bool b) { if (!b) device_valid_ = false; }
  void set_icrt_computed(bool) {}
  void set_host_updated(bool) {}
  int get_crt_spacing() const;
  void update_crt_spacing(int) {}
  // CRT residues [R][spacing] on the device (Polynomial::crt +
  // get_device_crt_residues); owned by this object.
  uint64_t* get_device_crt_residues();
  const uint64_t* device_coefs() const { return device_valid_ ? dev_->get() : nullptr; }
  void crt() { get_device_crt_residues(); }
  void icrt() {}
  // Reduction mod (phi = x^n + 1, global_mod) on the GPU
  // (Polynomial::reduce, src/cuda_functions.cu:1040-1078, which runs on the
  // host in the reference).
  void reduce();
  void release() { dev_.reset(); res_.reset(); device_valid_ = false; }
  void* get_stream() const { return nullptr; }

  // ---- statics (src/polynomial.h:898-1380)
  static void gen_crt_primes(ZZ q, uint64_t degree);
  static void DivRem(Polynomial a, Polynomial b, Polynomial& quot, Polynomial& rem);
  static Polynomial InvMod(Polynomial a, Polynomial b);
  static void BuildNthCyclotomic(Polynomial* phi, unsigned int n);
  static void random(Polynomial* a, int degree);
  static bool check_special_rem_format(const Polynomial* p);

  // limb-major device image helpers for the Ciphertext / Yashe layers
  static int ring_degree();
  void from_device(const uint64_t* limbs, int L, int n);
  // Takes a device result [L][n]: downloads it into the host coefficients and
  // keeps the buffer as the valid device copy.
  void adopt_device(std::shared_ptr<cuyashe::DeviceBuffer> buf, int L, int n);
  // Device copy as [L][n] with q's limb count and the ring degree.
  const uint64_t* device_ring_image();
  // The device copy when it is valid with shape [L][n], else null.
  std::shared_ptr<cuyashe::DeviceBuffer> device_buffer(int L, int n) const {
    return (device_valid_ && dev_L_ == L && dev_n_ == n) ? dev_ : nullptr;
  }

 protected:
  std::vector<ZZ> coefs_;
  ZZ mod_;
  const Polynomial* phi_ = nullptr;
  std::shared_ptr<cuyashe::DeviceBuffer> dev_;
  int dev_L_ = 0, dev_n_ = 0;
  bool device_valid_ = false;
  std::shared_ptr<cuyashe::DeviceBuffer> res_;
  int res_spacing_ = 0;
  void touched() { device_valid_ = false; res_.reset(); }
  // scalar operator on the device when this is a ring element (see the
  // operators); false when it is not and the host path applies
  bool device_scalar(const ZZ& b, int op);
  friend class Ciphertext;
  friend class Yashe;
};
