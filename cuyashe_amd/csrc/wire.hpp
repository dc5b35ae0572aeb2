// Byte format for ring elements (ciphertexts, keys, messages) leaving or
// entering the device: what a client of a serving process sends, and what a
// process writes to keep keys across runs. The reference has no such format
// (its only serialisation is the Python 2 to_JSON of
// experiments/modq/polynomial.py:158-165, decimal coefficients in JSON), so
// this one is designed for the library's own layout:
//
//   offset  size        field
//   0       4           magic "YHW1"
//   4       2           version (1)
//   6       2           kind (YH_WIRE_* in yashe_hip.h)
//   8       4           n (ring degree, power of two)
//   12      4           L (64-bit limbs of q)
//   16      8           t (plaintext modulus, 0 if none)
//   24      4           w_bits (keyswitch digit width, 0 if none)
//   28      4           count (polynomials in the payload)
//   32      8 L         q, limb 0 least significant
//   32+8L   8 L n count payload: polynomial-major, then limb-major
//                       ([count][L][n], the device layout), each
//                       coefficient in [0, q) ([0, t) for plaintexts)
//   end-4   4           CRC-32 (IEEE 802.3, as zlib.crc32) of every byte
//                       before it
// All integers little-endian. Decoding checks every header field against the
// ring the caller names, the exact length and the checksum before it writes
// the destination, then every coefficient's range (a value out of range
// clears the destination and fails).
#pragma once
#include <stdint.h>
#include <string.h>

#include <string>

namespace yhw {

constexpr uint32_t MAGIC = 0x31574859u;  // "YHW1" little-endian
constexpr int VERSION = 1;
constexpr int MAX_LIMBS = 8;
constexpr size_t HEADER = 32;

struct Crc32Table {
  uint32_t v[256];
  Crc32Table() {
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      v[i] = c;
    }
  }
};

inline uint32_t crc32(const uint8_t* p, size_t len) {
  static const Crc32Table table;  // thread-safe one-time initialisation
  uint32_t c = 0xFFFFFFFFu;
  for (size_t i = 0; i < len; ++i) c = table.v[(c ^ p[i]) & 0xFF] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

inline void put_u16(uint8_t* p, uint16_t v) { p[0] = (uint8_t)v; p[1] = (uint8_t)(v >> 8); }
inline void put_u32(uint8_t* p, uint32_t v) { for (int i = 0; i < 4; ++i) p[i] = (uint8_t)(v >> (8 * i)); }
inline void put_u64(uint8_t* p, uint64_t v) { for (int i = 0; i < 8; ++i) p[i] = (uint8_t)(v >> (8 * i)); }
inline uint16_t get_u16(const uint8_t* p) { return (uint16_t)(p[0] | (p[1] << 8)); }
inline uint32_t get_u32(const uint8_t* p) {
  uint32_t v = 0;
  for (int i = 3; i >= 0; --i) v = (v << 8) | p[i];
  return v;
}
inline uint64_t get_u64(const uint8_t* p) {
  uint64_t v = 0;
  for (int i = 7; i >= 0; --i) v = (v << 8) | p[i];
  return v;
}

struct Params {
  int n = 0;
  int L = 0;
  uint64_t q[MAX_LIMBS] = {};
  uint64_t t = 0;
  int w_bits = 0;
};

struct Header {
  int version = 0, kind = 0, count = 0;
  Params p;
};

inline size_t encoded_size(int n, int L, int count) {
  return HEADER + 8 * (size_t)L + 8 * (size_t)L * (size_t)n * (size_t)count + 4;
}

// a < b over L limbs
inline bool less_limbs(const uint64_t* a, const uint64_t* b, int L) {
  for (int l = L - 1; l >= 0; --l)
    if (a[l] != b[l]) return a[l] < b[l];
  return false;
}

// Checks the ring parameters themselves; "" when valid.
inline std::string check_params(const Params& p) {
  if (p.n < 2 || p.n > 16384 || (p.n & (p.n - 1))) return "n must be a power of two in [2, 16384]";
  if (p.L < 1 || p.L > MAX_LIMBS) return "q must have 1 to 8 limbs";
  if (p.q[p.L - 1] == 0) return "top limb of q is zero";
  if (p.L == 1 && p.q[0] < 2) return "q must be > 1";
  if (p.w_bits < 0 || p.w_bits > 32) return "w_bits out of range";
  return "";
}

// Coefficient j of limb-major polynomial data (L limbs, n coefficients) is
// below the bound (q, or t for plaintexts).
inline std::string check_range(const uint64_t* poly, const Params& p, bool plain) {
  for (int j = 0; j < p.n; ++j) {
    uint64_t c[MAX_LIMBS];
    for (int l = 0; l < p.L; ++l) c[l] = poly[(size_t)l * p.n + j];
    if (plain && p.t) {
      for (int l = 1; l < p.L; ++l)
        if (c[l]) return "plaintext coefficient not below t";
      if (c[0] >= p.t) return "plaintext coefficient not below t";
    } else if (!less_limbs(c, p.q, p.L)) {
      return "coefficient not below q";
    }
  }
  return "";
}

inline std::string encode(const Params& p, int kind, bool plain, const uint64_t* src, int count, uint8_t* out,
                          size_t cap, size_t* len) {
  std::string e = check_params(p);
  if (!e.empty()) return e;
  if (count < 0 || kind < 1 || kind > 0xFFFF) return "bad count or kind";
  const size_t need = encoded_size(p.n, p.L, count);
  if (len) *len = need;
  if (!out) return "";  // size query
  if (cap < need) return "output buffer too small";
  const size_t poly = (size_t)p.L * p.n;
  for (int i = 0; i < count; ++i) {
    e = check_range(src + i * poly, p, plain);
    if (!e.empty()) return e;
  }
  put_u32(out, MAGIC);
  put_u16(out + 4, VERSION);
  put_u16(out + 6, (uint16_t)kind);
  put_u32(out + 8, (uint32_t)p.n);
  put_u32(out + 12, (uint32_t)p.L);
  put_u64(out + 16, p.t);
  put_u32(out + 24, (uint32_t)p.w_bits);
  put_u32(out + 28, (uint32_t)count);
  uint8_t* w = out + HEADER;
  for (int l = 0; l < p.L; ++l, w += 8) put_u64(w, p.q[l]);
  for (size_t k = 0; k < poly * (size_t)count; ++k, w += 8) put_u64(w, src[k]);
  put_u32(w, crc32(out, need - 4));
  return "";
}

inline std::string peek(const uint8_t* in, size_t len, Header* h) {
  if (!in || len < HEADER) return "input shorter than the header";
  if (get_u32(in) != MAGIC) return "not a ring element blob (bad magic)";
  h->version = get_u16(in + 4);
  if (h->version != VERSION) return "unsupported format version";
  h->kind = get_u16(in + 6);
  h->p.n = (int)get_u32(in + 8);
  const uint32_t L = get_u32(in + 12);
  if (L < 1 || L > (uint32_t)MAX_LIMBS) return "q must have 1 to 8 limbs";
  h->p.L = (int)L;
  h->p.t = get_u64(in + 16);
  h->p.w_bits = (int)get_u32(in + 24);
  const uint32_t count = get_u32(in + 28);
  if (count > (1u << 30)) return "count out of range";
  h->count = (int)count;
  if (len < HEADER + 8 * (size_t)L) return "input shorter than the header";
  for (uint32_t l = 0; l < L; ++l) h->p.q[l] = get_u64(in + HEADER + 8 * l);
  return check_params(h->p);
}

// Decodes into dst ([count][L][n]); `expect` names the ring, `kind` the
// expected kind (0: any).
inline std::string decode(const Params& expect, int kind, bool plain, const uint8_t* in, size_t len, uint64_t* dst,
                          int cap, int* count) {
  Header h;
  std::string e = peek(in, len, &h);
  if (!e.empty()) return e;
  if (kind && h.kind != kind) return "blob holds another kind of element";
  if (h.p.n != expect.n || h.p.L != expect.L || h.p.t != expect.t || h.p.w_bits != expect.w_bits)
    return "blob belongs to another ring (n, q limbs, t or w_bits differ)";
  for (int l = 0; l < h.p.L; ++l)
    if (h.p.q[l] != expect.q[l]) return "blob belongs to another ring (q differs)";
  const size_t need = encoded_size(h.p.n, h.p.L, h.count);
  if (len != need) return "length does not match the header";
  if (get_u32(in + need - 4) != crc32(in, need - 4)) return "checksum mismatch";
  if (count) *count = h.count;
  if (!dst) return "";  // count query
  if (h.count > cap) return "destination too small";
  const size_t poly = (size_t)h.p.L * h.p.n;
  const uint8_t* r = in + HEADER + 8 * (size_t)h.p.L;
  for (size_t k = 0; k < poly * (size_t)h.count; ++k) dst[k] = get_u64(r + 8 * k);
  for (int i = 0; i < h.count; ++i) {
    e = check_range(dst + i * poly, h.p, plain);
    if (!e.empty()) {
      memset(dst, 0, poly * (size_t)h.count * sizeof(uint64_t));
      return e;
    }
  }
  return "";
}

}  // namespace yhw
