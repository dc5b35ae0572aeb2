// NTL-style ZZ for the host API.
//
// The reference's host classes take NTL::ZZ everywhere (src/polynomial.h:7,
// src/yashe.h:4). NTL is not part of this build; this header supplies ZZ as
// the project's own bignum (cuyashe_amd/csrc/bigint.hpp) with the free
// functions the reference calls: conv, to_ZZ, power2_ZZ, NumBits, NumBytes,
// LeftShift, RightShift, PowerMod, InvMod, MulMod, AddMod, DivRem, log.
// Division and remainder follow NTL: floor quotient, remainder with the sign
// of the divisor.
#pragma once
#include <stdint.h>

#include <cmath>
#include <iostream>
#include <string>

#include "../csrc/bigint.hpp"

using ZZ = yh::BigInt;

inline ZZ to_ZZ(long v) { return ZZ((int64_t)v); }
inline ZZ to_ZZ(unsigned long v) { return ZZ((uint64_t)v); }
inline ZZ to_ZZ(int v) { return ZZ((int64_t)v); }
inline ZZ to_ZZ(const char* s) { return ZZ::from_string(s); }
inline ZZ to_ZZ(const std::string& s) { return ZZ::from_string(s); }

template <class T>
T conv(const ZZ& a);
template <>
inline ZZ conv<ZZ>(const ZZ& a) { return a; }
template <>
inline uint64_t conv<uint64_t>(const ZZ& a) { return a.to_u64(); }
template <>
inline long conv<long>(const ZZ& a) { return (long)a.to_i64(); }
template <class T>
T conv(const char* s);
template <>
inline ZZ conv<ZZ>(const char* s) { return ZZ::from_string(s); }
template <class T>
T conv(uint64_t v);
template <>
inline ZZ conv<ZZ>(uint64_t v) { return ZZ(v); }

inline ZZ power2_ZZ(long e) { return ZZ::pow2((int)e); }
inline void power2(ZZ& out, long e) { out = ZZ::pow2((int)e); }
inline long NumBits(const ZZ& a) { return a.bits(); }
inline long NumBytes(const ZZ& a) { return (a.bits() + 7) / 8; }
inline ZZ LeftShift(const ZZ& a, long s) { return a << (int)s; }
inline ZZ RightShift(const ZZ& a, long s) { return a >> (int)s; }
inline ZZ abs(const ZZ& a) { return a.abs(); }
inline ZZ PowerMod(const ZZ& b, const ZZ& e, const ZZ& m) { return ZZ::pow_mod(b, e, m); }
inline ZZ PowerMod(const ZZ& b, long e, const ZZ& m) { return ZZ::pow_mod(b, ZZ((int64_t)e), m); }
inline ZZ InvMod(const ZZ& a, const ZZ& m) { return ZZ::inv_mod(a, m); }
inline ZZ MulMod(const ZZ& a, const ZZ& b, const ZZ& m) { return (a * b) % m; }
inline ZZ AddMod(const ZZ& a, const ZZ& b, const ZZ& m) { return (a + b) % m; }
inline void DivRem(ZZ& q, ZZ& r, const ZZ& a, const ZZ& b) { ZZ::divmod_floor(a, b, q, r); }
// natural logarithm (NTL::log), used by callers to size the digit count
inline double log(const ZZ& a) {
  int b = a.bits();
  if (b <= 60) return std::log((double)a.to_u64());
  ZZ t = a >> (b - 60);
  return std::log((double)t.to_u64()) + (b - 60) * std::log(2.0);
}

inline std::ostream& operator<<(std::ostream& os, const ZZ& a) { return os << a.to_string(); }
