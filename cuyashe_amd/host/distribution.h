// Seeded samplers for key generation and encryption.
//
// The reference samples with cuRAND (src/distribution.cpp:4-63,
// src/cuda_distribution.cu:11-35; its "Gaussian" is uniform mod 7 and a
// layout bug fills only part of the buffer). These are host-side samplers
// with the distributions the configurations state: ternary uniform
// {-1, 0, 1} (NARROW), the discrete Gaussian D_{Z, sigma} restricted to
// [-bound, bound] (DISCRETE_GAUSSIAN: P(x) proportional to
// exp(-x^2 / (2 sigma^2)), drawn by inverse transform of its cumulative
// table) and uniform mod q (UNIFORMLY), all driven by a seeded xoshiro256**
// stream so runs are reproducible.
#pragma once
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "zz.h"

enum kind_t { DISCRETE_GAUSSIAN, BINARY, NARROW, UNIFORMLY, KINDS_COUNT };

class Xoshiro256 {
 public:
  explicit Xoshiro256(uint64_t seed) {
    uint64_t z = seed;
    for (auto& v : s_) {  // splitmix64 seeding
      z += 0x9E3779B97F4A7C15ull;
      uint64_t x = z;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      v = x ^ (x >> 31);
    }
  }
  uint64_t next() {
    const uint64_t result = rotl(s_[1] * 5, 7) * 9;
    const uint64_t t = s_[1] << 17;
    s_[2] ^= s_[0];
    s_[3] ^= s_[1];
    s_[1] ^= s_[2];
    s_[0] ^= s_[3];
    s_[2] ^= t;
    s_[3] = rotl(s_[3], 45);
    return result;
  }
  // uniform in [0, m) by rejection
  uint64_t below(uint64_t m) {
    uint64_t lim = UINT64_MAX - UINT64_MAX % m;
    uint64_t x;
    do { x = next(); } while (x >= lim);
    return x % m;
  }
  double uniform01() { return (next() >> 11) * (1.0 / 9007199254740992.0); }

 private:
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t s_[4];
};

class Distribution {
 public:
  Distribution(kind_t kind = UNIFORMLY, double sigma = 8.0, int bound = 48, uint64_t seed = 1234)
      : kind_(kind), sigma_(sigma), bound_(bound), rng_(seed) {}
  // n signed samples (UNIFORMLY: uniform in [0, q) via sample_mod)
  std::vector<int64_t> sample(int n) {
    std::vector<int64_t> out(n);
    for (auto& v : out) {
      switch (kind_) {
        case NARROW: v = (int64_t)rng_.below(3) - 1; break;
        case BINARY: v = (int64_t)rng_.below(2); break;
        case DISCRETE_GAUSSIAN: v = gaussian(); break;
        default: v = (int64_t)rng_.below(1ull << 62); break;
      }
    }
    return out;
  }
  // n samples reduced into [0, q)
  std::vector<ZZ> sample_mod(int n, const ZZ& q) {
    std::vector<ZZ> out(n);
    if (kind_ == UNIFORMLY) {
      int L = q.num_limbs();
      std::vector<uint64_t> w(L + 1);
      for (auto& v : out) {
        for (auto& x : w) x = rng_.next();
        v = ZZ::from_limbs(w.data(), L + 1) % q;
      }
      return out;
    }
    auto s = sample(n);
    for (int i = 0; i < n; ++i) out[i] = ZZ(s[i]) % q;
    return out;
  }

 private:
  int64_t gaussian() {
    if (cdf_.empty()) {
      double acc = 0;
      for (int x = -bound_; x <= bound_; ++x) {
        acc += std::exp(-(double)x * x / (2.0 * sigma_ * sigma_));
        cdf_.push_back(acc);
      }
      for (auto& c : cdf_) c /= acc;
    }
    const double u = rng_.uniform01();
    size_t i = (size_t)(std::upper_bound(cdf_.begin(), cdf_.end(), u) - cdf_.begin());
    if (i >= cdf_.size()) i = cdf_.size() - 1;
    return (int64_t)i - bound_;
  }

  kind_t kind_;
  double sigma_;
  int bound_;
  Xoshiro256 rng_;
  std::vector<double> cdf_;
};
