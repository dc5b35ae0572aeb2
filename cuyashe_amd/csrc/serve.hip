// Continuous-batching request loop (yh_server_*): homomorphic additions,
// multiply+keyswitches and encrypt-then-decrypt round trips on resident keys
// and ciphertexts of several rings.
//
// One host thread runs the loop: it admits arrived requests in arrival order
// while the HBM reserved for work in flight stays within a budget and the
// ring has a free result slot, groups the admitted requests by (ring,
// operation), launches each group as ONE batched operation over device
// pointer lists on that ring's own stream (groups of different rings overlap
// on the device), decrypts every result on the device and copies the
// decryptions back, and retires groups whose completion event has fired.
// Encryption noise (ternary s, discrete Gaussian e) is drawn on the device
// from a counter-based generator keyed by (seed, request id, coefficient), so
// a request's ciphertext does not depend on the batch it ran in.
//
// The reference has no request loop: it runs one Ciphertext operator at a
// time (src/ciphertext.cpp:38-107, Yashe::encrypt / decrypt at
// src/yashe.cpp:104-197), each followed by a device synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/yashe_hip.h"
#include "bigint.hpp"
#include "internal.hpp"

namespace {

using yhi::fail;

constexpr int GAUSS_BOUND = 48;  // D_{Z, 8} restricted to [-48, 48]
constexpr double GAUSS_SIGMA = 8.0;
constexpr int GAUSS_VALUES = 2 * GAUSS_BOUND + 1;
constexpr int MAX_LIMBS = 8;
constexpr int LISTS = 6;  // pointer lists per group (see launch_group)

struct NoiseArgs {
  uint64_t q[MAX_LIMBS];
  uint64_t seed;
  int L;
  int n;
  // thresholds floor(F(k) 2^64) of the cumulative distribution F at the
  // values -48 .. 47; the sample is (number of thresholds <= u) - 48
  uint64_t cdf[GAUSS_VALUES - 1];
};

// SplitMix64 finaliser (Steele, Lea and Flood, OOPSLA 2014): a bijective
// 64-bit mix whose outputs on consecutive inputs pass BigCrush.
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Uniform word `which` of coefficient j of request rid.
__device__ __forceinline__ uint64_t noise_word(uint64_t seed, uint64_t rid, int j, int which) {
  return mix64(mix64(seed ^ mix64(rid)) + 2 * (uint64_t)j + (uint64_t)which);
}

// v in (-q, q) as L limbs of v mod q, limb stride `stride`.
__device__ __forceinline__ void put_signed(uint64_t* out, size_t stride, int v, const NoiseArgs& na) {
  if (v >= 0) {
    out[0] = (uint64_t)v;
    for (int l = 1; l < na.L; ++l) out[(size_t)l * stride] = 0;
    return;
  }
  uint64_t borrow = (uint64_t)(-v);
  for (int l = 0; l < na.L; ++l) {
    const uint64_t d = na.q[l] - borrow;
    borrow = na.q[l] < borrow ? 1 : 0;
    out[(size_t)l * stride] = d;
  }
}

// Inputs of the encryptions of one group (Yashe::encrypt,
// src/yashe.cpp:104-142, c = h s + e + floor(q/t) m): for message m of the
// group and coefficient j, the ternary s_j, the Gaussian e_j and m_j itself,
// each as an [L][n] ring element. grid = (ceil(n / 256), messages).
__global__ void k_encrypt_inputs(NoiseArgs na, const uint64_t* rids, const uint64_t* msg, uint64_t* S, uint64_t* E,
                                 uint64_t* DM) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int m = blockIdx.y;
  if (j >= na.n) return;
  const uint64_t rid = rids[m];
  const int s = (int)(noise_word(na.seed, rid, j, 0) % 3) - 1;
  const uint64_t u = noise_word(na.seed, rid, j, 1);
  int lo = 0, hi = GAUSS_VALUES - 1;  // count of thresholds <= u, in [0, 96]
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (na.cdf[mid] <= u) lo = mid + 1;
    else hi = mid;
  }
  const int e = lo - GAUSS_BOUND;
  const size_t base = (size_t)m * na.L * na.n + j;
  put_signed(S + base, na.n, s, na);
  put_signed(E + base, na.n, e, na);
  DM[base] = msg[(size_t)m * na.n + j];
  for (int l = 1; l < na.L; ++l) DM[base + (size_t)l * na.n] = 0;
}

#define SV_HIP(expr)                                                                     \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess) {                                                              \
      (void)hipGetLastError();                                                           \
      return fail(YH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    }                                                                                    \
  } while (0)

#define SV_TRY(expr)                \
  do {                              \
    if (int s_ = (expr)) return s_; \
  } while (0)

// Device and pinned host buffers of one group in flight, reused by later
// groups of the same (ring, operation) once it has retired.
struct Stage {
  uint64_t* d_tab = nullptr;  // LISTS pointer lists of max_batch words
  uint64_t* h_tab = nullptr;
  uint64_t* d_M = nullptr;    // decryptions [max_batch][n]
  uint64_t* h_M = nullptr;
  uint64_t* d_S = nullptr;    // encryption inputs and h s, [max_batch][L][n] each
  uint64_t* d_E = nullptr;
  uint64_t* d_DM = nullptr;
  uint64_t* d_HS = nullptr;
  uint64_t* d_msg = nullptr;  // messages [max_batch][n]
  uint64_t* h_msg = nullptr;
  hipEvent_t ev_op = nullptr;    // the operation's result is complete
  hipEvent_t ev_done = nullptr;  // the decryptions are on the host
  ~Stage() {
    for (uint64_t* d : {d_tab, d_M, d_S, d_E, d_DM, d_HS, d_msg})
      if (d) (void)hipFree(d);
    for (uint64_t* h : {h_tab, h_M, h_msg})
      if (h) (void)hipHostFree(h);
    if (ev_op) (void)hipEventDestroy(ev_op);
    if (ev_done) (void)hipEventDestroy(ev_done);
  }
};

struct RingState {
  yh_serve_ring cfg{};
  yh_ring_info info{};
  hipStream_t stream = nullptr;
  std::vector<uint64_t> delta;  // floor(q / t)
  NoiseArgs noise{};
  size_t elt = 0;               // bytes of one ring element
  uint64_t* d_out = nullptr;    // result slots [slots][L][n]
  std::deque<int> free;
  size_t fp[3] = {};            // admission footprint per request kind
  std::vector<Stage*> idle[3];  // stages free for reuse, per kind
  std::vector<std::unique_ptr<Stage>> owned;
  RingState() = default;
  RingState(const RingState&) = delete;
  RingState(RingState&&) = default;
};

struct Group {
  int ring = 0, kind = 0;
  std::vector<int> reqs;
  std::vector<int> slots;
  Stage* st = nullptr;
  size_t reserved = 0;
};

}  // namespace

struct yh_server {
  yh_ctx* ctx = nullptr;
  int device = 0;
  std::vector<RingState> rings;
  size_t budget = 0;
  int max_batch = 0;
  int depth = 0;  // groups in flight per (ring, kind)
  uint64_t seed = 0;
  hipEvent_t start = nullptr;
};

namespace {

int stage_for(yh_server* s, RingState& rs, int kind, Stage** out) {
  if (!rs.idle[kind].empty()) {
    *out = rs.idle[kind].back();
    rs.idle[kind].pop_back();
    return YH_OK;
  }
  auto st = std::make_unique<Stage>();
  const size_t mb = s->max_batch, n = rs.info.n;
  SV_HIP(hipMalloc(&st->d_tab, LISTS * mb * sizeof(uint64_t)));
  SV_HIP(hipHostMalloc(&st->h_tab, LISTS * mb * sizeof(uint64_t), hipHostMallocDefault));
  SV_HIP(hipMalloc(&st->d_M, mb * n * sizeof(uint64_t)));
  SV_HIP(hipHostMalloc(&st->h_M, mb * n * sizeof(uint64_t), hipHostMallocDefault));
  if (kind == YH_REQ_ENCDEC) {
    for (uint64_t** d : {&st->d_S, &st->d_E, &st->d_DM, &st->d_HS}) SV_HIP(hipMalloc(d, mb * rs.elt));
    SV_HIP(hipMalloc(&st->d_msg, mb * n * sizeof(uint64_t)));
    SV_HIP(hipHostMalloc(&st->h_msg, mb * n * sizeof(uint64_t), hipHostMallocDefault));
  }
  SV_HIP(hipEventCreate(&st->ev_op));
  SV_HIP(hipEventCreate(&st->ev_done));
  *out = st.get();
  rs.owned.push_back(std::move(st));
  return YH_OK;
}

size_t footprint(const yh_server* s, const yh_request& r) {
  const RingState& rs = s->rings[r.ring];
  return rs.fp[r.kind] + rs.elt + (size_t)rs.info.n * sizeof(uint64_t);
}

// Enqueues one group on its ring's stream: the operation into the result
// slots, then the decryption of every result and its copy to the host.
int launch_group(yh_server* s, yh_request* reqs, Group& g) {
  RingState& rs = s->rings[g.ring];
  yh_ring* ring = rs.cfg.ring;
  const int k = (int)g.reqs.size();
  const size_t mb = s->max_batch, n = rs.info.n, ew = rs.elt / sizeof(uint64_t);
  SV_TRY(stage_for(s, rs, g.kind, &g.st));
  Stage& st = *g.st;
  uint64_t* h = st.h_tab;
  using PL = uint64_t* const*;
  using CPL = const uint64_t* const*;
  PL d0 = (PL)st.d_tab, d1 = (PL)(st.d_tab + mb), d2 = (PL)(st.d_tab + 2 * mb), d3 = (PL)(st.d_tab + 3 * mb),
     d4 = (PL)(st.d_tab + 4 * mb);
  for (int m = 0; m < k; ++m) {
    h[2 * mb + m] = (uint64_t)(rs.d_out + (size_t)g.slots[m] * ew);  // results
    h[4 * mb + m] = (uint64_t)(st.d_M + (size_t)m * n);             // decryptions
  }
  if (g.kind == YH_REQ_ENCDEC) {
    for (int m = 0; m < k; ++m) {
      const yh_request& r = reqs[g.reqs[m]];
      h[m] = (uint64_t)(st.d_HS + (size_t)m * ew);
      h[mb + m] = (uint64_t)(st.d_E + (size_t)m * ew);
      h[3 * mb + m] = (uint64_t)(st.d_DM + (size_t)m * ew);
      h[5 * mb + m] = r.rid;
      std::memcpy(st.h_msg + (size_t)m * n, r.msg, n * sizeof(uint64_t));
    }
    SV_HIP(hipMemcpyAsync(st.d_tab, h, LISTS * mb * sizeof(uint64_t), hipMemcpyHostToDevice, rs.stream));
    SV_HIP(hipMemcpyAsync(st.d_msg, st.h_msg, k * n * sizeof(uint64_t), hipMemcpyHostToDevice, rs.stream));
    NoiseArgs na = rs.noise;
    na.seed = s->seed;
    hipLaunchKernelGGL(k_encrypt_inputs, dim3((unsigned)((n + 255) / 256), k), dim3(256), 0, rs.stream, na,
                       (const uint64_t*)(st.d_tab + 5 * mb), (const uint64_t*)st.d_msg, st.d_S, st.d_E, st.d_DM);
    SV_HIP(hipGetLastError());
    // c = h s + e + floor(q/t) m, written into the result slots
    SV_TRY(yh_ring_scalar(ring, st.d_DM, rs.delta.data(), (int)rs.delta.size(), st.d_DM, k, YH_MUL));
    SV_TRY(yh_ring_mul_bcast(ring, rs.cfg.d_h, st.d_S, st.d_HS, k));
    SV_TRY(yh_ring_addsub_list(ring, (CPL)d0, (CPL)d1, d2, k, YH_ADD));
    SV_TRY(yh_ring_addsub_list(ring, (CPL)d2, (CPL)d3, d2, k, YH_ADD));
  } else {
    for (int m = 0; m < k; ++m) {
      const yh_request& r = reqs[g.reqs[m]];
      h[m] = (uint64_t)(rs.cfg.d_pool + (size_t)r.a * ew);
      h[mb + m] = (uint64_t)(rs.cfg.d_pool + (size_t)r.b * ew);
    }
    SV_HIP(hipMemcpyAsync(st.d_tab, h, LISTS * mb * sizeof(uint64_t), hipMemcpyHostToDevice, rs.stream));
    if (g.kind == YH_REQ_ADD) SV_TRY(yh_ring_addsub_list(ring, (CPL)d0, (CPL)d1, d2, k, YH_ADD));
    else SV_TRY(yh_yashe_mul_keyswitch_list(ring, rs.cfg.evk, (CPL)d0, (CPL)d1, d2, k));
  }
  SV_HIP(hipEventRecord(st.ev_op, rs.stream));
  SV_TRY(yh_yashe_decrypt_list(ring, rs.cfg.d_f, (CPL)d2, d4, k));
  SV_HIP(hipMemcpyAsync(st.h_M, st.d_M, k * n * sizeof(uint64_t), hipMemcpyDeviceToHost, rs.stream));
  SV_HIP(hipEventRecord(st.ev_done, rs.stream));
  return YH_OK;
}

int retire_group(yh_server* s, yh_request* reqs, Group& g) {
  RingState& rs = s->rings[g.ring];
  Stage& st = *g.st;
  float ms = 0;
  SV_HIP(hipEventElapsedTime(&ms, s->start, st.ev_op));
  const size_t n = rs.info.n, ew = rs.elt / sizeof(uint64_t);
  for (size_t m = 0; m < g.reqs.size(); ++m) {
    yh_request& r = reqs[g.reqs[m]];
    r.done = ms * 1e-3;
    std::memcpy(r.result, st.h_M + m * n, n * sizeof(uint64_t));
    if (r.output) {  // the result ciphertext itself, before its slot is reused
      SV_HIP(hipMemcpyAsync(r.output, rs.d_out + (size_t)g.slots[m] * ew, rs.elt, hipMemcpyDeviceToHost, rs.stream));
      SV_HIP(hipStreamSynchronize(rs.stream));
    }
  }
  for (int sl : g.slots) rs.free.push_back(sl);
  rs.idle[g.kind].push_back(g.st);
  g.st = nullptr;
  return YH_OK;
}

int validate(const yh_server* s, const yh_request* reqs, int count) {
  for (int i = 0; i < count; ++i) {
    const yh_request& r = reqs[i];
    const std::string at = "request " + std::to_string(i) + ": ";
    if (r.kind < 0 || r.kind > 2) return fail(YH_ERR_ARG, at + "unknown kind");
    if (r.ring < 0 || r.ring >= (int)s->rings.size()) return fail(YH_ERR_ARG, at + "unknown ring");
    const RingState& rs = s->rings[r.ring];
    if (!r.result) return fail(YH_ERR_ARG, at + "null result buffer");
    if (!(r.arrival >= 0) || !std::isfinite(r.arrival)) return fail(YH_ERR_ARG, at + "bad arrival time");
    if (r.kind == YH_REQ_ENCDEC) {
      if (!r.msg) return fail(YH_ERR_ARG, at + "null message");
      for (int j = 0; j < rs.info.n; ++j)
        if (r.msg[j] >= rs.info.t) return fail(YH_ERR_ARG, at + "message coefficient not below t");
    } else {
      if (r.a < 0 || r.b < 0 || r.a >= rs.cfg.pool || r.b >= rs.cfg.pool)
        return fail(YH_ERR_ARG, at + "pool index out of range");
      if (r.kind == YH_REQ_MULKS && !rs.cfg.evk) return fail(YH_ERR_ARG, at + "ring has no keyswitch key");
    }
    if (footprint(s, r) > s->budget) return fail(YH_ERR_ARG, at + "needs more HBM than the whole budget");
  }
  return YH_OK;
}

// Waits for every group still in flight (error paths).
void drain(yh_server* s) {
  for (RingState& rs : s->rings) (void)hipStreamSynchronize(rs.stream);
}

}  // namespace

extern "C" {

int yh_server_create(yh_ctx* ctx, const yh_serve_ring* rings, int nrings, size_t hbm_budget, int max_batch,
                     int depth, uint64_t seed, yh_server** out) {
  if (!ctx || !rings || nrings <= 0 || !out) return fail(YH_ERR_ARG, "null or empty argument");
  if (max_batch < 1 || max_batch > 4096) return fail(YH_ERR_ARG, "max_batch must be in [1, 4096]");
  if (depth < 1 || depth > 64) return fail(YH_ERR_ARG, "depth must be in [1, 64]");
  for (int k = 0; k < nrings; ++k) {
    const yh_serve_ring& c = rings[k];
    if (!c.ring || !c.d_f || !c.d_h || !c.d_pool || c.pool < 1 || c.slots < 1)
      return fail(YH_ERR_ARG, "ring " + std::to_string(k) + ": null key or pool, or empty pool or slots");
    if (yhi::ring_ctx(c.ring) != ctx) return fail(YH_ERR_ARG, "ring from another context");
    if (c.evk && yhi::evk_ring(c.evk) != c.ring) return fail(YH_ERR_ARG, "keyswitch key of another ring");
  }
  auto s = std::make_unique<yh_server>();
  s->ctx = ctx;
  s->device = yhi::ctx_device(ctx);
  s->budget = hbm_budget;
  s->max_batch = max_batch;
  s->depth = depth;
  s->seed = seed;
  SV_HIP(hipSetDevice(s->device));
  // the thresholds of D_{Z, 8} on [-48, 48]
  std::vector<double> w(GAUSS_VALUES);
  for (int i = 0; i < GAUSS_VALUES; ++i) {
    const double x = i - GAUSS_BOUND;
    w[i] = std::exp(-x * x / (2.0 * GAUSS_SIGMA * GAUSS_SIGMA));
  }
  const double total = std::accumulate(w.begin(), w.end(), 0.0);
  uint64_t cdf[GAUSS_VALUES - 1];
  double acc = 0;
  for (int i = 0; i + 1 < GAUSS_VALUES; ++i) {
    acc += w[i];
    const double v = std::ldexp(acc / total, 64);
    cdf[i] = v >= 18446744073709551615.0 ? ~0ull : (uint64_t)v;
  }
  s->rings.resize(nrings);
  for (int k = 0; k < nrings; ++k) {
    RingState& rs = s->rings[k];
    rs.cfg = rings[k];
    SV_TRY(yh_ring_get_info(rs.cfg.ring, &rs.info));
    if (rs.info.t < 2) return fail(YH_ERR_ARG, "ring " + std::to_string(k) + " has no plaintext modulus t");
    SV_TRY(yh_ring_use_own_stream(rs.cfg.ring));
    rs.stream = yhi::ring_stream(rs.cfg.ring);
    int L = 0;
    const uint64_t* q = yhi::ring_q(rs.cfg.ring, &L);
    const yh::BigInt Q = yh::BigInt::from_limbs(q, L);
    const yh::BigInt D = Q / yh::BigInt(rs.info.t);
    rs.delta.assign(L, 0);
    D.to_limbs(rs.delta.data(), L);
    rs.noise = NoiseArgs{};
    for (int l = 0; l < L; ++l) rs.noise.q[l] = q[l];
    rs.noise.L = L;
    rs.noise.n = rs.info.n;
    std::memcpy(rs.noise.cdf, cdf, sizeof(cdf));
    rs.elt = (size_t)L * rs.info.n * sizeof(uint64_t);
    rs.fp[YH_REQ_ADD] = yh_mem_footprint(rs.cfg.ring, 4);
    rs.fp[YH_REQ_MULKS] = yh_mem_footprint(rs.cfg.ring, 3);
    rs.fp[YH_REQ_ENCDEC] = 3 * rs.elt;
    SV_HIP(hipMalloc(&rs.d_out, (size_t)rs.cfg.slots * rs.elt));
    for (int sl = 0; sl < rs.cfg.slots; ++sl) rs.free.push_back(sl);
  }
  SV_HIP(hipEventCreate(&s->start));
  for (RingState& rs : s->rings) {
    yhi::hold_ring(rs.cfg.ring, 1);
    if (rs.cfg.evk) yhi::hold_evk(rs.cfg.evk, 1);
  }
  yhi::hold_ctx(ctx, 1);
  *out = s.release();
  return YH_OK;
}

int yh_server_set_budget(yh_server* s, size_t hbm_budget) {
  if (!s) return fail(YH_ERR_ARG, "null server");
  s->budget = hbm_budget;
  return YH_OK;
}

size_t yh_server_footprint(const yh_server* s, int ring, int kind) {
  if (!s || ring < 0 || ring >= (int)s->rings.size() || kind < 0 || kind > 2) return 0;
  yh_request r{};
  r.ring = ring;
  r.kind = kind;
  return footprint(s, r);
}

int yh_server_run(yh_server* s, yh_request* reqs, int count, yh_serve_stats* stats) {
  if (!s || (count > 0 && !reqs)) return fail(YH_ERR_ARG, "null argument");
  SV_TRY(validate(s, reqs, count));
  SV_HIP(hipSetDevice(s->device));
  std::vector<int> order(count);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return reqs[x].arrival < reqs[y].arrival; });
  std::vector<int> rank(count);  // position in arrival order
  for (int i = 0; i < count; ++i) rank[order[i]] = i;
  for (RingState& rs : s->rings) SV_HIP(hipStreamSynchronize(rs.stream));
  hipStream_t cs = yhi::ctx_stream(s->ctx);
  SV_HIP(hipStreamSynchronize(cs));
  SV_HIP(hipEventRecord(s->start, cs));
  SV_HIP(hipEventSynchronize(s->start));
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto now_s = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };
  // one queue of arrived, not yet admitted requests per (ring, kind)
  const int nkeys = 3 * (int)s->rings.size();
  std::vector<std::deque<int>> queue(nkeys);
  std::vector<int> key_inflight(nkeys, 0);
  std::vector<Group> inflight;
  std::vector<int> cand;
  size_t used = 0, peak = 0, queued = 0;
  long groups = 0, batched = 0;
  int max_inflight = 0;
  int next = 0;
  int rc = YH_OK;
  while (next < count || queued > 0 || !inflight.empty()) {
    const double now = now_s();
    while (next < count && reqs[order[next]].arrival <= now) {
      const yh_request& r = reqs[order[next]];
      queue[3 * r.ring + r.kind].push_back(order[next++]);
      ++queued;
    }
    // retire finished groups
    for (size_t k = 0; k < inflight.size();) {
      const hipError_t q = hipEventQuery(inflight[k].st->ev_done);
      if (q == hipErrorNotReady) {
        ++k;
        continue;
      }
      if (q != hipSuccess) {
        rc = fail(YH_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
        break;
      }
      if ((rc = retire_group(s, reqs, inflight[k]))) break;
      used -= inflight[k].reserved;
      key_inflight[3 * inflight[k].ring + inflight[k].kind]--;
      inflight[k] = std::move(inflight.back());
      inflight.pop_back();
    }
    if (rc) break;
    // admission: keys below their in-flight depth, oldest waiting request
    // first; each forms one group of up to max_batch requests from the front
    // of its queue while the budget and the ring's free slots allow. A
    // request that does not fit the budget stops admission (no later
    // arrival overtakes it for memory).
    cand.clear();
    for (int key = 0; key < nkeys; ++key)
      if (!queue[key].empty() && key_inflight[key] < s->depth) cand.push_back(key);
    std::sort(cand.begin(), cand.end(),
              [&](int x, int y) { return rank[queue[x].front()] < rank[queue[y].front()]; });
    bool launched = false, blocked = false;
    for (int key : cand) {
      RingState& rs = s->rings[key / 3];
      while (!queue[key].empty() && key_inflight[key] < s->depth && !blocked) {
        Group g;
        g.ring = key / 3;
        g.kind = key % 3;
        while (!queue[key].empty() && (int)g.reqs.size() < s->max_batch && !rs.free.empty()) {
          const size_t fp = footprint(s, reqs[queue[key].front()]);
          if (used + fp > s->budget) {
            blocked = true;
            break;
          }
          g.reqs.push_back(queue[key].front());
          g.slots.push_back(rs.free.front());
          g.reserved += fp;
          used += fp;
          queue[key].pop_front();
          rs.free.pop_front();
          --queued;
        }
        if (g.reqs.empty()) break;  // no budget or no free slot on this ring
        if ((rc = launch_group(s, reqs, g))) {
          for (int sl : g.slots) rs.free.push_back(sl);
          if (g.st) rs.idle[g.kind].push_back(g.st);
          used -= g.reserved;
          break;
        }
        ++groups;
        batched += (long)g.reqs.size();
        key_inflight[key]++;
        inflight.push_back(std::move(g));
        launched = true;
      }
      if (rc || blocked) break;
    }
    if (rc) break;
    peak = std::max(peak, used);
    max_inflight = std::max(max_inflight, (int)inflight.size());
    if (!launched) {
      if (inflight.empty() && queued == 0 && next < count) {
        // idle until shortly before the next arrival
        const double wait = reqs[order[next]].arrival - now_s() - 200e-6;
        if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      } else {
        std::this_thread::yield();
      }
    }
  }
  const double wall = now_s();
  if (rc) {
    drain(s);
    for (Group& g : inflight) {
      RingState& rs = s->rings[g.ring];
      for (int sl : g.slots) rs.free.push_back(sl);
      if (g.st) rs.idle[g.kind].push_back(g.st);
    }
    return rc;
  }
  if (stats) {
    stats->wall_s = wall;
    stats->groups = groups;
    stats->mean_batch = groups ? (double)batched / groups : 0.0;
    stats->peak_reserved = peak;
    stats->max_inflight = max_inflight;
  }
  return YH_OK;
}

int yh_server_destroy(yh_server* s) {
  if (!s) return YH_OK;
  (void)hipSetDevice(s->device);
  drain(s);
  for (RingState& rs : s->rings) {
    rs.owned.clear();
    if (rs.d_out) (void)hipFree(rs.d_out);
    yhi::hold_ring(rs.cfg.ring, -1);
    if (rs.cfg.evk) yhi::hold_evk(rs.cfg.evk, -1);
  }
  if (s->start) (void)hipEventDestroy(s->start);
  yhi::hold_ctx(s->ctx, -1);
  delete s;
  return YH_OK;
}

}  // extern "C"
