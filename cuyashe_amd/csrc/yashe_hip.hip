// C-ABI implementation of the YASHE polynomial-ring hot path for gfx950.
// Declarations and reference citations: include/yashe_hip.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstddef>
#include <cstring>
#include <fstream>
#include <sstream>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/yashe_hip.h"
#include "bigint.hpp"
#include "batch.hpp"
#include "internal.hpp"
#include "kernels.hpp"
#include "launch.hpp"
#include "ring_params.hpp"
#include "wire.hpp"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define YH_LAUNCH(expr)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ == hipErrorInvalidValue) return fail(YH_ERR_UNSUPPORTED, "shape out of range");  \
    if (e_ != hipSuccess) return fail(YH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define YH_HIP(expr)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (void)hipGetLastError(); /* consume, so the next call does not report it */           \
      return fail(YH_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    }                                                                                       \
  } while (0)

// Longest LDS-resident transform, and longest plain product (built from two
// transforms of the longest length and zero coefficients above them).
constexpr int YH_TRANSFORM_MAX = 16384;
constexpr long YH_PLAIN_MAX_N = 1L << 20;

#define YH_TRY_BEGIN try {
#define YH_TRY_END                                                   \
  }                                                                  \
  catch (const std::invalid_argument& e) { return fail(YH_ERR_ARG, e.what()); } \
  catch (const std::exception& e) { return fail(YH_ERR_INTERNAL, e.what()); }

struct DevTables {
  uint64_t* t16 = nullptr;  // radix-16 tables (one allocation)
  yhk::NttTab tab{};
};

// Uploads the radix-16 tables of length N into one device allocation.
int upload_ntt16(int N, uint64_t** buf, yhk::NttTab* tab);

}  // namespace

// Device scratch of yh_yashe_mul_keyswitch_static: one copy per GPU, held
// until the process ends. Host code never uses this name as an address (on
// the host it is only a stand-in); it uses the per-device address returned
// by hipGetSymbolAddress (static_scratch_ptr below).
static __device__ __align__(256) uint8_t scratch[YH_STATIC_SCRATCH_BYTES];

struct TimedLaunch {
  int kernel;
  hipEvent_t start, stop;
};

struct yh_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::map<int, DevTables> cyclic;  // N -> twiddles of the cyclic transform
  std::mutex mu;
  bool timing = false;
  unsigned timing_mask = ~0u;  // kernel ids (bit k = id k) timed while timing is on
  std::vector<TimedLaunch> pending;
  std::vector<hipEvent_t> event_pool;
  double total_ms[YH_KERNEL_COUNT] = {};
  long launches[YH_KERNEL_COUNT] = {};
  int cus = 0;           // compute units of the device
  // mul+keyswitch batches of at least this many ciphertexts run the fused
  // scale-and-digit kernel (one workgroup per ciphertext); smaller ones the
  // separate reconstruction and digit kernels, which spread one ciphertext
  // over more workgroups (0: eight per CU, two full rounds of the fused
  // kernel's four resident workgroups per CU)
  int fuse_min_batch = 0;
  std::string arch;      // gcnArchName
  bool capturing = false;
  // objects whose device memory the launches being captured reference; the
  // graph takes a reference on each when the capture ends
  std::vector<yh_ring*> cap_rings;
  std::vector<const yh_evk*> cap_evks;
  std::vector<yh_batch*> cap_batches;
  uint8_t* static_scratch = nullptr;  // this device's copy of the allocation-free entry point's array
  int live = 0;  // rings, batch plans and graphs created on this context and not yet destroyed
};

struct yh_graph {
  yh_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<yh_ring*> rings;
  std::vector<const yh_evk*> evks;
  std::vector<yh_batch*> batches;
};

struct yh_ring {
  yh_ctx* ctx = nullptr;
  std::unique_ptr<yh::RingParams> P;
  uint64_t* d_tables = nullptr;
  uint64_t* d_ntt = nullptr;   // radix-16 tables of length n
  yhk::RingDev rd{};
  uint32_t* d_y = nullptr;     // ICRT input scratch
  size_t y_bytes = 0;
  uint64_t* d_tmp = nullptr;   // ring-element scratch (mul+keyswitch intermediate)
  size_t tmp_bytes = 0;
  uint64_t* d_dig = nullptr;   // keyswitch digit transforms [ct][lw][n]
  size_t dig_bytes = 0;
  std::map<int, yh_ring*> sub;  // rings x^m + 1 over the same q, m < n (inverse recursion)
  hipStream_t stream = nullptr;  // the context stream, or the ring's own (yh_ring_use_own_stream)
  bool own_stream = false;
  // captured graphs holding pointers to this ring's tables and scratch
  // (while any exists the scratch cannot be reallocated), and batch plans
  // holding copies of its table pointers; while either exists the ring
  // cannot be destroyed
  int graph_users = 0;
  int plan_users = 0;
  int server_users = 0;  // request loops (yh_server) serving this ring
};

struct yh_event {
  int device = 0;
  hipEvent_t ev = nullptr;
};

struct yh_evk {
  yh_ring* ring = nullptr;
  uint64_t* d = nullptr;     // [lw][R_ks][n]: transform domain (n >= 16) or centred residues
  uint64_t* corr = nullptr;  // [R_ks][n]: correction of the centred digits (n >= 16)
  mutable int users = 0;     // captured graphs that read this key
  mutable int server_users = 0;  // request loops that use this key
};

namespace {

// Set while the allocation-free entry point runs: no event is created then.
thread_local bool t_no_alloc = false;

// Per-device address of `scratch`, looked up once per device (with that
// device current) when the first context on it is created.
std::mutex g_scratch_mu;
std::map<int, uint8_t*> g_scratch_ptr;

int static_scratch_ptr(int device, uint8_t** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  auto it = g_scratch_ptr.find(device);
  if (it != g_scratch_ptr.end()) { *out = it->second; return YH_OK; }
  void* p = nullptr;
  YH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) return fail(YH_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
  g_scratch_ptr[device] = (uint8_t*)p;
  *out = (uint8_t*)p;
  return YH_OK;
}

hipEvent_t take_event(yh_ctx* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Records an event pair around one launch when timing is on.
struct LaunchTimer {
  yh_ctx* ctx;
  int kernel;
  hipStream_t stream;
  hipEvent_t start = nullptr;
  LaunchTimer(yh_ctx* c, int k, hipStream_t s) : ctx(c), kernel(k), stream(s) {
    if (ctx->timing && !ctx->capturing && !t_no_alloc && (ctx->timing_mask >> kernel & 1u)) {
      start = take_event(ctx);
      if (start) (void)hipEventRecord(start, stream);
    }
  }
  ~LaunchTimer() {
    if (!start) return;
    hipEvent_t stop = take_event(ctx);
    if (!stop) return;
    (void)hipEventRecord(stop, stream);
    ctx->pending.push_back({kernel, start, stop});
  }
};

int set_dev(yh_ctx* ctx) {
  YH_HIP(hipSetDevice(ctx->device));
  return YH_OK;
}

// Selects the ring's device and, while the context stream is being captured,
// records that the graph will reference the ring's tables and scratch. Every
// ring entry point that launches work calls it first.
int ring_enter(yh_ring* r) {
  if (int s = set_dev(r->ctx)) return s;
  yh_ctx* c = r->ctx;
  if (c->capturing && r->stream == c->stream &&
      std::find(c->cap_rings.begin(), c->cap_rings.end(), r) == c->cap_rings.end())
    c->cap_rings.push_back(r);
  return YH_OK;
}

void note_evk(yh_ring* r, const yh_evk* k) {
  yh_ctx* c = r->ctx;
  if (c->capturing && r->stream == c->stream &&
      std::find(c->cap_evks.begin(), c->cap_evks.end(), k) == c->cap_evks.end())
    c->cap_evks.push_back(k);
}

// Grows one of the ring's scratch buffers (waits for the ring's work first).
// Growth frees the old buffer, so it is refused while a captured graph or a
// batch plan holds the ring (their launches keep the old addresses).
int ensure(yh_ring* r, void** buf, size_t* have, size_t need) {
  if (*have >= need) return YH_OK;
  if (r->ctx->capturing) return fail(YH_ERR_ARG, "scratch would grow during graph capture; run the sequence once first");
  if (r->graph_users > 0)
    return fail(YH_ERR_ARG,
                "ring scratch would be reallocated while a captured graph references it; run the largest batch "
                "before capturing, or destroy the graph first");
  if (*buf) {
    YH_HIP(hipStreamSynchronize(r->stream));
    YH_HIP(hipFree(*buf));
    *buf = nullptr;
    *have = 0;
  }
  YH_HIP(hipMalloc(buf, need));
  *have = need;
  return YH_OK;
}

// Device scratch of one fused operation: reconstruction rows, the
// mul+keyswitch intermediate and the keyswitch digit transforms. The
// allocating entry points take them from the ring (grown on demand), the
// allocation-free one from the library's static array.
struct Scratch {
  uint32_t* y = nullptr;
  uint64_t* tmp = nullptr;
  uint64_t* dig = nullptr;
};

size_t y_bytes_for(const yh_ring* r, int set, int batch) {
  return (size_t)batch * yhk::y_rows(r->rd.set_R[set]) * r->rd.n * sizeof(uint32_t);
}

int ring_y(yh_ring* r, size_t need, Scratch* s) {
  if (int st = ensure(r, (void**)&r->d_y, &r->y_bytes, need)) return st;
  s->y = r->d_y;
  return YH_OK;
}

int upload_ntt16(int N, uint64_t** buf, yhk::NttTab* tab) {
  yh::Ntt16Tables t = yh::ntt16_tables(N);
  std::vector<uint64_t> h;
  size_t o_f = h.size(); h.insert(h.end(), t.fwd.begin(), t.fwd.end());
  size_t o_i = h.size(); h.insert(h.end(), t.inv.begin(), t.inv.end());
  size_t o_t = h.size(); h.insert(h.end(), t.twist.begin(), t.twist.end());
  size_t o_u = h.size(); h.insert(h.end(), t.untwist_pw.begin(), t.untwist_pw.end());
  size_t o_m = h.size(); h.insert(h.end(), t.untwist_m.begin(), t.untwist_m.end());
  size_t o_f0 = h.size(); h.insert(h.end(), t.fwd0.begin(), t.fwd0.end());
  size_t o_ip = h.size(); h.insert(h.end(), t.inv0_pw.begin(), t.inv0_pw.end());
  size_t o_im = h.size(); h.insert(h.end(), t.inv0_m.begin(), t.inv0_m.end());
  YH_HIP(hipMalloc(buf, h.size() * sizeof(uint64_t)));
  YH_HIP(hipMemcpy(*buf, h.data(), h.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  tab->fwd = *buf + o_f;
  tab->inv = *buf + o_i;
  tab->twist = *buf + o_t;
  tab->untwist_pw = *buf + o_u;
  tab->untwist_m = *buf + o_m;
  tab->fwd0 = *buf + o_f0;
  tab->inv0_pw = *buf + o_ip;
  tab->inv0_m = *buf + o_im;
  tab->n_inv_pw = t.n_inv_pw;
  return YH_OK;
}

int get_cyclic(yh_ctx* ctx, int N, DevTables* out) {
  std::lock_guard<std::mutex> lk(ctx->mu);
  auto it = ctx->cyclic.find(N);
  if (it != ctx->cyclic.end()) { *out = it->second; return YH_OK; }
  if (ctx->capturing) return fail(YH_ERR_ARG, "transform tables would be built during graph capture");
  DevTables t;
  if (int st = upload_ntt16(N, &t.t16, &t.tab)) return st;
  ctx->cyclic[N] = t;
  *out = t;
  return YH_OK;
}

int ilog2i(long v) { int l = 0; while ((1L << l) < v) ++l; return l; }

using yhl::lds_bytes;

// Launch the fused product kernel for `batch` pairs on N-point transforms.
int launch_prod(yh_ring* r, yhk::ProdArgs pa, int N, int batch, bool negacyclic) {
  pa.batch = batch;
  const int R = r->rd.set_R[pa.set];
  LaunchTimer timer(r->ctx, YH_KERNEL_PROD, r->stream);
  if (N >= 16)
    YH_LAUNCH(yhl::prod(pa.logN, negacyclic, dim3(yhk::xcd_grid(batch, R)), dim3(N / 16), r->stream, r->rd, pa));
  else
    YH_LAUNCH(yhl::prod(pa.logN, negacyclic, dim3(R, batch), dim3(64), r->stream, r->rd, pa));
  return YH_OK;
}

int launch_icrt(yh_ring* r, const yhk::IcrtArgs& ia, int batch) {
  LaunchTimer timer(r->ctx, YH_KERNEL_ICRT, r->stream);
  const int LQ = r->rd.LQ;
  if ((ia.mode == yhk::EPI_MODQ || ia.mode == yhk::EPI_SCALE) &&
      yhk::icrt_pipelined(LQ, r->rd.set_R[ia.set], ia.mode)) {
    const int per = YH_ICRT_BLOCK * YH_ICRT_PIPE_COEFS;
    YH_LAUNCH(yhl::icrt_pipe(LQ, ia.mode, dim3((ia.ncoef + per - 1) / per, batch), dim3(YH_ICRT_BLOCK), r->stream,
                             r->rd, ia));
    return YH_OK;
  }
  // decryption and the other mod-q and scale-and-round shapes: a kernel per
  // epilogue (k_icrt holds every epilogue's registers: 200 VGPRs at eight
  // limbs against 62 for mod q alone and 98 for scale-and-round; 7-10 % more
  // mul+keyswitch/s on 255- and 511-bit rings); k_icrt keeps the raw integer
  if (ia.mode == yhk::EPI_DECRYPT || ((ia.mode == yhk::EPI_MODQ || ia.mode == yhk::EPI_SCALE) &&
                                       !yhk::icrt_pipelined(LQ, r->rd.set_R[ia.set], ia.mode))) {
    YH_LAUNCH(yhl::icrt_epi(LQ, ia.mode, dim3((ia.ncoef + YH_ICRT_BLOCK - 1) / YH_ICRT_BLOCK, batch),
                             dim3(YH_ICRT_BLOCK), r->stream, r->rd, ia));
    return YH_OK;
  }
  const int per = YH_ICRT_BLOCK * (yhk::icrt_pipelined(LQ, r->rd.set_R[ia.set], ia.mode) ? 2 : YH_ICRT_COEFS);
  YH_LAUNCH(yhl::icrt(LQ, dim3((ia.ncoef + per - 1) / per, batch), dim3(YH_ICRT_BLOCK), r->stream, r->rd, ia));
  return YH_OK;
}

int addsub_impl(yh_ring* r, const uint64_t* a, const uint64_t* b, uint64_t* c, int batch, int sub,
                yhk::AddsubLists lists) {
  const size_t count = (size_t)batch * r->rd.n;
  YH_LAUNCH(yhl::addsub_q(r->rd.LQ, dim3((unsigned)((count + 255) / 256)), dim3(256), r->stream, r->rd, a, b, c,
                          r->rd.n, count, sub, lists));
  return YH_OK;
}

yhk::IcrtArgs icrt_args(const uint32_t* Y, int N, int ncoef, int set, int mode, uint64_t* out, int out_limbs,
                        size_t out_stride, size_t out_pstride, int positive) {
  yhk::IcrtArgs ia{};
  ia.Y = Y;
  ia.N = N;
  ia.ncoef = ncoef;
  ia.set = set;
  ia.mode = mode;
  ia.out = out;
  ia.out_limbs = out_limbs;
  ia.out_stride = out_stride;
  ia.out_pstride = out_pstride;
  ia.positive = positive;
  return ia;
}

// Ring product with one of the reconstruction epilogues; sc.y holds at least
// y_bytes_for(r, 0, batch) bytes.
// rows_only: only the product kernel runs (its rows stay in sc.y for a
// fused consumer); out is not written.
int ring_product(yh_ring* r, const Scratch& sc, const uint64_t* a, const uint64_t* b, uint64_t* out, int batch,
                 int mode, const uint64_t* const* a_list = nullptr, const uint64_t* const* b_list = nullptr,
                 uint64_t* const* out_list = nullptr, bool rows_only = false) {
  const int n = r->rd.n, LQ = r->rd.LQ;
  int st;
  yhk::ProdArgs pa{};
  pa.A = a; pa.B = b;
  pa.lin_a = pa.lin_b = LQ;
  pa.ncoef_a = pa.ncoef_b = n;
  pa.lstride_a = pa.lstride_b = n;
  pa.ostride_a = pa.ostride_b = (size_t)LQ * n;
  pa.Y = sc.y;
  pa.logN = r->rd.logn;
  pa.tab = r->rd.tab;
  pa.set = 0;
  pa.a_list = a_list;
  pa.b_list = b_list;
  st = launch_prod(r, pa, n, batch, true);
  if (st || rows_only) return st;
  yhk::IcrtArgs ia = icrt_args(sc.y, n, n, 0, mode, out, LQ, n, (size_t)LQ * n, 0);
  ia.out_list = out_list;
  ia.reverse = 1;  // k_prod wrote the rows from the first pair to the last
  return launch_icrt(r, ia, batch);
}

size_t dig_bytes_for(const yh_ring* r, int batch) {
  return r->rd.n >= 16 ? (size_t)batch * r->rd.lw * r->rd.n * sizeof(uint64_t) : 0;
}

// Keyswitch of `batch` ciphertexts; sc.y holds y_bytes_for(r, 1, batch) bytes
// and sc.dig dig_bytes_for(r, batch). With `fused` (fusable_mulks) c is not
// read: the digit transforms come from the product rows in sc.y, which
// ring_product's k_prod left there, through k_ks_scale_digits.
int keyswitch_impl(yh_ring* r, const Scratch& sc, const yh_evk* k, const uint64_t* c, uint64_t* out, int batch,
                   uint64_t* const* out_list = nullptr, bool fused = false) {
  const int n = r->rd.n, LQ = r->rd.LQ, Rks = r->rd.set_R[1];
  note_evk(r, k);
  yhk::KsArgs ka{};
  ka.C = c;
  ka.ostride = (size_t)LQ * n;
  ka.evk = k->d;
  ka.corr = k->corr;
  ka.Y = sc.y;
  ka.batch = batch;
  // k_ks_digits writes the digit transforms from the first ciphertext to
  // the last, k_ks_scale_digits from the last to the first; k_ks_mac reads
  // them the other way round and writes its rows in its own order, which
  // the reconstruction reverses again
  bool mac_reverse = true;
  if (n >= 16) {
    ka.D = sc.dig;
    const bool w32 = r->rd.w_bits == 32;
    if (fused) {
      ka.P = sc.y;
      LaunchTimer timer(r->ctx, YH_KERNEL_KS_DIGITS, r->stream);
      YH_LAUNCH(yhl::ks_scale_digits(r->rd.logn, LQ, dim3(yhk::xcd_grid(batch, 1)), r->stream, r->rd, ka));
      mac_reverse = false;
    } else {
      LaunchTimer timer(r->ctx, YH_KERNEL_KS_DIGITS, r->stream);
      YH_LAUNCH(yhl::ks_digits(r->rd.logn, w32, dim3(yhk::xcd_grid(batch, r->rd.lw)), r->stream, r->rd, ka));
    }
    ka.reverse = mac_reverse ? 1 : 0;
    {
      LaunchTimer timer(r->ctx, YH_KERNEL_KEYSWITCH, r->stream);
      YH_LAUNCH(yhl::ks_mac(r->rd.logn, dim3(yhk::xcd_grid(batch, Rks)), r->stream, r->rd, ka));
    }
  } else {
    if (fused) return fail(YH_ERR_INTERNAL, "fused keyswitch below n = 16");
    LaunchTimer timer(r->ctx, YH_KERNEL_KEYSWITCH, r->stream);
    YH_LAUNCH(yhl::ks_small(dim3(Rks, batch), r->stream, r->rd, ka, k->d));
  }
  yhk::IcrtArgs ia = icrt_args(sc.y, n, n, 1, yhk::EPI_MODQ, out, LQ, n, (size_t)LQ * n, 0);
  ia.out_list = out_list;
  ia.reverse = mac_reverse ? 0 : 1;
  return launch_icrt(r, ia, batch);
}

// mul+keyswitch calls whose scale-and-round runs inside the digit
// transforms (k_ks_scale_digits): q of one or two limbs, w = 2^32, n >= 1024,
// product rows that the reconstruction holds in registers, and a batch that
// fills the device twice over with one workgroup per ciphertext. Measured at
// n = 4096, q = 2^127 - 1 against the separate kernels (the single-epilogue
// scale-and-round and k_ks_digits): slower below 1536 ciphertexts (512: 0.50
// vs 0.49 ms, one ciphertext 0.15 vs 0.09 ms), equal at 2048 (1.83 ms),
// about 1 % faster at 4096.
bool fusable_mulks(const yh_ring* r, int batch) {
  const int min_batch = r->ctx->fuse_min_batch > 0 ? r->ctx->fuse_min_batch : 8 * r->ctx->cus;
  return batch >= min_batch && r->rd.logn >= yhl::FUSED_MIN_LOGN && r->rd.LQ <= 2 && r->rd.w_bits == 32 &&
         yhk::icrt_pipelined(r->rd.LQ, r->rd.set_R[0], yhk::EPI_SCALE);
}

// YASHE multiply followed by the keyswitch; operands from arrays or
// pointer lists as in ring_product, output to `out` or `out_list`. sc as
// for mulks_scratch (sc.tmp is unused on fused rings).
int mul_keyswitch_impl(yh_ring* r, const Scratch& sc, const yh_evk* k, const uint64_t* c1, const uint64_t* c2,
                       uint64_t* out, int batch, const uint64_t* const* c1_list = nullptr,
                       const uint64_t* const* c2_list = nullptr, uint64_t* const* out_list = nullptr) {
  if (fusable_mulks(r, batch)) {
    if (int st = ring_product(r, sc, c1, c2, nullptr, batch, yhk::EPI_SCALE, c1_list, c2_list, nullptr, true))
      return st;
    return keyswitch_impl(r, sc, k, nullptr, out, batch, out_list, true);
  }
  if (int st = ring_product(r, sc, c1, c2, sc.tmp, batch, yhk::EPI_SCALE, c1_list, c2_list, nullptr)) return st;
  return keyswitch_impl(r, sc, k, sc.tmp, out, batch, out_list);
}

// Scratch of the allocating keyswitch: reconstruction rows and digit
// transforms from the ring.
int ks_scratch(yh_ring* r, int batch, Scratch* sc) {
  if (int st = ring_y(r, y_bytes_for(r, 1, batch), sc)) return st;
  if (r->rd.n >= 16) {
    if (int st = ensure(r, (void**)&r->d_dig, &r->dig_bytes, dig_bytes_for(r, batch))) return st;
    sc->dig = r->d_dig;
  }
  return YH_OK;
}

// y rows for both the product and the keyswitch of a mul+keyswitch, and the
// intermediate ring element.
int mulks_scratch(yh_ring* r, int batch, Scratch* sc) {
  const size_t yb = std::max(y_bytes_for(r, 0, batch), y_bytes_for(r, 1, batch));
  if (int st = ring_y(r, yb, sc)) return st;
  if (int st = ensure(r, (void**)&r->d_tmp, &r->tmp_bytes, (size_t)batch * r->rd.LQ * r->rd.n * sizeof(uint64_t)))
    return st;
  sc->tmp = r->d_tmp;
  if (r->rd.n >= 16) {
    if (int st = ensure(r, (void**)&r->d_dig, &r->dig_bytes, dig_bytes_for(r, batch))) return st;
    sc->dig = r->d_dig;
  }
  return YH_OK;
}

}  // namespace

// ====================================================================== C ABI
extern "C" {

static int drain_timing(yh_ctx* ctx);

const char* yh_last_error(void) { return g_err.c_str(); }
int yh_version(void) { return 100; }

int yh_ctx_create(int device, yh_ctx** out) {
  if (!out) return fail(YH_ERR_ARG, "null out");
  int count = 0;
  YH_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) return fail(YH_ERR_ARG, "no such device");
  YH_HIP(hipSetDevice(device));
  // Transforms keep a whole length-16384 vector (+1/16 padding) in LDS:
  // 139264 bytes, within gfx950's 160 KiB per workgroup.
  yhl::prod_attrs();
  yhl::ks_attrs();
  yhl::misc_attrs();
  yhl::batch_attrs();
  auto* c = new yh_ctx();
  c->device = device;
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
      c->cus = prop.multiProcessorCount;
      c->arch = prop.gcnArchName;
    }
    if (c->cus <= 0) c->cus = 256;
  }
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return fail(YH_ERR_HIP, hipGetErrorString(e)); }
  if (int st = static_scratch_ptr(device, &c->static_scratch)) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    return st;
  }
  *out = c;
  return YH_OK;
}

int yh_ctx_destroy(yh_ctx* ctx) {
  if (!ctx) return YH_OK;
  if (ctx->live > 0)
    return fail(YH_ERR_ARG, "context still has " + std::to_string(ctx->live) +
                                " ring(s), batch plan(s), graph(s) or request loop(s); destroy them first");
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  drain_timing(ctx);
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  for (auto& kv : ctx->cyclic) (void)hipFree(kv.second.t16);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return YH_OK;
}

void* yh_ctx_stream(yh_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

static int drain_timing(yh_ctx* ctx) {
  for (auto& t : ctx->pending) {
    YH_HIP(hipEventSynchronize(t.stop));
    float ms = 0;
    YH_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
    ctx->total_ms[t.kernel] += ms;
    ctx->launches[t.kernel] += 1;
    ctx->event_pool.push_back(t.start);
    ctx->event_pool.push_back(t.stop);
  }
  ctx->pending.clear();
  return YH_OK;
}

int yh_ctx_set_timing(yh_ctx* ctx, int enable) {
  if (!ctx) return fail(YH_ERR_ARG, "null ctx");
  if (int s = set_dev(ctx)) return s;
  if (int s = drain_timing(ctx)) return s;
  ctx->timing = enable != 0;
  return YH_OK;
}

int yh_ctx_set_timing_mask(yh_ctx* ctx, unsigned mask) {
  if (!ctx) return fail(YH_ERR_ARG, "null ctx");
  if (int s = set_dev(ctx)) return s;
  if (int s = drain_timing(ctx)) return s;
  ctx->timing_mask = mask;
  return YH_OK;
}

int yh_ctx_set_fuse_min_batch(yh_ctx* ctx, int min_batch) {
  if (!ctx || min_batch < 0) return fail(YH_ERR_ARG, "bad argument");
  ctx->fuse_min_batch = min_batch;
  return YH_OK;
}

int yh_ctx_reset_timing(yh_ctx* ctx) {
  if (!ctx) return fail(YH_ERR_ARG, "null ctx");
  if (int s = set_dev(ctx)) return s;
  if (int s = drain_timing(ctx)) return s;
  for (int k = 0; k < YH_KERNEL_COUNT; ++k) { ctx->total_ms[k] = 0; ctx->launches[k] = 0; }
  return YH_OK;
}

int yh_ctx_kernel_time(yh_ctx* ctx, int kernel_id, double* total_ms, long* launches) {
  if (!ctx || kernel_id < 0 || kernel_id >= YH_KERNEL_COUNT) return fail(YH_ERR_ARG, "bad argument");
  if (int s = set_dev(ctx)) return s;
  if (int s = drain_timing(ctx)) return s;
  if (total_ms) *total_ms = ctx->total_ms[kernel_id];
  if (launches) *launches = ctx->launches[kernel_id];
  return YH_OK;
}

int yh_ctx_sync(yh_ctx* ctx) {
  if (!ctx) return fail(YH_ERR_ARG, "null ctx");
  if (ctx->capturing) return fail(YH_ERR_ARG, "synchronize during graph capture");
  if (int s = set_dev(ctx)) return s;
  YH_HIP(hipStreamSynchronize(ctx->stream));
  return YH_OK;
}

int yh_malloc(yh_ctx* ctx, size_t bytes, void** dptr) {
  if (!ctx || !dptr) return fail(YH_ERR_ARG, "null argument");
  if (int s = set_dev(ctx)) return s;
  hipError_t e = hipMalloc(dptr, bytes ? bytes : 1);
  if (e == hipErrorOutOfMemory) return fail(YH_ERR_NOMEM, "out of device memory");
  if (e != hipSuccess) return fail(YH_ERR_HIP, hipGetErrorString(e));
  return YH_OK;
}

int yh_free(yh_ctx* ctx, void* dptr) {
  if (!ctx) return fail(YH_ERR_ARG, "null ctx");
  if (!dptr) return YH_OK;
  if (ctx->capturing) return fail(YH_ERR_ARG, "free during graph capture");
  if (int s = set_dev(ctx)) return s;
  YH_HIP(hipStreamSynchronize(ctx->stream));
  YH_HIP(hipFree(dptr));
  return YH_OK;
}

static int memcpy_kind(yh_ctx* ctx, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  if (!ctx) return fail(YH_ERR_ARG, "null ctx");
  if (ctx->capturing) return fail(YH_ERR_ARG, "synchronous copy during graph capture");
  if (bytes == 0) return YH_OK;
  if (!dst || !src) return fail(YH_ERR_ARG, "null pointer");
  if (int s = set_dev(ctx)) return s;
  YH_HIP(hipMemcpyAsync(dst, src, bytes, kind, ctx->stream));
  YH_HIP(hipStreamSynchronize(ctx->stream));
  return YH_OK;
}
int yh_memcpy_htod(yh_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return memcpy_kind(ctx, dst, src, bytes, hipMemcpyHostToDevice);
}
int yh_memcpy_dtoh(yh_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return memcpy_kind(ctx, dst, src, bytes, hipMemcpyDeviceToHost);
}
int yh_memcpy_dtod(yh_ctx* ctx, void* dst, const void* src, size_t bytes) {
  return memcpy_kind(ctx, dst, src, bytes, hipMemcpyDeviceToDevice);
}
int yh_memset(yh_ctx* ctx, void* dst, int value, size_t bytes) {
  if (!ctx || !dst) return fail(YH_ERR_ARG, "null argument");
  if (ctx->capturing) return fail(YH_ERR_ARG, "memset nodes are not captured");
  if (int s = set_dev(ctx)) return s;
  YH_HIP(hipMemsetAsync(dst, value, bytes, ctx->stream));
  return YH_OK;
}

int yh_ring_create(yh_ctx* ctx, int n, const uint64_t* q, int q_limbs, uint64_t t, int w_bits, yh_ring** out) {
  if (!ctx || !q || !out || q_limbs <= 0) return fail(YH_ERR_ARG, "null or empty argument");
  YH_TRY_BEGIN
  if (int s = set_dev(ctx)) return s;
  auto r = std::make_unique<yh_ring>();
  r->ctx = ctx;
  r->stream = ctx->stream;
  r->P = std::make_unique<yh::RingParams>(n, yh::BigInt::from_limbs(q, q_limbs), t, w_bits);
  const yh::RingParams& P = *r->P;
  // one table buffer: pack every array, remember offsets
  std::vector<uint64_t> h;
  auto put = [&h](const std::vector<uint64_t>& v) { size_t o = h.size(); h.insert(h.end(), v.begin(), v.end()); return o; };
  size_t o_pr = put(P.primes), o_b = put(P.barrett), o_pw = put(P.pow32);
  size_t o_pr1 = put(P.primes_ks), o_b1 = put(P.barrett_ks), o_pw1 = put(P.pow32_ks);
  size_t o_q = put(P.q_limbs), o_mu = put(P.mu_limbs);
  size_t o_mi0 = put(P.full.Mi_limbs), o_inv0 = put(P.full.invMi), o_m0 = put(P.full.M_limbs);
  size_t o_mi1 = put(P.ks.Mi_limbs), o_inv1 = put(P.ks.invMi), o_m1 = put(P.ks.M_limbs);
  size_t o_mq0 = put(P.full.Mq_limbs), o_nmq0 = put(P.full.negMq_limbs);
  size_t o_mq1 = put(P.ks.Mq_limbs), o_nmq1 = put(P.ks.negMq_limbs);
  size_t o_al = put(P.scale.alpha_limbs), o_be = put(P.scale.beta_limbs);
  // constants of the small primes, one 32-byte row per modulus (row 0 = P unused)
  auto small_rows = [](const std::vector<uint64_t>& pr, const std::vector<uint64_t>& inv) {
    std::vector<uint64_t> w(pr.size() * sizeof(yhk::SmallPrime) / sizeof(uint64_t), 0);
    for (size_t i = 1; i < pr.size(); ++i) {
      const yhk::SmallPrime k = yhk::make_small_prime((uint32_t)pr[i], (uint32_t)inv[i]);
      memcpy(&w[i * sizeof(yhk::SmallPrime) / sizeof(uint64_t)], &k, sizeof k);
    }
    return w;
  };
  size_t o_sp0 = put(small_rows(P.primes, P.full.invMi)), o_sp1 = put(small_rows(P.primes_ks, P.ks.invMi));
  size_t o_nbm = put(P.scale.negbetaM), o_na1 = put(P.scale.negalpha1);
  YH_HIP(hipMalloc(&r->d_tables, h.size() * sizeof(uint64_t)));
  YH_HIP(hipMemcpy(r->d_tables, h.data(), h.size() * sizeof(uint64_t), hipMemcpyHostToDevice));
  uint64_t* d = r->d_tables;
  yhk::RingDev& rd = r->rd;
  rd.n = P.n; rd.logn = P.logn; rd.LQ = P.LQ; rd.LW = P.LW;
  rd.R = (int)P.primes.size(); rd.w_bits = P.w_bits; rd.lw = P.lw;
  rd.t = P.t;
  if (P.n >= 16) {
    if (int st = upload_ntt16(P.n, &r->d_ntt, &rd.tab)) return st;
  }
  rd.primes[0] = d + o_pr; rd.barrett[0] = d + o_b; rd.pow32[0] = d + o_pw;
  rd.primes[1] = d + o_pr1; rd.barrett[1] = d + o_b1; rd.pow32[1] = d + o_pw1;
  rd.q = d + o_q; rd.mu = d + o_mu;
  rd.set_R[0] = P.full.R; rd.Mi[0] = d + o_mi0; rd.invMi[0] = d + o_inv0; rd.M[0] = d + o_m0;
  rd.set_R[1] = P.ks.R; rd.Mi[1] = d + o_mi1; rd.invMi[1] = d + o_inv1; rd.M[1] = d + o_m1;
  rd.Mq[0] = d + o_mq0; rd.negMq[0] = d + o_nmq0;
  rd.Mq[1] = d + o_mq1; rd.negMq[1] = d + o_nmq1;
  rd.alpha = d + o_al; rd.beta = d + o_be; rd.negbetaM = d + o_nbm; rd.negalpha1 = d + o_na1;
  rd.fast_div = P.fast_div ? 1 : 0;
  rd.q_shift = P.q_shift;
  rd.qmu = P.qmu;
  rd.sp[0] = reinterpret_cast<const yhk::SmallPrime*>(d + o_sp0);
  rd.sp[1] = reinterpret_cast<const yhk::SmallPrime*>(d + o_sp1);
  ctx->live++;
  *out = r.release();
  return YH_OK;
  YH_TRY_END
}

int yh_ring_destroy(yh_ring* r) {
  if (!r) return YH_OK;
  if (r->graph_users + r->plan_users + r->server_users > 0)
    return fail(YH_ERR_ARG, "ring is referenced by " + std::to_string(r->graph_users) + " captured graph(s), " +
                                std::to_string(r->plan_users) + " batch plan(s) and " +
                                std::to_string(r->server_users) + " request loop(s); destroy them first");
  (void)hipSetDevice(r->ctx->device);
  (void)hipStreamSynchronize(r->stream);
  for (auto& kv : r->sub) yh_ring_destroy(kv.second);
  r->ctx->live--;
  if (r->own_stream) (void)hipStreamDestroy(r->stream);
  (void)hipFree(r->d_tables);
  if (r->d_ntt) (void)hipFree(r->d_ntt);
  if (r->d_y) (void)hipFree(r->d_y);
  if (r->d_tmp) (void)hipFree(r->d_tmp);
  if (r->d_dig) (void)hipFree(r->d_dig);
  delete r;
  return YH_OK;
}

// The R product moduli, then (when cap allows) the R_ks keyswitch moduli.
static int copy_moduli(const yh::RingParams& P, uint64_t* out, int cap) {
  const int R = (int)P.primes.size(), Rk = P.w_bits ? (int)P.primes_ks.size() : 0;
  if (cap < R) return fail(YH_ERR_ARG, "capacity too small");
  for (int i = 0; i < R; ++i) out[i] = P.primes[i];
  if (cap >= R + Rk)
    for (int i = 0; i < Rk; ++i) out[R + i] = P.primes_ks[i];
  return YH_OK;
}

static void fill_info(const yh::RingParams& P, yh_ring_info* info) {
  info->n = P.n;
  info->q_limbs = P.LQ;
  info->q_bits = P.q.bits();
  info->R = (int)P.primes.size();
  info->R_ks = P.w_bits ? P.ks.R : 0;
  info->w_bits = P.w_bits;
  info->lw = P.lw;
  info->work_limbs = P.LW;
  info->m_bits = P.full.M.bits();
  info->m_ks_bits = P.ks.M.bits();
  info->t = P.t;
  info->prime_max = P.pmax;
}

int yh_ring_get_info(const yh_ring* r, yh_ring_info* info) {
  if (!r || !info) return fail(YH_ERR_ARG, "null argument");
  fill_info(*r->P, info);
  return YH_OK;
}

int yh_ring_plan(int n, const uint64_t* q, int q_limbs, uint64_t t, int w_bits, yh_ring_info* info,
                 uint64_t* primes, int cap) {
  if (!q || !info || q_limbs <= 0) return fail(YH_ERR_ARG, "null or empty argument");
  YH_TRY_BEGIN
  yh::RingParams P(n, yh::BigInt::from_limbs(q, q_limbs), t, w_bits);
  fill_info(P, info);
  if (primes) {
    if (int st = copy_moduli(P, primes, cap)) return st;
  }
  return YH_OK;
  YH_TRY_END
}

int yh_ring_get_primes(const yh_ring* r, uint64_t* out, int cap) {
  if (!r || !out) return fail(YH_ERR_ARG, "null argument");
  return copy_moduli(*r->P, out, cap);
}

int yh_ring_get_modulus(const yh_ring* r, int set, uint64_t* out, int cap) {
  if (!r || !out || (set != 0 && set != 1)) return fail(YH_ERR_ARG, "bad argument");
  const yh::IcrtSet& s = set ? r->P->ks : r->P->full;
  if (cap < r->P->LW) return fail(YH_ERR_ARG, "capacity too small");
  for (int l = 0; l < r->P->LW; ++l) out[l] = s.M_limbs[l];
  return YH_OK;
}

int yh_crt(yh_ring* r, const uint64_t* d_coefs, int limbs, int ncoef, int spacing, uint64_t* d_res) {
  if (!r || !d_res || (ncoef > 0 && !d_coefs)) return fail(YH_ERR_ARG, "null argument");
  if (limbs < 1 || limbs > r->rd.LW) return fail(YH_ERR_ARG, "limbs out of range");
  if (ncoef < 0 || spacing < ncoef || spacing <= 0) return fail(YH_ERR_ARG, "bad spacing");
  if (int s = ring_enter(r)) return s;
  dim3 block(256), grid((spacing + 255) / 256, r->rd.R);
  YH_LAUNCH(yhl::crt(grid, block, r->stream, r->rd, d_coefs, limbs, ncoef, (size_t)ncoef, d_res, spacing));
  return YH_OK;
}

// ICRT of residues into d_coefs, or (d_coefs null) into the ring's staging
// buffer d_tmp.
static int icrt_impl(yh_ring* r, const uint64_t* d_res, int spacing, int ncoef, uint64_t* d_coefs, int out_limbs) {
  if (!r || !d_res) return fail(YH_ERR_ARG, "null argument");
  if (ncoef <= 0 || ncoef > spacing || out_limbs < 1) return fail(YH_ERR_ARG, "bad sizes");
  if (int s = ring_enter(r)) return s;
  Scratch sc;
  if (int st = ring_y(r, (size_t)yhk::y_rows(r->rd.R) * spacing * sizeof(uint32_t), &sc)) return st;
  if (!d_coefs) {
    if (int st = ensure(r, (void**)&r->d_tmp, &r->tmp_bytes, (size_t)out_limbs * ncoef * sizeof(uint64_t))) return st;
    d_coefs = r->d_tmp;
  }
  dim3 block(256), grid((spacing + 255) / 256, r->rd.R);
  YH_LAUNCH(yhl::residue_to_y(grid, block, r->stream, r->rd, d_res, spacing, sc.y, 0));
  return launch_icrt(r, icrt_args(sc.y, spacing, ncoef, 0, yhk::EPI_RAW, d_coefs, out_limbs, ncoef, 0, 1), 1);
}

int yh_icrt(yh_ring* r, const uint64_t* d_res, int spacing, int ncoef, uint64_t* d_coefs, int out_limbs) {
  if (!d_coefs) return fail(YH_ERR_ARG, "null argument");
  return icrt_impl(r, d_res, spacing, ncoef, d_coefs, out_limbs);
}

int yh_ntt(yh_ctx* ctx, uint64_t* d_data, int N, int count, int direction) {
  if (!ctx || !d_data) return fail(YH_ERR_ARG, "null argument");
  if (N < 16 || N > 16384 || (N & (N - 1))) return fail(YH_ERR_UNSUPPORTED, "N must be a power of two in [16, 16384]");
  if (count <= 0) return fail(YH_ERR_ARG, "count must be positive");
  if (int s = set_dev(ctx)) return s;
  DevTables t;
  if (int s = get_cyclic(ctx, N, &t)) return s;
  uint64_t Ninv = gl::inv((uint64_t)N);
  const int inv = direction == YH_INVERSE ? 1 : 0;
  YH_LAUNCH(yhl::ntt_cyclic(ilog2i(N), count, ctx->stream, d_data, t.tab, Ninv, inv));
  return YH_OK;
}

int yh_pointwise_mulmod(yh_ctx* ctx, uint64_t* d_a, const uint64_t* d_b, size_t count) {
  if (!ctx || !d_a || !d_b) return fail(YH_ERR_ARG, "null argument");
  if (count == 0) return YH_OK;
  if (int s = set_dev(ctx)) return s;
  YH_LAUNCH(yhl::pointwise(dim3((unsigned)((count + 255) / 256)), dim3(256), ctx->stream, d_a, d_b, count));
  return YH_OK;
}

int yh_residue_addsub(yh_ring* r, uint64_t* d_c, const uint64_t* d_a, const uint64_t* d_b, int spacing, int op) {
  if (!r || !d_c || !d_a || !d_b || spacing <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (int s = ring_enter(r)) return s;
  dim3 block(256), grid((spacing + 255) / 256, r->rd.R);
  YH_LAUNCH(yhl::residue_addsub(grid, block, r->stream, r->rd, d_c, d_a, d_b, spacing, op == YH_SUB ? 1 : 0));
  return YH_OK;
}

int yh_poly_mul_plain(yh_ring* r, const uint64_t* d_a, int deg_a, const uint64_t* d_b, int deg_b, int min_N,
                      uint64_t* d_out, int out_limbs, int* N_out) {
  if (!r || !d_a || !d_b || !d_out) return fail(YH_ERR_ARG, "null argument");
  const int n = r->rd.n;
  if (deg_a < 0 || deg_b < 0 || deg_a >= n || deg_b >= n) return fail(YH_ERR_ARG, "operand degrees must be in [0, n)");
  if (out_limbs < 1) return fail(YH_ERR_ARG, "out_limbs must be positive");
  if (min_N < 1 || (min_N & (min_N - 1))) return fail(YH_ERR_ARG, "min_N must be a power of two");
  const long N = yh_poly_mul_plain_length(deg_a, deg_b, min_N);
  if (N > YH_PLAIN_MAX_N) return fail(YH_ERR_UNSUPPORTED, "plain product longer than 2^20");
  if (int s = ring_enter(r)) return s;
  const int R = r->rd.R, LQ = r->rd.LQ;
  Scratch sc;
  if (N > YH_TRANSFORM_MAX) {
    // Every operand has at most n <= h coefficients (h the longest transform)
    // and deg a + deg b <= 2n - 2 < N, so the length-N product is the linear
    // one. It is assembled from a cyclic and a negacyclic product of length h;
    // the negacyclic one is needed only when the product reaches x^h, and then
    // n = h and the ring's own transform tables are the length-h ones.
    const int h = YH_TRANSFORM_MAX;
    const bool wraps = deg_a + deg_b >= h;
    if (wraps && r->rd.n != h) return fail(YH_ERR_ARG, "operand degrees must be in [0, n)");
    const size_t rows = yhk::y_rows(R);
    int st = ring_y(r, rows * (2 * (size_t)h + N) * sizeof(uint32_t), &sc);
    if (st) return st;
    uint32_t* Yu = sc.y;
    uint32_t* Yv = Yu + rows * h;
    uint32_t* Yc = Yv + rows * h;
    DevTables tw{};
    if (int s = get_cyclic(r->ctx, h, &tw)) return s;
    yhk::ProdArgs pa{};
    pa.A = d_a; pa.B = d_b;
    pa.lin_a = pa.lin_b = LQ;
    pa.ncoef_a = deg_a + 1; pa.ncoef_b = deg_b + 1;
    pa.lstride_a = deg_a + 1; pa.lstride_b = deg_b + 1;
    pa.logN = ilog2i(h);
    pa.set = 0;
    pa.Y = Yu;
    pa.tab = tw.tab;
    if ((st = launch_prod(r, pa, h, 1, false))) return st;
    if (wraps) {
      pa.Y = Yv;
      pa.tab = r->rd.tab;
      if ((st = launch_prod(r, pa, h, 1, true))) return st;
    }
    YH_LAUNCH(yhl::join_halves(dim3((unsigned)((N + 255) / 256), R), dim3(256), r->stream, r->rd, Yu,
                               wraps ? Yv : nullptr, Yc, h, (int)N));
    if (N_out) *N_out = (int)N;
    return launch_icrt(r, icrt_args(Yc, (int)N, (int)N, 0, yhk::EPI_RAW, d_out, out_limbs, N, 0, 1), 1);
  }
  int st = ring_y(r, (size_t)yhk::y_rows(R) * N * sizeof(uint32_t), &sc);
  if (st) return st;
  DevTables tw{};
  if (N >= 16) {
    if (int s = get_cyclic(r->ctx, (int)N, &tw)) return s;
  }
  yhk::ProdArgs pa{};
  pa.A = d_a; pa.B = d_b;
  pa.lin_a = pa.lin_b = LQ;
  pa.ncoef_a = deg_a + 1; pa.ncoef_b = deg_b + 1;
  pa.lstride_a = deg_a + 1; pa.lstride_b = deg_b + 1;
  pa.ostride_a = pa.ostride_b = 0;
  pa.Y = sc.y;
  pa.logN = ilog2i(N);
  pa.tab = tw.tab;
  pa.set = 0;
  st = launch_prod(r, pa, (int)N, 1, false);
  if (st) return st;
  if (N_out) *N_out = (int)N;
  return launch_icrt(r, icrt_args(sc.y, (int)N, (int)N, 0, yhk::EPI_RAW, d_out, out_limbs, N, 0, 1), 1);
}

int yh_poly_mul_plain_length(int deg_a, int deg_b, int min_N) {
  // 2^ceil(log2(deg a + deg b)) as in src/common.h:139, at least min_N and 2,
  // and long enough to hold every operand coefficient (deg_b = 0 with deg_a a
  // power of two)
  long N = 1;
  while (N < (long)deg_a + deg_b) N <<= 1;
  if (N < min_N) N = min_N;
  if (N < 2) N = 2;
  while (N < (long)deg_a + 1 || N < (long)deg_b + 1) N <<= 1;
  return (int)N;
}

int yh_ring_mul(yh_ring* r, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out, int batch) {
  if (!r || !d_a || !d_b || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (int s = ring_enter(r)) return s;
  Scratch sc;
  if (int st = ring_y(r, y_bytes_for(r, 0, batch), &sc)) return st;
  return ring_product(r, sc, d_a, d_b, d_out, batch, yhk::EPI_MODQ);
}

int yh_ring_addsub(yh_ring* r, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out, int batch, int op) {
  if (!r || !d_a || !d_b || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (int s = ring_enter(r)) return s;
  return addsub_impl(r, d_a, d_b, d_out, batch, op == YH_SUB ? 1 : 0, yhk::AddsubLists{nullptr, nullptr, nullptr});
}

int yh_ring_reduce(yh_ring* r, const uint64_t* d_c, int in_limbs, uint64_t* d_out, int batch) {
  if (!r || !d_c || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (in_limbs < 1 || in_limbs > r->rd.LW) return fail(YH_ERR_ARG, "in_limbs out of range");
  if (int s = ring_enter(r)) return s;
  YH_LAUNCH(yhl::reduce(r->rd.LQ, dim3((r->rd.n + 255) / 256, batch), dim3(256), r->stream, r->rd, d_c, in_limbs,
                        d_out));
  return YH_OK;
}

// The scalar (host limbs, any width) reduced mod m, as limbs into v[0..L).
static int scalar_mod(const uint64_t* s, int s_limbs, const yh::BigInt& m, uint64_t* v, int L) {
  if (!s || s_limbs < 1 || s_limbs > 1024) return fail(YH_ERR_ARG, "scalar must have 1 to 1024 limbs");
  yh::BigInt x = yh::BigInt::from_limbs(s, s_limbs) % m;
  x.to_limbs(v, L);
  return YH_OK;
}

int yh_ring_scalar(yh_ring* r, const uint64_t* d_a, const uint64_t* scalar, int scalar_limbs, uint64_t* d_out,
                   int batch, int op) {
  if (!r || !d_a || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (op != YH_ADD && op != YH_SUB && op != YH_MUL) return fail(YH_ERR_ARG, "op must be YH_ADD, YH_SUB or YH_MUL");
  YH_TRY_BEGIN
  yhk::ScalarArg sa{};
  if (int st = scalar_mod(scalar, scalar_limbs, r->P->q, sa.v, r->rd.LQ)) return st;
  if (int s = ring_enter(r)) return s;
  const size_t count = (size_t)batch * r->rd.n;
  const dim3 block(256), grid((unsigned)((count + 255) / 256));
  YH_LAUNCH(yhl::scalar_q(r->rd.LQ, grid, block, r->stream, r->rd, d_a, d_out, r->rd.n, count, op, sa));
  return YH_OK;
  YH_TRY_END
}

int yh_residue_scalar(yh_ring* r, uint64_t* d_c, const uint64_t* d_a, int spacing, int op,
                      const uint64_t* d_scalar_res) {
  if (!r || !d_c || !d_a || !d_scalar_res || spacing <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (op != YH_ADD && op != YH_SUB && op != YH_MUL) return fail(YH_ERR_ARG, "op must be YH_ADD, YH_SUB or YH_MUL");
  if (int s = ring_enter(r)) return s;
  const dim3 block(256), grid((spacing + 255) / 256, r->rd.R);
  YH_LAUNCH(yhl::residue_scalar(grid, block, r->stream, r->rd, d_c, d_a, spacing, op, d_scalar_res));
  return YH_OK;
}

static_assert(sizeof(yh_bn_t) == sizeof(yhk::BnT) && offsetof(yh_bn_t, dp) == offsetof(yhk::BnT, dp),
              "yh_bn_t mirrors the device view");

int yh_crt_bn(yh_ring* r, const yh_bn_t* d_coefs, int ncoef, int spacing, uint64_t* d_res) {
  if (!r || !d_res || (ncoef > 0 && !d_coefs)) return fail(YH_ERR_ARG, "null argument");
  if (ncoef < 0 || spacing < ncoef || spacing <= 0) return fail(YH_ERR_ARG, "bad spacing");
  if (int s = ring_enter(r)) return s;
  const dim3 block(256), grid((spacing + 255) / 256, r->rd.R);
  YH_LAUNCH(yhl::crt_bn(grid, block, r->stream, r->rd, reinterpret_cast<const yhk::BnT*>(d_coefs), ncoef, d_res,
                        spacing));
  return YH_OK;
}

int yh_icrt_bn(yh_ring* r, const uint64_t* d_res, int spacing, int ncoef, yh_bn_t* d_coefs) {
  if (!r || !d_res || !d_coefs) return fail(YH_ERR_ARG, "null argument");
  if (ncoef <= 0 || ncoef > spacing) return fail(YH_ERR_ARG, "bad sizes");
  const int LW = r->rd.LW;
  if (int st = icrt_impl(r, d_res, spacing, ncoef, nullptr, LW)) return st;  // into the ring's staging buffer
  YH_LAUNCH(yhl::to_bn(dim3((ncoef + 255) / 256), dim3(256), r->stream, r->d_tmp, LW, ncoef,
                       reinterpret_cast<yhk::BnT*>(d_coefs)));
  return YH_OK;
}

int yh_yashe_mul(yh_ring* r, const uint64_t* d_c1, const uint64_t* d_c2, uint64_t* d_out, int batch) {
  if (!r || !d_c1 || !d_c2 || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (r->rd.t == 0) return fail(YH_ERR_ARG, "ring has no plaintext modulus t");
  if (int s = ring_enter(r)) return s;
  Scratch sc;
  if (int st = ring_y(r, y_bytes_for(r, 0, batch), &sc)) return st;
  return ring_product(r, sc, d_c1, d_c2, d_out, batch, yhk::EPI_SCALE);
}

int yh_evk_create(yh_ring* r, const uint64_t* d_gamma, yh_evk** out) {
  if (!r || !d_gamma || !out) return fail(YH_ERR_ARG, "null argument");
  if (r->rd.w_bits == 0) return fail(YH_ERR_ARG, "ring has no keyswitch base");
  if (r->ctx->capturing) return fail(YH_ERR_ARG, "yh_evk_create allocates; create keys before capturing");
  if (int s = set_dev(r->ctx)) return s;
  auto* k = new yh_evk();
  k->ring = r;
  size_t bytes = (size_t)r->rd.lw * r->rd.set_R[1] * r->rd.n * sizeof(uint64_t);
  hipError_t e = hipMalloc(&k->d, bytes);
  if (e != hipSuccess) { delete k; return fail(YH_ERR_NOMEM, hipGetErrorString(e)); }
  if (r->rd.n >= 16) {
    e = hipMalloc(&k->corr, (size_t)r->rd.set_R[1] * r->rd.n * sizeof(uint64_t));
    if (e != hipSuccess) { (void)hipFree(k->d); delete k; return fail(YH_ERR_NOMEM, hipGetErrorString(e)); }
    e = yhl::evk_prepare(r->rd.logn, r->stream, r->rd, d_gamma, k->d, k->corr);
  } else {
    e = yhl::evk_residues(r->stream, r->rd, d_gamma, k->d);
  }
  if (e != hipSuccess) {
    (void)hipFree(k->d);
    if (k->corr) (void)hipFree(k->corr);
    delete k;
    return fail(YH_ERR_HIP, hipGetErrorString(e));
  }
  *out = k;
  return YH_OK;
}

int yh_evk_destroy(yh_evk* k) {
  if (!k) return YH_OK;
  if (k->users + k->server_users > 0)
    return fail(YH_ERR_ARG, "keyswitch key is referenced by " + std::to_string(k->users) + " captured graph(s) and " +
                                std::to_string(k->server_users) + " request loop(s); destroy them first");
  (void)hipSetDevice(k->ring->ctx->device);
  (void)hipStreamSynchronize(k->ring->stream);
  (void)hipFree(k->d);
  if (k->corr) (void)hipFree(k->corr);
  delete k;
  return YH_OK;
}

int yh_keyswitch(yh_ring* r, const yh_evk* k, const uint64_t* d_c, uint64_t* d_out, int batch) {
  if (!r || !k || !d_c || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (k->ring != r) return fail(YH_ERR_ARG, "key belongs to another ring");
  if (d_c == d_out) return fail(YH_ERR_ARG, "keyswitch cannot run in place");
  if (int s = ring_enter(r)) return s;
  Scratch sc;
  if (int st = ks_scratch(r, batch, &sc)) return st;
  return keyswitch_impl(r, sc, k, d_c, d_out, batch);
}

int yh_yashe_mul_keyswitch(yh_ring* r, const yh_evk* k, const uint64_t* d_c1, const uint64_t* d_c2,
                           uint64_t* d_out, int batch) {
  if (!r || !k || !d_c1 || !d_c2 || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (k->ring != r) return fail(YH_ERR_ARG, "key belongs to another ring");
  if (r->rd.t == 0) return fail(YH_ERR_ARG, "ring has no plaintext modulus t");
  if (int s = ring_enter(r)) return s;
  Scratch sc;
  if (int st = mulks_scratch(r, batch, &sc)) return st;
  return mul_keyswitch_impl(r, sc, k, d_c1, d_c2, d_out, batch);
}

int yh_yashe_mul_keyswitch_static(yh_ring* r, const yh_evk* k, const uint64_t* d_c1, const uint64_t* d_c2,
                                  uint64_t* d_out, int batch) {
  if (!r || !k || !d_c1 || !d_c2 || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (k->ring != r) return fail(YH_ERR_ARG, "key belongs to another ring");
  if (r->rd.t == 0) return fail(YH_ERR_ARG, "ring has no plaintext modulus t");
  if (r->rd.n < 16) return fail(YH_ERR_UNSUPPORTED, "the allocation-free entry point takes n >= 16");
  // layout of the static array: reconstruction rows | intermediate | digit transforms
  auto al = [](size_t v) { return (v + 255) & ~(size_t)255; };
  const size_t yb = al(std::max(y_bytes_for(r, 0, batch), y_bytes_for(r, 1, batch)));
  const size_t tb = al((size_t)batch * r->rd.LQ * r->rd.n * sizeof(uint64_t));
  const size_t db = al(dig_bytes_for(r, batch));
  if (yb + tb + db > (size_t)YH_STATIC_SCRATCH_BYTES)
    return fail(YH_ERR_ARG, "batch needs " + std::to_string(yb + tb + db) + " bytes of scratch, more than the " +
                                std::to_string((size_t)YH_STATIC_SCRATCH_BYTES) +
                                " of the static array; split the batch");
  if (int s = ring_enter(r)) return s;
  uint8_t* base = r->ctx->static_scratch;
  if (!base) return fail(YH_ERR_INTERNAL, "static scratch address not set up");
  Scratch sc;
  sc.y = (uint32_t*)base;
  sc.tmp = (uint64_t*)(base + yb);
  sc.dig = (uint64_t*)(base + yb + tb);
  t_no_alloc = true;
  int st = mul_keyswitch_impl(r, sc, k, d_c1, d_c2, d_out, batch);
  t_no_alloc = false;
  return st;
}

static int decrypt_impl(yh_ring* r, const uint64_t* d_f, const uint64_t* d_c, uint64_t* d_m, int batch,
                        const uint64_t* const* c_list, uint64_t* const* m_list) {
  const int n = r->rd.n, LQ = r->rd.LQ;
  Scratch sc;
  int st = ring_y(r, y_bytes_for(r, 0, batch), &sc);
  if (st) return st;
  yhk::ProdArgs pa{};
  pa.A = d_f; pa.B = d_c;
  pa.lin_a = pa.lin_b = LQ;
  pa.ncoef_a = pa.ncoef_b = n;
  pa.lstride_a = pa.lstride_b = n;
  pa.ostride_a = 0;
  pa.ostride_b = (size_t)LQ * n;
  pa.Y = sc.y;
  pa.logN = r->rd.logn;
  pa.tab = r->rd.tab;
  pa.set = 0;
  pa.b_list = c_list;
  st = launch_prod(r, pa, n, batch, true);
  if (st) return st;
  yhk::IcrtArgs ia = icrt_args(sc.y, n, n, 0, yhk::EPI_DECRYPT, d_m, 1, n, n, 0);
  ia.out_list = m_list;
  ia.reverse = 1;  // k_prod wrote the rows from the first pair to the last
  return launch_icrt(r, ia, batch);
}

int yh_yashe_decrypt(yh_ring* r, const uint64_t* d_f, const uint64_t* d_c, uint64_t* d_m, int batch) {
  if (!r || !d_f || !d_c || !d_m || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (r->rd.t == 0) return fail(YH_ERR_ARG, "ring has no plaintext modulus t");
  if (int s = ring_enter(r)) return s;
  return decrypt_impl(r, d_f, d_c, d_m, batch, nullptr, nullptr);
}

// ---------------------------------------------------------------- pointer-list variants
int yh_ring_addsub_list(yh_ring* r, const uint64_t* const* d_a, const uint64_t* const* d_b, uint64_t* const* d_out,
                        int batch, int op) {
  if (!r || !d_a || !d_b || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (int s = ring_enter(r)) return s;
  return addsub_impl(r, nullptr, nullptr, nullptr, batch, op == YH_SUB ? 1 : 0, yhk::AddsubLists{d_a, d_b, d_out});
}

int yh_yashe_mul_keyswitch_list(yh_ring* r, const yh_evk* k, const uint64_t* const* d_c1,
                                const uint64_t* const* d_c2, uint64_t* const* d_out, int batch) {
  if (!r || !k || !d_c1 || !d_c2 || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (k->ring != r) return fail(YH_ERR_ARG, "key belongs to another ring");
  if (r->rd.t == 0) return fail(YH_ERR_ARG, "ring has no plaintext modulus t");
  if (int s = ring_enter(r)) return s;
  Scratch sc;
  if (int st = mulks_scratch(r, batch, &sc)) return st;
  return mul_keyswitch_impl(r, sc, k, nullptr, nullptr, nullptr, batch, d_c1, d_c2, d_out);
}

int yh_yashe_decrypt_list(yh_ring* r, const uint64_t* d_f, const uint64_t* const* d_c, uint64_t* const* d_m,
                          int batch) {
  if (!r || !d_f || !d_c || !d_m || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (r->rd.t == 0) return fail(YH_ERR_ARG, "ring has no plaintext modulus t");
  if (int s = ring_enter(r)) return s;
  return decrypt_impl(r, d_f, nullptr, nullptr, batch, d_c, d_m);
}

// ---------------------------------------------------------------- streams and events
int yh_ring_use_own_stream(yh_ring* r) {
  if (!r) return fail(YH_ERR_ARG, "null ring");
  if (r->own_stream) return YH_OK;
  if (r->ctx->capturing) return fail(YH_ERR_ARG, "cannot change streams during graph capture");
  if (int s = set_dev(r->ctx)) return s;
  YH_HIP(hipStreamSynchronize(r->stream));
  hipStream_t st;
  YH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  r->stream = st;
  r->own_stream = true;
  return YH_OK;
}

int yh_ring_sync(yh_ring* r) {
  if (!r) return fail(YH_ERR_ARG, "null ring");
  if (r->ctx->capturing && !r->own_stream) return fail(YH_ERR_ARG, "synchronize during graph capture");
  if (int s = set_dev(r->ctx)) return s;
  YH_HIP(hipStreamSynchronize(r->stream));
  return YH_OK;
}

int yh_event_create(yh_ctx* ctx, yh_event** out) {
  if (!ctx || !out) return fail(YH_ERR_ARG, "null argument");
  if (int s = set_dev(ctx)) return s;
  auto e = std::make_unique<yh_event>();
  e->device = ctx->device;
  YH_HIP(hipEventCreate(&e->ev));
  *out = e.release();
  return YH_OK;
}

int yh_event_record(yh_event* e, yh_ctx* ctx, yh_ring* ring) {
  if (!e || (!ctx && !ring)) return fail(YH_ERR_ARG, "null argument");
  YH_HIP(hipSetDevice(e->device));
  YH_HIP(hipEventRecord(e->ev, ring ? ring->stream : ctx->stream));
  return YH_OK;
}

int yh_event_query(yh_event* e) {
  if (!e) return fail(YH_ERR_ARG, "null event");
  hipError_t r = hipEventQuery(e->ev);
  if (r == hipSuccess) return 1;
  if (r == hipErrorNotReady) return 0;
  (void)hipGetLastError();
  return fail(YH_ERR_HIP, hipGetErrorString(r));
}

int yh_event_sync(yh_event* e) {
  if (!e) return fail(YH_ERR_ARG, "null event");
  YH_HIP(hipEventSynchronize(e->ev));
  return YH_OK;
}

int yh_event_elapsed_ms(yh_event* start, yh_event* stop, float* ms) {
  if (!start || !stop || !ms) return fail(YH_ERR_ARG, "null argument");
  YH_HIP(hipEventElapsedTime(ms, start->ev, stop->ev));
  return YH_OK;
}

int yh_event_destroy(yh_event* e) {
  if (!e) return YH_OK;
  (void)hipSetDevice(e->device);
  (void)hipEventDestroy(e->ev);
  delete e;
  return YH_OK;
}

int yh_ring_mul_bcast(yh_ring* r, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out, int batch) {
  if (!r || !d_a || !d_b || !d_out || batch <= 0) return fail(YH_ERR_ARG, "bad argument");
  if (int s = ring_enter(r)) return s;
  const int n = r->rd.n, LQ = r->rd.LQ;
  Scratch sc;
  int st = ring_y(r, y_bytes_for(r, 0, batch), &sc);
  if (st) return st;
  yhk::ProdArgs pa{};
  pa.A = d_a; pa.B = d_b;
  pa.lin_a = pa.lin_b = LQ;
  pa.ncoef_a = pa.ncoef_b = n;
  pa.lstride_a = pa.lstride_b = n;
  pa.ostride_a = 0;
  pa.ostride_b = (size_t)LQ * n;
  pa.Y = sc.y;
  pa.logN = r->rd.logn;
  pa.tab = r->rd.tab;
  pa.set = 0;
  st = launch_prod(r, pa, n, batch, true);
  if (st) return st;
  return launch_icrt(r, icrt_args(sc.y, n, n, 0, yhk::EPI_MODQ, d_out, LQ, n, (size_t)LQ * n, 0), batch);
}

static int sub_ring(yh_ring* r, int m, yh_ring** out) {
  if (m == r->rd.n) { *out = r; return YH_OK; }
  auto it = r->sub.find(m);
  if (it != r->sub.end()) {
    it->second->stream = r->stream;  // same stream as the parent's work
    *out = it->second;
    return YH_OK;
  }
  std::vector<uint64_t> ql(r->P->q_limbs);
  yh_ring* s = nullptr;
  int st = yh_ring_create(r->ctx, m, ql.data(), (int)ql.size(), 0, 0, &s);
  if (st) return st;
  s->stream = r->stream;
  r->sub[m] = s;
  *out = s;
  return YH_OK;
}


int yh_ring_inverse(yh_ring* r, const uint64_t* d_f, uint64_t* d_out) {
  if (!r || !d_f || !d_out) return fail(YH_ERR_ARG, "null argument");
  if (r->ctx->capturing) return fail(YH_ERR_ARG, "yh_ring_inverse cannot be captured");
  if (int s = set_dev(r->ctx)) return s;
  YH_TRY_BEGIN
  const int n = r->rd.n, LQ = r->rd.LQ;
  const size_t elt = (size_t)LQ * n * sizeof(uint64_t);
  // level buffers: F_l (size n_l), G_l = F_l(-x); one allocation
  std::vector<int> sizes;
  for (int m = n; m >= 2; m /= 2) sizes.push_back(m);
  uint64_t* pool = nullptr;
  YH_HIP(hipMalloc(&pool, 6 * elt));
  std::unique_ptr<uint64_t, void (*)(uint64_t*)> guard(pool, [](uint64_t* p) { (void)hipFree(p); });
  uint64_t* F = pool;                // current F_l
  uint64_t* Gs = pool + 1 * (elt / 8);  // G_l for every level, packed (total < 2 elt)
  uint64_t* T = pool + 3 * (elt / 8);   // product scratch
  uint64_t* U = pool + 4 * (elt / 8);   // inverse of the level below, spread
  YH_HIP(hipMemcpyAsync(F, d_f, elt, hipMemcpyDeviceToDevice, r->stream));
  std::vector<size_t> goff;
  size_t off = 0;
  for (int m : sizes) {
    yh_ring* s = nullptr;
    if (int st = sub_ring(r, m, &s)) return st;
    uint64_t* G = Gs + off;
    goff.push_back(off);
    off += (size_t)LQ * m;
    YH_LAUNCH(yhl::negate_odd(LQ, dim3((m + 255) / 256), dim3(256), s->stream, s->rd, F, G, m));
    if (int st = yh_ring_mul(s, F, G, T, 1)) return st;  // N(F)(x^2)
    YH_LAUNCH(yhl::take_even(dim3((m / 2 + 255) / 256), dim3(256), r->stream, T, F, m, LQ));
  }
  // F now holds the scalar N(...)(y) in Z_q[y]/(y + 1): invert it on the host
  std::vector<uint64_t> h(LQ);
  YH_HIP(hipMemcpyAsync(h.data(), F, LQ * sizeof(uint64_t), hipMemcpyDeviceToHost, r->stream));
  YH_HIP(hipStreamSynchronize(r->stream));
  yh::BigInt c0 = yh::BigInt::from_limbs(h.data(), LQ);
  if (c0.is_zero()) return fail(YH_ERR_ARG, "not invertible");
  yh::BigInt ci;
  try {
    ci = yh::BigInt::inv_mod(c0, r->P->q);
  } catch (const std::domain_error&) {
    return fail(YH_ERR_ARG, "not invertible");
  }
  ci.to_limbs(h.data(), LQ);
  YH_HIP(hipMemcpyAsync(F, h.data(), LQ * sizeof(uint64_t), hipMemcpyHostToDevice, r->stream));
  // back up: Inv_l = G_l * Inv_{l+1}(x^2)
  for (int k = (int)sizes.size() - 1; k >= 0; --k) {
    const int m = sizes[k];
    yh_ring* s = nullptr;
    if (int st = sub_ring(r, m, &s)) return st;
    YH_LAUNCH(yhl::spread_even(dim3((m + 255) / 256), dim3(256), r->stream, F, U, m, LQ));
    uint64_t* dst = k == 0 ? d_out : F;
    if (int st = yh_ring_mul(s, Gs + goff[k], U, dst, 1)) return st;
  }
  YH_HIP(hipStreamSynchronize(r->stream));
  return YH_OK;
  YH_TRY_END
}

struct yh_batch {
  yh_ctx* ctx = nullptr;
  yh_batch_config cfg{};
  int ninst = 0, nbundles = 0, grid = 0, W = 0, maxlq = 0;
  void* d_mem = nullptr;
  yhk::BatchArgs args{};
  std::vector<yh_ring*> rings;  // distinct rings of the items (each holds a reference while the plan lives)
  int users = 0;                // captured graphs that launch this plan
  ~yh_batch() {
    if (d_mem) (void)hipFree(d_mem);
    for (yh_ring* r : rings) r->plan_users--;
  }
};

static void batch_users(yh_batch* b, int d) { b->users += d; }

int yh_graph_begin(yh_ctx* ctx) {
  if (!ctx) return fail(YH_ERR_ARG, "null ctx");
  if (ctx->capturing) return fail(YH_ERR_ARG, "already capturing");
  if (int s = set_dev(ctx)) return s;
  if (int s = drain_timing(ctx)) return s;
  YH_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  ctx->cap_rings.clear();
  ctx->cap_evks.clear();
  ctx->cap_batches.clear();
  return YH_OK;
}

int yh_graph_end(yh_ctx* ctx, yh_graph** out) {
  if (!ctx || !out) return fail(YH_ERR_ARG, "null argument");
  if (!ctx->capturing) return fail(YH_ERR_ARG, "not capturing");
  ctx->capturing = false;
  auto g = std::make_unique<yh_graph>();
  g->ctx = ctx;
  g->rings.swap(ctx->cap_rings);
  g->evks.swap(ctx->cap_evks);
  g->batches.swap(ctx->cap_batches);
  YH_HIP(hipStreamEndCapture(ctx->stream, &g->graph));
  YH_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  // the graph's launches hold these objects' device addresses from now on
  for (yh_ring* r : g->rings) r->graph_users++;
  for (const yh_evk* k : g->evks) k->users++;
  for (yh_batch* b : g->batches) batch_users(b, +1);
  ctx->live++;
  *out = g.release();
  return YH_OK;
}

int yh_graph_launch(yh_graph* g) {
  if (!g) return fail(YH_ERR_ARG, "null graph");
  if (int s = set_dev(g->ctx)) return s;
  YH_HIP(hipGraphLaunch(g->exec, g->ctx->stream));
  return YH_OK;
}

int yh_graph_destroy(yh_graph* g) {
  if (!g) return YH_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  for (yh_ring* r : g->rings) r->graph_users--;
  for (const yh_evk* k : g->evks) k->users--;
  for (yh_batch* b : g->batches) batch_users(b, -1);
  g->ctx->live--;
  delete g;
  return YH_OK;
}



extern "C++" {
namespace {

int validate_items(const yh_batch_item* items, int count, int* max_T) {
  if (!items || count <= 0) return fail(YH_ERR_ARG, "empty batch");
  int mt = 1;
  for (int k = 0; k < count; ++k) {
    const yh_batch_item& it = items[k];
    if (!it.ring || !it.a || !it.b || !it.out) return fail(YH_ERR_ARG, "null pointer in batch item");
    const int ln = it.ring->rd.logn;
    if (ln < yhk::BATCH_MIN_LOGN || ln > yhk::BATCH_MAX_LOGN)
      return fail(YH_ERR_ARG, "batched products take 16 <= n <= 16384");
    if (it.ring->ctx != items[0].ring->ctx) return fail(YH_ERR_ARG, "batch items from different contexts");
    mt = std::max(mt, (1 << ln) / 16);
  }
  *max_T = mt;
  return YH_OK;
}

int check_config(const yh_batch_config& c, int max_T) {
  const int W = c.wg_threads;
  if (W < 64 || W > yhk::BATCH_MAX_THREADS || (W & (W - 1)) || W < max_T)
    return fail(YH_ERR_ARG, "wg_threads must be a power of two in [max(64, n/16), 1024]");
  if (c.wgs_per_cu < 1 || c.wgs_per_cu > 16) return fail(YH_ERR_ARG, "wgs_per_cu out of range");
  if (c.queues < 1 || c.queues > 64) return fail(YH_ERR_ARG, "queues out of range");
  if (c.waves < 2 || c.waves > 4) return fail(YH_ERR_ARG, "waves must be 2, 3 or 4");
  if (W > yhk::BATCH_TUNED_THREADS && c.waves != 4)
    return fail(YH_ERR_ARG, "1024-thread workgroups run four waves per SIMD (waves = 4)");
  return YH_OK;
}

// Key of a batch composition: device plus the histogram of (log2 n, limbs).
std::string batch_signature(const yh_ctx* ctx, const yh_batch_item* items, int count) {
  std::map<std::pair<int, int>, int> h;
  for (int k = 0; k < count; ++k) h[{items[k].ring->rd.logn, items[k].ring->rd.LQ}]++;
  std::ostringstream o;
  o << ctx->arch << ":cu" << ctx->cus << "|";
  bool first = true;
  for (auto& kv : h) {
    o << (first ? "" : ",") << "n" << (1 << kv.first.first) << "L" << kv.first.second << "x" << kv.second;
    first = false;
  }
  return o.str();
}

bool cache_lookup(const char* path, const std::string& key, yh_batch_config* out) {
  if (!path) return false;
  std::ifstream f(path);
  if (!f) return false;
  std::string line, want = "\"" + key + "\":";
  while (std::getline(f, line)) {
    size_t pos = line.find(want);
    if (pos == std::string::npos) continue;
    yh_batch_config c{};
    double ms = 0;
    if (sscanf(line.c_str() + pos + want.size(),
               " {\"wg_threads\": %d, \"wgs_per_cu\": %d, \"queues\": %d, \"waves\": %d, \"ms\": %lf}",
               &c.wg_threads, &c.wgs_per_cu, &c.queues, &c.waves, &ms) == 5) {
      *out = c;
      return true;
    }
  }
  return false;
}

// The cache is a JSON object with one entry per line; new entries are added
// before the closing brace (the file is rewritten whole).
void cache_store(const char* path, const std::string& key, const yh_batch_config& c, double ms) {
  if (!path) return;
  std::vector<std::string> entries;
  {
    std::ifstream f(path);
    std::string line;
    while (f && std::getline(f, line)) {
      if (line.size() > 2 && line[0] == ' ' && line.find("\"" + key + "\":") == std::string::npos) {
        while (!line.empty() && (line.back() == ',' || line.back() == ' ')) line.pop_back();
        entries.push_back(line);
      }
    }
  }
  char buf[256];
  snprintf(buf, sizeof(buf), "{\"wg_threads\": %d, \"wgs_per_cu\": %d, \"queues\": %d, \"waves\": %d, \"ms\": %.4f}",
           c.wg_threads, c.wgs_per_cu, c.queues, c.waves, ms);
  entries.push_back("  \"" + key + "\": " + buf);
  std::string tmp = std::string(path) + ".tmp";
  {
    std::ofstream f(tmp);
    f << "{\n";
    for (size_t i = 0; i < entries.size(); ++i) f << entries[i] << (i + 1 < entries.size() ? ",\n" : "\n");
    f << "}\n";
  }
  std::rename(tmp.c_str(), path);
}

}  // namespace
}  // extern "C++"

int yh_batch_default_config(const yh_batch_item* items, int count, yh_batch_config* cfg) {
  if (!cfg) return fail(YH_ERR_ARG, "null cfg");
  int mt = 0;
  if (int s = validate_items(items, count, &mt)) return s;
  cfg->wg_threads = std::max(256, mt);
  cfg->wgs_per_cu = cfg->wg_threads > yhk::BATCH_TUNED_THREADS ? 1 : 2;
  cfg->queues = yhk::XCDS;
  int maxlq = 0;
  for (int k = 0; k < count; ++k) maxlq = std::max(maxlq, items[k].ring->rd.LQ);
  cfg->waves = maxlq > 4 && cfg->wg_threads <= yhk::BATCH_TUNED_THREADS ? 2 : 4;
  return YH_OK;
}

int yh_batch_create(yh_ctx* ctx, const yh_batch_item* items, int count, const yh_batch_config* cfg_in, yh_batch** out) {
  if (!ctx || !out) return fail(YH_ERR_ARG, "null argument");
  if (ctx->capturing) return fail(YH_ERR_ARG, "yh_batch_create allocates; plan before capturing");
  int mt = 0;
  if (int s = validate_items(items, count, &mt)) return s;
  if (items[0].ring->ctx != ctx) return fail(YH_ERR_ARG, "batch items from another context");
  yh_batch_config cfg;
  if (cfg_in) cfg = *cfg_in;
  else yh_batch_default_config(items, count, &cfg);
  if (int s = check_config(cfg, mt)) return s;
  if (int s = set_dev(ctx)) return s;
  YH_TRY_BEGIN
  const int W = cfg.wg_threads;
  const int nq = cfg.queues;
  // distinct rings
  std::vector<yh_ring*> rings;
  std::map<yh_ring*, int> ridx;
  for (int k = 0; k < count; ++k)
    if (!ridx.count(items[k].ring)) { ridx[items[k].ring] = (int)rings.size(); rings.push_back(items[k].ring); }
  // products to queues: largest first, each to the least loaded queue
  std::vector<int> order(count);
  for (int k = 0; k < count; ++k) order[k] = k;
  auto cost = [&](int k) {
    const yh_ring* r = items[k].ring;
    return (double)r->rd.R * (double)r->rd.n * (r->rd.logn + 2);
  };
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost(x) > cost(y); });
  std::vector<double> load(nq, 0.0);
  std::vector<std::vector<int>> qinst(nq);
  for (int k : order) {
    int best = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[best] += cost(k);
    qinst[best].push_back(k);
  }
  // tasks and bundles, queue by queue; inside a queue larger rings first and
  // the primes of one product adjacent
  std::vector<yhk::BatchTask> tasks;
  std::vector<yhk::BatchBundle> bundles;
  std::vector<int> qstart(nq + 1, 0);
  // rings from the largest n down; a bundle holds tasks of one ring
  std::vector<int> ring_order(rings.size());
  for (size_t i = 0; i < rings.size(); ++i) ring_order[i] = (int)i;
  std::stable_sort(ring_order.begin(), ring_order.end(), [&](int x, int y) {
    return rings[x]->rd.logn != rings[y]->rd.logn ? rings[x]->rd.logn > rings[y]->rd.logn
                                                  : rings[x]->rd.R > rings[y]->rd.R;
  });
  for (int q = 0; q < nq; ++q) {
    qstart[q] = (int)bundles.size();
    for (int ri : ring_order) {
      const int ln = rings[ri]->rd.logn;
      const int cap = W / ((1 << ln) / 16);
      // tasks of modulus 0 (P, a different residue path) get bundles of
      // their own, so every bundle takes one path uniformly
      for (int gold = 1; gold >= 0; --gold) {
        int open = -1;
        for (int k : qinst[q]) {
          if (items[k].ring != rings[ri]) continue;
          for (int i = gold ? 0 : 1; i < (gold ? 1 : rings[ri]->rd.R); ++i) {
            if (open < 0 || bundles[open].count == cap) {
              bundles.push_back({ln, 0, (int)tasks.size(), ri, gold});
              open = (int)bundles.size() - 1;
            }
            tasks.push_back({k, i});
            bundles[open].count++;
          }
        }
      }
    }
  }
  qstart[nq] = (int)bundles.size();
  // one device allocation: rings | inst | tasks | bundles | qstart | head | done | Y
  auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
  size_t off = 0;
  const size_t o_rings = off; off = align(off + rings.size() * sizeof(yhk::RingDev));
  const size_t o_inst = off; off = align(off + count * sizeof(yhk::BatchInst));
  const size_t o_tasks = off; off = align(off + tasks.size() * sizeof(yhk::BatchTask));
  const size_t o_bund = off; off = align(off + bundles.size() * sizeof(yhk::BatchBundle));
  const size_t o_qs = off; off = align(off + qstart.size() * sizeof(int));
  const size_t o_head = off; off = align(off + nq * sizeof(int));
  const size_t o_done = off; off = align(off + count * sizeof(int));
  const size_t host_bytes = o_head;
  std::vector<size_t> o_y(count);
  for (int k = 0; k < count; ++k) {
    o_y[k] = off;
    off = align(off + (size_t)yhk::y_rows(items[k].ring->rd.R) * items[k].ring->rd.n * sizeof(uint32_t));
  }
  auto b = std::make_unique<yh_batch>();
  b->ctx = ctx;
  b->cfg = cfg;
  b->W = W;
  b->ninst = count;
  b->nbundles = (int)bundles.size();
  for (int k = 0; k < count; ++k) b->maxlq = std::max(b->maxlq, items[k].ring->rd.LQ);
  YH_HIP(hipMalloc(&b->d_mem, off));
  char* base = (char*)b->d_mem;
  std::vector<char> h(host_bytes, 0);
  for (size_t i = 0; i < rings.size(); ++i)
    memcpy(h.data() + o_rings + i * sizeof(yhk::RingDev), &rings[i]->rd, sizeof(yhk::RingDev));
  for (int k = 0; k < count; ++k) {
    yhk::BatchInst in{items[k].a, items[k].b, items[k].out, (uint32_t*)(base + o_y[k]), ridx[items[k].ring], 0};
    (void)in.ring;
    memcpy(h.data() + o_inst + k * sizeof(yhk::BatchInst), &in, sizeof(in));
  }
  if (!tasks.empty()) memcpy(h.data() + o_tasks, tasks.data(), tasks.size() * sizeof(yhk::BatchTask));
  if (!bundles.empty()) memcpy(h.data() + o_bund, bundles.data(), bundles.size() * sizeof(yhk::BatchBundle));
  memcpy(h.data() + o_qs, qstart.data(), qstart.size() * sizeof(int));
  YH_HIP(hipMemcpy(base, h.data(), host_bytes, hipMemcpyHostToDevice));
  yhk::BatchArgs& a = b->args;
  a.rings = (const yhk::RingDev*)(base + o_rings);
  a.inst = (const yhk::BatchInst*)(base + o_inst);
  a.tasks = (const yhk::BatchTask*)(base + o_tasks);
  a.bundles = (const yhk::BatchBundle*)(base + o_bund);
  a.qstart = (const int*)(base + o_qs);
  a.head = (int*)(base + o_head);
  a.done = (int*)(base + o_done);
  a.nqueues = nq;
  b->grid = std::min(ctx->cus * cfg.wgs_per_cu, std::max(b->nbundles, 1));
  // the plan holds copies of each ring's table pointers
  b->rings = rings;
  for (yh_ring* r : rings) r->plan_users++;
  ctx->live++;
  *out = b.release();
  return YH_OK;
  YH_TRY_END
}

int yh_batch_run(yh_batch* b) {
  if (!b) return fail(YH_ERR_ARG, "null batch");
  if (int s = set_dev(b->ctx)) return s;
  hipStream_t st = b->ctx->stream;
  if (b->ctx->capturing &&
      std::find(b->ctx->cap_batches.begin(), b->ctx->cap_batches.end(), b) == b->ctx->cap_batches.end())
    b->ctx->cap_batches.push_back(b);
  YH_LAUNCH(yhl::batch_reset(dim3(std::max(1, std::min(64, (b->ninst + 255) / 256))), dim3(256), st, b->args.head,
                             b->args.nqueues, b->args.done, b->ninst));
  LaunchTimer timer(b->ctx, YH_KERNEL_BATCH, b->ctx->stream);
  YH_LAUNCH(yhl::batch_prod(b->maxlq, b->cfg.waves, dim3(b->grid), dim3(b->W), yhk::batch_lds_bytes(b->W), st,
                            b->args));
  return YH_OK;
}

int yh_batch_destroy(yh_batch* b) {
  if (!b) return YH_OK;
  if (b->users > 0)
    return fail(YH_ERR_ARG, "batch plan is launched by " + std::to_string(b->users) +
                                " captured graph(s); destroy them first");
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  b->ctx->live--;
  delete b;
  return YH_OK;
}

int yh_batch_autotune(yh_ctx* ctx, const yh_batch_item* items, int count, const char* cache_path, int reps,
                      yh_batch_config* out) {
  if (!ctx || !out) return fail(YH_ERR_ARG, "null argument");
  if (ctx->capturing) return fail(YH_ERR_ARG, "autotune during graph capture");
  int mt = 0;
  if (int s = validate_items(items, count, &mt)) return s;
  if (int s = set_dev(ctx)) return s;
  const std::string key = batch_signature(ctx, items, count);
  if (cache_lookup(cache_path, key, out) && check_config(*out, mt) == YH_OK) return YH_OK;
  reps = std::max(1, reps);
  if (int s = drain_timing(ctx)) return s;
  const bool was_timing = ctx->timing;
  ctx->timing = false;
  hipEvent_t e0, e1;
  YH_HIP(hipEventCreate(&e0));
  YH_HIP(hipEventCreate(&e1));
  double best_ms = 1e30;
  yh_batch_config best{};
  int rc = YH_OK;
  int maxlq = 0;
  for (int k = 0; k < count; ++k) maxlq = std::max(maxlq, items[k].ring->rd.LQ);
  std::vector<yh_batch_config> cand;
  for (int W = std::max(64, mt); W <= yhk::BATCH_MAX_THREADS; W *= 2)
    for (int waves : {2, 3, 4}) {
      // the kernels that exist
      if (W > yhk::BATCH_TUNED_THREADS ? waves != 4 : (maxlq > 4) != (waves == 2)) continue;
      for (int wgs : {1, 2, 3, 4}) {
        if ((W / 64) * wgs > 4 * waves) continue;  // more workgroups than can be resident
        for (int nq : {1, yhk::XCDS}) cand.push_back({W, wgs, nq, waves});
      }
    }
  for (const yh_batch_config& c : cand) {
    yh_batch* b = nullptr;
    if ((rc = yh_batch_create(ctx, items, count, &c, &b))) break;
    std::vector<float> t;
    rc = yh_batch_run(b);  // warm-up
    for (int r = 0; r < reps && rc == YH_OK; ++r) {
      if (hipEventRecord(e0, ctx->stream) != hipSuccess) { rc = fail(YH_ERR_HIP, "event record"); break; }
      if ((rc = yh_batch_run(b))) break;
      if (hipEventRecord(e1, ctx->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) {
        rc = fail(YH_ERR_HIP, "event sync");
        break;
      }
      float ms = 0;
      (void)hipEventElapsedTime(&ms, e0, e1);
      t.push_back(ms);
    }
    yh_batch_destroy(b);
    if (rc) break;
    std::sort(t.begin(), t.end());
    const double med = t[t.size() / 2];
    if (med < best_ms) { best_ms = med; best = c; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->timing = was_timing;
  if (rc) return rc;
  *out = best;
  cache_store(cache_path, key, best, best_ms);
  return YH_OK;
}

static int wire_params(int n, const uint64_t* q, int q_limbs, uint64_t t, int w_bits, yhw::Params* p) {
  if (!q || q_limbs < 1 || q_limbs > yhw::MAX_LIMBS) return fail(YH_ERR_ARG, "q must have 1 to 8 limbs");
  p->n = n;
  p->L = q_limbs;
  for (int l = 0; l < q_limbs; ++l) p->q[l] = q[l];
  p->t = t;
  p->w_bits = w_bits;
  std::string e = yhw::check_params(*p);
  if (!e.empty()) return fail(YH_ERR_ARG, e.c_str());
  return YH_OK;
}

size_t yh_wire_size(int n, int q_limbs, int count) {
  if (n < 1 || q_limbs < 1 || count < 0) return 0;
  return yhw::encoded_size(n, q_limbs, count);
}

int yh_wire_encode(int n, const uint64_t* q, int q_limbs, uint64_t t, int w_bits, int kind, const uint64_t* h_src,
                   int count, void* out, size_t cap, size_t* len) {
  yhw::Params p;
  if (int st = wire_params(n, q, q_limbs, t, w_bits, &p)) return st;
  if (out && count > 0 && !h_src) return fail(YH_ERR_ARG, "null source");
  std::string e = yhw::encode(p, kind, kind == YH_WIRE_PLAINTEXT, h_src, count, (uint8_t*)out, cap, len);
  if (!e.empty()) return fail(YH_ERR_ARG, e.c_str());
  return YH_OK;
}

int yh_wire_decode(int n, const uint64_t* q, int q_limbs, uint64_t t, int w_bits, int kind, const void* in,
                   size_t len, uint64_t* h_dst, int cap, int* count) {
  yhw::Params p;
  if (int st = wire_params(n, q, q_limbs, t, w_bits, &p)) return st;
  std::string e = yhw::decode(p, kind, kind == YH_WIRE_PLAINTEXT, (const uint8_t*)in, len, h_dst, cap, count);
  if (!e.empty()) return fail(YH_ERR_ARG, e.c_str());
  return YH_OK;
}

int yh_wire_peek(const void* in, size_t len, yh_wire_header* hdr) {
  if (!hdr) return fail(YH_ERR_ARG, "null header");
  yhw::Header h;
  std::string e = yhw::peek((const uint8_t*)in, len, &h);
  if (!e.empty()) return fail(YH_ERR_ARG, e.c_str());
  hdr->version = h.version;
  hdr->kind = h.kind;
  hdr->n = h.p.n;
  hdr->q_limbs = h.p.L;
  hdr->t = h.p.t;
  hdr->w_bits = h.p.w_bits;
  hdr->count = h.count;
  for (int l = 0; l < 8; ++l) hdr->q[l] = l < h.p.L ? h.p.q[l] : 0;
  return YH_OK;
}

size_t yh_mem_footprint(const yh_ring* r, int op) {
  if (!r) return 0;
  size_t n = r->rd.n, LQ = r->rd.LQ, R = r->rd.R, Rks = r->rd.set_R[1];
  size_t ct = LQ * n * sizeof(uint64_t);
  size_t y = yhk::y_rows((int)R) * n * sizeof(uint32_t), yks = yhk::y_rows((int)Rks) * n * sizeof(uint32_t);
  // keyswitch: reconstruction rows plus the digit transforms
  yks += (size_t)r->rd.lw * n * sizeof(uint64_t);
  switch (op) {
    case 0: return 3 * ct + y;
    case 1: return 3 * ct + y;
    case 2: return 2 * ct + yks;
    case 3: return 4 * ct + (y > yks ? y : yks);
    case 4: return 3 * ct;
    default: return 0;
  }
}

}  // extern "C"

// ---------------------------------------------------------------- internals for other translation units
namespace yhi {
int fail(int code, const std::string& msg) { return ::fail(code, msg); }
int ctx_device(const yh_ctx* ctx) { return ctx->device; }
hipStream_t ctx_stream(const yh_ctx* ctx) { return ctx->stream; }
hipStream_t ring_stream(const yh_ring* ring) { return ring->stream; }
const uint64_t* ring_q(const yh_ring* ring, int* limbs) {
  *limbs = ring->P->LQ;
  return ring->P->q_limbs.data();
}
yh_ctx* ring_ctx(const yh_ring* ring) { return ring->ctx; }
const yh_ring* evk_ring(const yh_evk* evk) { return evk->ring; }
void hold_ring(yh_ring* ring, int delta) { ring->server_users += delta; }
void hold_evk(const yh_evk* evk, int delta) { evk->server_users += delta; }
void hold_ctx(yh_ctx* ctx, int delta) { ctx->live += delta; }
}  // namespace yhi
