// Bodies of the reference's device-layer entry points (declared in
// reference_shim.h) over the gfx950 C ABI. Each body states the reference
// function it replaces; failures abort with the library's message, as the
// reference asserts.
#include "reference_shim.h"

#include <cstdio>
#include <cstdlib>

namespace {

yh_ctx* g_ctx = nullptr;
yh_ring* g_ring = nullptr;
std::vector<uint64_t> g_primes;

void ok(int status, const char* what) {
  if (status != YH_OK) {
    std::fprintf(stderr, "%s: %s\n", what, yh_last_error());
    std::abort();
  }
}

int log2_exact(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

}  // namespace

int CUDAFunctions::N = 0;
int CUDAFunctions::std_bn_t_alloc = 32;  // STD_BNT_WORDS_ALLOC, src/settings.h:30
cuyasheint_t* CUDAFunctions::d_inner_results = nullptr;
cuyasheint_t* CUDAFunctions::d_inner_results_used = nullptr;
cuyasheint_t CUDAFunctions::wN = 0;
cuyasheint_t* CUDAFunctions::d_W = nullptr;
cuyasheint_t* CUDAFunctions::d_WInv = nullptr;

namespace reference_shim {

yh_ctx* context() {
  if (!g_ctx) ok(yh_ctx_create(0, &g_ctx), "yh_ctx_create");
  return g_ctx;
}

yh_ring* ring() {
  if (!g_ring) {
    std::fprintf(stderr, "reference shim: gen_crt_primes was not called\n");
    std::abort();
  }
  return g_ring;
}

void gen_crt_primes(const uint64_t* q_limbs, int q_words, int degree) {
  int n = 2;
  while (n < degree) n <<= 1;
  if (g_ring) ok(yh_ring_destroy(g_ring), "yh_ring_destroy");
  g_ring = nullptr;
  ok(yh_ring_create(context(), n, q_limbs, q_words, 0, 0, &g_ring), "yh_ring_create");
  yh_ring_info info;
  ok(yh_ring_get_info(g_ring, &info), "yh_ring_get_info");
  g_primes.assign(info.R, 0);
  ok(yh_ring_get_primes(g_ring, g_primes.data(), info.R), "yh_ring_get_primes");
}

const std::vector<uint64_t>& crt_primes() { return g_primes; }

void release() {
  if (g_ring) yh_ring_destroy(g_ring);
  g_ring = nullptr;
  if (g_ctx) yh_ctx_destroy(g_ctx);
  g_ctx = nullptr;
}

}  // namespace reference_shim

using reference_shim::context;
using reference_shim::ring;

// CUDAFunctions::init (src/cuda_functions.cu:878-956): the reference builds
// the twiddle tables of length N and the ICRT scratch; the library builds and
// caches its tables per length and ring on first use.
void CUDAFunctions::init(int n) {
  if (n <= 0 || (n & (n - 1))) {
    std::fprintf(stderr, "CUDAFunctions::init: N must be a power of two\n");
    std::abort();
  }
  CUDAFunctions::N = n;
  (void)context();
}

// write_crt_primes (src/cuda_bn.cu:1047-1147): the ring's tables were written
// when reference_shim::gen_crt_primes created it.
void CUDAFunctions::write_crt_primes() { (void)ring(); }

// callPolynomialMul (src/cuda_functions.cu:617-738): per residue, forward
// transforms of both operands, pointwise product, inverse transform scaled by
// N^-1 (the reference's NTTScale divides by N); the cyclic convolution mod
// x^N - 1 of each residue row, computed mod P. Returns a new device buffer of
// N * NPolis words, synchronised, as the reference does. The operands are
// expected aligned at spacing N (realign_* are unused in the reference's NTT
// path too).
cuyasheint_t* CUDAFunctions::callPolynomialMul(cudaStream_t, cuyasheint_t* a, const bool, const int,
                                               cuyasheint_t* b, const bool, const int, const int N,
                                               const int NPolis) {
  if (N <= 0 || (N & (N - 1)) || N != CUDAFunctions::N) {
    std::fprintf(stderr, "callPolynomialMul: N must be a power of two equal to CUDAFunctions::N\n");
    std::abort();
  }
  yh_ctx* ctx = context();
  const size_t size = (size_t)N * NPolis;
  void* mem = nullptr;
  ok(yh_malloc(ctx, 2 * size * sizeof(cuyasheint_t), &mem), "yh_malloc");
  cuyasheint_t* d_result = (cuyasheint_t*)mem;
  cuyasheint_t* d_b = d_result + size;
  ok(yh_memcpy_dtod(ctx, d_result, a, size * sizeof(cuyasheint_t)), "yh_memcpy_dtod");
  ok(yh_memcpy_dtod(ctx, d_b, b, size * sizeof(cuyasheint_t)), "yh_memcpy_dtod");
  ok(yh_ntt(ctx, d_result, N, NPolis, YH_FORWARD), "yh_ntt");
  ok(yh_ntt(ctx, d_b, N, NPolis, YH_FORWARD), "yh_ntt");
  ok(yh_pointwise_mulmod(ctx, d_result, d_b, size), "yh_pointwise_mulmod");
  ok(yh_ntt(ctx, d_result, N, NPolis, YH_INVERSE), "yh_ntt");
  ok(yh_ctx_sync(ctx), "yh_ctx_sync");
  return d_result;  // d_b, behind it in the same allocation, is scratch
}

// callPolynomialAddSub (src/cuda_functions.cu:115-126): c = a +- b on
// residues [NPolis][size / NPolis], reduced mod each prime (the reference
// omits the reduction).
void CUDAFunctions::callPolynomialAddSub(cuyasheint_t* c, cuyasheint_t* a, cuyasheint_t* b, int size, int OP,
                                         cudaStream_t) {
  yh_ring_info info;
  ok(yh_ring_get_info(ring(), &info), "yh_ring_get_info");
  ok(yh_residue_addsub(ring(), c, a, b, size / info.R, OP == SUB ? YH_SUB : YH_ADD), "yh_residue_addsub");
}

// callPolynomialAddSubInPlace (src/cuda_functions.cu:128-165): a = a +- b.
void CUDAFunctions::callPolynomialAddSubInPlace(cudaStream_t stream, cuyasheint_t* a, cuyasheint_t* b, int size,
                                                int OP) {
  callPolynomialAddSub(a, a, b, size, OP, stream);
}

// callPolynomialOPInteger (src/cuda_functions.cu:477-548): ADD / SUB the
// scalar's residue on coefficient 0, MUL every residue; the scalar is given
// by its residues integer_array[NPolis] on the device. Returns a new buffer.
cuyasheint_t* CUDAFunctions::callPolynomialOPInteger(const int opcode, cudaStream_t, cuyasheint_t* a,
                                                     cuyasheint_t* integer_array, const int N, const int NPolis) {
  void* out = nullptr;
  ok(yh_malloc(context(), (size_t)N * NPolis * sizeof(cuyasheint_t), &out), "yh_malloc");
  const int op = opcode == MUL ? YH_MUL : opcode == SUB ? YH_SUB : YH_ADD;
  ok(yh_residue_scalar(ring(), (uint64_t*)out, a, N, op, integer_array), "yh_residue_scalar");
  return (cuyasheint_t*)out;
}

// callNTT (src/cuda_functions.cu:447-458): NPolis transforms of length N;
// the result lands where the reference's ping-pong leaves it: in dataI when
// log2(N) is even, in dataO when it is odd.
void CUDAFunctions::callNTT(const int N, const int NPolis, cuyasheint_t* dataI, cuyasheint_t* dataO, const int type) {
  yh_ctx* ctx = context();
  ok(yh_ntt(ctx, dataI, N, NPolis, type == FORWARD ? YH_FORWARD : YH_INVERSE), "yh_ntt");
  if (log2_exact(N) & 1)
    ok(yh_memcpy_dtod(ctx, dataO, dataI, (size_t)N * NPolis * sizeof(cuyasheint_t)), "yh_memcpy_dtod");
}

// callCRT (src/cuda_bn.cu:970-1002): residues [NPolis][N] of the first
// used_coefs coefficients (bn_t on the device), zero beyond.
void callCRT(bn_t* coefs, const int used_coefs, cuyasheint_t* d_polyCRT, const int N, const int NPolis, cudaStream_t) {
  if (NPolis != (int)reference_shim::crt_primes().size()) {
    std::fprintf(stderr, "callCRT: NPolis differs from the prime set\n");
    std::abort();
  }
  ok(yh_crt_bn(ring(), coefs, used_coefs, N, d_polyCRT), "yh_crt_bn");
}

// callICRT (src/cuda_bn.cu:1004-1045): the N coefficients in [0, M) from
// the residues [NPolis][N], written into the bn_t array.
void callICRT(bn_t* coefs, cuyasheint_t* d_polyCRT, const int N, const int NPolis, cudaStream_t) {
  if (NPolis != (int)reference_shim::crt_primes().size()) {
    std::fprintf(stderr, "callICRT: NPolis differs from the prime set\n");
    std::abort();
  }
  ok(yh_icrt_bn(ring(), d_polyCRT, N, N, coefs), "yh_icrt_bn");
  ok(yh_ctx_sync(context()), "yh_ctx_sync");
}
