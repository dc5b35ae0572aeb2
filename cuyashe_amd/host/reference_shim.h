// The reference's device-layer interface, restated so that its unchanged
// callers (Polynomial, Integer, common_multiplication) bind to the gfx950 C
// ABI (include/yashe_hip.h) instead of CUDA kernels. A maintainer replaces
// src/cuda_functions.cu and the CRT/ICRT entry points of src/cuda_bn.cu with
// reference_shim.cpp and keeps these declarations, which match
// src/cuda_functions.h:19-84 and src/cuda_bn.h:18-23,86-87 (plus the
// add_mode_t / ntt_mode_t values of src/settings.h:36-37).
//
// Conventions kept from the reference: callee-allocated results
// (callPolynomialMul, callPolynomialOPInteger), residues [NPolis][N] of the
// primes in Polynomial::CRTPrimes, bn_t coefficients on the device, a device
// synchronisation at the end of callPolynomialMul. The cudaStream_t
// arguments are accepted and ignored: every call runs on the library
// context's stream, in call order.
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/yashe_hip.h"

typedef uint64_t cuyasheint_t;
typedef void* cudaStream_t;
enum add_mode_t { ADD, SUB, MUL, DIV, MOD };
enum ntt_mode_t { INVERSE, FORWARD };

// The same bytes as the reference's bn_st {alloc, used, sign, dp}.
typedef yh_bn_t bn_t;

class CUDAFunctions {
 public:
  static int N;
  static int std_bn_t_alloc;
  static cuyasheint_t* d_inner_results;
  static cuyasheint_t* d_inner_results_used;
  static cuyasheint_t wN;
  static cuyasheint_t* d_W;
  static cuyasheint_t* d_WInv;

  static void callPolynomialAddSub(cuyasheint_t* c, cuyasheint_t* a, cuyasheint_t* b, int size, int OP,
                                   cudaStream_t stream);
  static void callPolynomialAddSubInPlace(cudaStream_t stream, cuyasheint_t* a, cuyasheint_t* b, int size, int OP);
  static cuyasheint_t* callPolynomialMul(cudaStream_t stream, cuyasheint_t* a, const bool realign_A, const int A_N,
                                         cuyasheint_t* b, const bool realign_B, const int B_N, const int N,
                                         const int NPolis);
  static cuyasheint_t* callPolynomialOPInteger(const int opcode, cudaStream_t stream, cuyasheint_t* a,
                                               cuyasheint_t* integer_array, const int N, const int NPolis);
  static void callNTT(const int N, const int NPolis, cuyasheint_t* dataI, cuyasheint_t* dataO, const int type);
  static void init(int N);
  static void write_crt_primes();
};

void callCRT(bn_t* coefs, const int used_coefs, cuyasheint_t* d_polyCRT, const int N, const int NPolis,
             cudaStream_t stream);
void callICRT(bn_t* coefs, cuyasheint_t* d_polyCRT, const int N, const int NPolis, cudaStream_t stream);

// The device part of Polynomial::gen_crt_primes (src/polynomial.h:898-955):
// the ring of (degree, q) whose primes make every product exact. Its moduli
// become the shim's CRT prime set (what Polynomial::CRTPrimes holds in the
// reference); the library context is created on first use.
namespace reference_shim {
yh_ctx* context();
yh_ring* ring();
void gen_crt_primes(const uint64_t* q_limbs, int q_words, int degree);
const std::vector<uint64_t>& crt_primes();
void release();
}  // namespace reference_shim
