// HIP kernels for gfx950 (CDNA4): CRT, Goldilocks NTT, pointwise products,
// ICRT and the YASHE epilogues.
//
// Reference kernels these replace (behaviour, not structure):
//  * cuCRT                src/cuda_bn.cu:803-835     -> crt_residue()
//  * NTT / NTTIteration   src/cuda_functions.cu:404-458 (one radix-2 stage per
//                         launch)                    -> ntt_fwd()/ntt_inv(): a
//                         whole transform per workgroup, LDS-resident, four
//                         radix-2 stages per register pass
//  * polynomialNTTMul     src/cuda_functions.cu:460-474 -> fused pointwise step
//  * NTTScale             src/cuda_functions.cu:397-402 -> multiply by N^-1 mod P
//  * cuPreICRT + cuICRT   src/cuda_bn.cu:854-968     -> k_icrt (fixed-point
//                         quotient, no per-coefficient Barrett by M)
//  * Polynomial::reduce   src/cuda_functions.cu:1040-1160 (host fold + mod q)
//                         -> negacyclic transforms make the fold implicit; the
//                         mod-q / scale-and-round epilogues of k_icrt
//  * Ciphertext::operator* rounding  src/ciphertext.cpp:55-63 -> EPI_SCALE
//  * Ciphertext::worddecomp          src/ciphertext.cpp:109-130 -> digit loads
//                         inside k_ks_digits (each digit transformed once,
//                         shared by every keyswitch modulus)
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gl64.hpp"
#include "modp.hpp"
#include "ntt16.hpp"

namespace yhk {

constexpr int E = 16;  // elements per thread in every NTT pass (4 radix-2 stages)

// Minimum waves per SIMD requested from the compiler for the transform
// kernels (2 caps them at 256 VGPRs).
//
// Register pressure: every address a transform computes from threadIdx.x
// (LDS slots, table and operand offsets) is invariant across the transforms
// of one thread, and the compiler keeps them all live (over 100 VGPRs). The
// kernels therefore hand each transform a copy of the thread index that an
// empty asm makes opaque (`opaque_tid`): the addresses are recomputed (a few
// cheap instructions) and the registers are free for more waves.
#ifndef YH_MIN_WAVES
#define YH_MIN_WAVES 2
#endif
// The fused product kernel runs at four waves per SIMD (119 VGPRs at
// n = 4096, no spills; five waves spill); the keyswitch kernels fit four or
// more without spills.
#ifndef YH_PROD_WAVES
#define YH_PROD_WAVES 4
#endif
// The keyswitch multiply-accumulate (k_ks_mac) loads l_w digit and key
// transforms per thread before its inverse transform.
#ifndef YH_KSMAC_WAVES
#define YH_KSMAC_WAVES 4
#endif
// Coefficients whose l_w products are accumulated side by side (four: no
// spills at 128 VGPRs; eight spilled 11 dwords per lane and ran 1 % slower).
#ifndef YH_KSMAC_GROUP
#define YH_KSMAC_GROUP 4
#endif
// The digit transforms (k_ks_digits, 79 VGPRs) may run six waves; forcing
// eight (64 VGPRs, small spills) measured slower. At n = 4096 the LDS of a
// transform (34 KB) admits four workgroups per CU, so every transform kernel
// there runs four waves per SIMD. The reconstructions have no LDS; their
// registers set their waves (k_icrt_pipe asks for YH_ICRT_PIPE_WAVES).
#ifndef YH_KSDIG_WAVES
#define YH_KSDIG_WAVES 6
#endif
// k_ks_scale_digits: the 16 coefficients' 32-bit words (64 VGPRs for a
// two-limb q) stay in registers through the transforms
#ifndef YH_KSSD_WAVES
#define YH_KSSD_WAVES 4
#endif
// Rows of the next YH_KSSD_AHEAD coefficients requested before the current
// one is reconstructed (one: 128 VGPRs, 6 dwords spilled; two spills 96)
#ifndef YH_KSSD_AHEAD
#define YH_KSSD_AHEAD 1
#endif
// coefficients per thread of the reconstruction kernel outside its
// pipelined path (two measured slower than one at n = 4096, two limbs)
#ifndef YH_ICRT_COEFS
#define YH_ICRT_COEFS 1
#endif
// threads per workgroup of the reconstruction kernel
#ifndef YH_ICRT_BLOCK
#define YH_ICRT_BLOCK 256
#endif

// Device view of one ring's tables (all pointers are device memory).
struct RingDev {
  int n, logn, LQ, LW;
  int R, w_bits, lw;
  uint64_t t;
  NttTab tab;               // length-n transform tables (negacyclic)
  const uint64_t* q;        // [LQ]
  const uint64_t* mu;       // [LW+2-LQ]
  // Two modulus sets, each P followed by small primes: 0 = ring products
  // (primes sized for one product per convolution), 1 = keyswitch (sized
  // for the l_w-digit multiply-accumulate). Per set:
  int set_R[2];
  const uint64_t* primes[2];
  const uint64_t* barrett[2];  // floor(2^64 / p_i)
  const uint64_t* pow32[2];    // [R][2*LW]: 2^(32k) mod p_i
  const uint64_t* Mi[2];    // [R][LW]
  const uint64_t* invMi[2]; // [R]
  const SmallPrime* sp[2];  // [R]: constants of the small primes (row 0 unused)
  const uint64_t* M[2];     // [LW]
  // Epilogues mod q without the full integer (per set): M_i mod q [R][LQ]
  // and q - (M mod q) [LQ]; scale-and-round splits of the product set
  // t M_i = q alpha_i + beta_i: alpha_i mod q and beta_i [R][LQ],
  // q - beta_M and q - ((alpha_M + 1) mod q) [LQ].
  const uint64_t* Mq[2];
  const uint64_t* negMq[2];
  const uint64_t* alpha;
  const uint64_t* beta;
  const uint64_t* negbetaM;
  const uint64_t* negalpha1;
  // one-word division of the epilogues (rings with bits(q) >= 64 LQ - 26)
  int fast_div;
  int q_shift;
  uint64_t qmu;
};

// threadIdx.x, recomputed after `dep` is ready (see the register note above).
template <typename T>
__device__ __forceinline__ int opaque_tid(const T& dep) {
  int g = threadIdx.x;
  asm volatile("" : "+v"(g) : "v"(dep));
  return g;
}

// ---------------------------------------------------------------- workgroup mapping
// MI355X deals consecutive workgroup ids round-robin over its 8 XCDs, each
// with a private L2. The fused kernels run one workgroup per (operand, prime);
// this mapping keeps all primes of one operand on one XCD so its coefficient
// reads (and the shared key/twiddle tables) hit that XCD's L2.
constexpr int XCDS = 8;

__host__ __device__ inline int xcd_grid(int batch, int R) { return ((batch + XCDS - 1) / XCDS) * XCDS * R; }

// Workgroup w -> (operand, prime); false for padding workgroups.
__device__ __forceinline__ bool xcd_map(int w, int R, int batch, int& pair, int& prime) {
  const int x = w % XCDS, k = w / XCDS;
  pair = (k / R) * XCDS + x;
  prime = k % R;
  return pair < batch;
}

// The same map with the operands of each XCD taken from the last to the
// first (operand groups in reverse order; an operand keeps its XCD).
__device__ __forceinline__ bool xcd_map_rev(int w, int R, int batch, int& pair, int& prime) {
  const int x = w % XCDS, k = w / XCDS;
  const int groups = (batch + XCDS - 1) / XCDS;
  pair = (groups - 1 - k / R) * XCDS + x;
  prime = k % R;
  return pair < batch;
}

// ---------------------------------------------------------------- small mod p

// x mod p for x < 2^64 with b = floor(2^64/p) (p < 2^40).
__device__ __forceinline__ uint64_t mod_p(uint64_t x, uint64_t p, uint64_t b) {
  uint64_t qh = gl::umulhi(x, b);
  uint64_t r = x - qh * p;
  return r >= p ? r - p : r;
}

// Residue r in [0, p) as a Goldilocks element of its centred lift in
// (-p/2, p/2]; centring halves the magnitude entering the convolution.
__device__ __forceinline__ uint64_t center(uint32_t r, uint32_t p) {
  return r > (p >> 1) ? gl::P - (p - r) : r;
}

// centred residue r times w (the negacyclic twist), as a Goldilocks element;
// a negative residue multiplies |r| by -w (w != 0, so P - w is canonical)
__device__ __forceinline__ uint64_t twist_centered(uint32_t r, uint32_t p, uint64_t w) {
  const bool neg = r > (p >> 1);
  return mul_small(neg ? p - r : r, neg ? gl::P - w : w);
}

// a 2^e mod P for a < 2^ABITS (ABITS <= 32) and 0 <= e < 96 known at
// compile time: a plain shift while the product stays below 2^63 < P.
template <int ABITS>
__device__ __forceinline__ uint64_t shift_small(uint32_t a, int e) {
  return e + ABITS <= 63 ? (uint64_t)a << e : gl::mul_pow2((uint64_t)a, e);
}

// Input slot c of a merged negacyclic transform (ntt16_fwd<LOGN, true>): the
// centred residue of r in [0, p) times 2^(78 c), canonical. 2^e with
// e >= 96 is -2^(e-96); the result is negative when exactly one of the
// centred value and that sign is, and 0 stays 0.
__device__ __forceinline__ uint64_t twist_slot(uint32_t r, uint32_t p, int c) {
  int e = (78 * c) % 192;
  const bool ng = e >= 96;
  if (ng) e -= 96;
  const uint32_t h = p >> 1;
  const bool neg = r > h;
  const uint32_t a = neg ? p - r : r;     // < 2^27
  const bool f = ng ? r - 1u < h : neg;  // ng: negative iff the centred value is > 0
  const uint64_t v = shift_small<27>(a, e);
  return f ? gl::P - v : v;
}

// The same for a keyswitch digit d in [0, 2^w) centred as d - 2^(w-1)
// (half = 2^(w-1) <= 2^31).
__device__ __forceinline__ uint64_t twist_slot_digit(uint64_t d, uint64_t half, int c) {
  int e = (78 * c) % 192;
  const bool ng = e >= 96;
  if (ng) e -= 96;
  const bool neg = d < half;
  const uint32_t a = (uint32_t)(neg ? half - d : d - half);  // <= 2^31
  const bool f = ng ? d > half : neg;
  const uint64_t v = shift_small<32>(a, e);
  return f ? gl::P - v : v;
}

// CRT of one coefficient (LIN 64-bit limbs, limb-major with stride `ls`):
// sum_k word_k * (2^(32k) mod p) with 32x32-bit products (the table entries
// are below p < 2^28), then mod_small. Matches bn_mod1_low
// (src/cuda_bn.cu:290-317): the residue x mod p.
__device__ __forceinline__ uint32_t crt_residue(const uint64_t* a, int lin, size_t ls, const uint64_t* pw,
                                                const SmallPrime& k) {
  uint64_t acc = 0;
  for (int l = 0; l < lin; ++l) {
    const uint64_t v = a[(size_t)l * ls];
    acc += (uint64_t)(uint32_t)v * (uint32_t)pw[2 * l] + (uint64_t)(uint32_t)(v >> 32) * (uint32_t)pw[2 * l + 1];
    // each product < 2^60; reducing every 4 limbs keeps the sum < 2^63
    if ((l & 3) == 3) acc = mod_small(acc, k);
  }
  return mod_small(acc, k);
}

// ---------------------------------------------------------------- modulus 0 = P
// The first CRT modulus is the Goldilocks prime itself: the transform gives
// the convolution mod P exactly for any operand size, so its residues are
// canonical values in [0, P) (not centred) and its y is 64 bits wide.

// Residue mod P of a coefficient of `lin` 64-bit limbs (stride ls): Horner
// steps r 2^64 + limb, each one 128-bit reduction.
__device__ __forceinline__ uint64_t crt_residue_P(const uint64_t* a, int lin, size_t ls) {
  uint64_t r = 0;
  for (int l = lin - 1; l >= 0; --l) r = gl::reduce128(r, a[(size_t)l * ls]);
  return r;
}

// Rows of the reconstruction scratch per operand: modulus 0 keeps its 64-bit
// y in rows 0 (low half) and 1 (high half), modulus i >= 1 in row i + 1.
__host__ __device__ constexpr int y_rows(int R) { return R + 1; }

// Stores y of modulus i at coefficient j; Yp is the operand's first row.
__device__ __forceinline__ void store_y(uint32_t* Yp, size_t N, int i, int j, uint64_t y) {
  if (i == 0) {
    stg(Yp + j, (uint32_t)y);
    stg(Yp + N + j, (uint32_t)(y >> 32));
  } else {
    stg(Yp + (size_t)(i + 1) * N + j, (uint32_t)y);
  }
}

// 64-bit words through a buffer resource: the workgroup-uniform base and
// word offset sit in scalar registers and the thread's byte offset in one
// vector register, so a load takes no vector address arithmetic. The
// resource spans `bytes` from `base`: every load's scalar plus vector offset
// must lie below it (reads beyond the range return 0).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t words_rsrc(const uint64_t* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(base), (short)0, (int)bytes, 0x00020000);
}
// An operand of `lin` limbs at stride `ls` words, n words each, spans less
// than 2^31 bytes (it fits a buffer resource and its 32-bit offsets).
__device__ __forceinline__ bool full_extent(int lin, size_t ls, int n) {
  return ((size_t)(lin - 1) * ls + (size_t)n) * 8 < ((size_t)1 << 31);
}
__device__ __forceinline__ uint64_t word_load(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, soff, 0);
  return (uint64_t)v[0] | ((uint64_t)v[1] << 32);
}
__device__ __forceinline__ void word_store(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff, uint64_t w) {
  const __attribute__((ext_vector_type(2))) unsigned int d = {(unsigned)w, (unsigned)(w >> 32)};
  __builtin_amdgcn_raw_buffer_store_b64(d, rs, voff, soff, 0);
}

// 32-bit rows of one operand through a buffer resource: the workgroup-
// uniform base and row offset sit in scalar registers and the thread's byte
// offset in one vector register, so a row load takes no vector address
// arithmetic. The resource spans `bytes` from `base`, which covers every
// offset (scalar plus vector) read through it; reads beyond it return 0.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rows_rsrc(const uint32_t* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t*>(base), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t row_load(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff) {
  return __builtin_amdgcn_raw_buffer_load_b32(rs, voff, soff, 0);
}
__device__ __forceinline__ void row_store(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t soff, uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b32(v, rs, voff, soff, 0);
}

// ---------------------------------------------------------------- fused products
//
// k_prod: one workgroup per (pair, prime). CRT both operands on load, forward
// transforms, pointwise product in registers, inverse transform, and the
// residue y = (c mod p) * (M/p)^-1 mod p written as u32 for the ICRT.
// Negacyclic (ring product mod x^n + 1) when T/Ti are the psi tables, cyclic
// (plain product mod x^N - 1) with the omega tables.
// Operand layout: limb-major [lin][lstride] per operand, operand `pair` at
// offset pair*ostride; coefficients j >= ncoef read as zero.
struct ProdArgs {
  const uint64_t* A; const uint64_t* B;
  // optional per-pair operand addresses (device arrays of device pointers);
  // when set they replace A + pair*ostride_a / B + pair*ostride_b
  const uint64_t* const* a_list;
  const uint64_t* const* b_list;
  int lin_a, lin_b;
  int ncoef_a, ncoef_b;
  size_t lstride_a, lstride_b;
  size_t ostride_a, ostride_b;
  uint32_t* Y;          // [pair][R][N]
  int logN;
  NttTab tab;           // tables of the length-N transform
  int set;
  int batch;
};

// CRT of 16 coefficients into the entry layout of ntt16_fwd (j = g + c*N/16),
// centred, and for negacyclic transforms twisted by psi^j.
// The coefficients are loaded and reduced in two groups of eight: the empty
// asm makes the second group's loads wait for the first group's residues, so
// at most eight coefficients' limbs are in registers at once (the compiler
// would otherwise issue all loads up front and spill while the first
// operand's transform is held in registers).
#ifndef YH_LOAD_GROUP_B
#define YH_LOAD_GROUP_B 4
#endif
template <int G = 8>
__device__ __forceinline__ void group_fence(uint64_t* x) {
  if constexpr (G == 4) {
    asm volatile("" : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3])::"memory");
  } else {
    asm volatile("" : "+v"(x[0]), "+v"(x[1]), "+v"(x[2]), "+v"(x[3]), "+v"(x[4]), "+v"(x[5]), "+v"(x[6]),
                 "+v"(x[7])::"memory");
  }
}

// MERGED: negacyclic input of ntt16_fwd<LOGN, true> (slot shifts only; the
// twist table is not read).
// Within a group the limbs are the outer loop, so the group's eight loads of
// one limb are in flight together (crt_residue's steps, eight coefficients
// side by side).
// FULL: the operand has all n coefficients (ncoef >= n), so no load needs
// the guard j < ncoef (its exec-mask branches cost about 4 % of k_prod's
// instructions), and its limbs span less than 2^31 bytes, so they are read
// through a buffer resource.
template <int LOGN, bool NEGA, bool MERGED = false, int G = 8, bool FULL = false>
__device__ __forceinline__ void load_crt(uint64_t* x, const uint64_t* a, int lin, int ncoef, size_t ls,
                                         const uint64_t* pw, const SmallPrime& k, const uint64_t* twist,
                                         int g = threadIdx.x) {
  constexpr int stride = 1 << (LOGN - 4);
  // FULL: the operand's limbs through a buffer resource (scalar base and
  // offsets, one vector offset per thread); the caller checks that its
  // extent fits the resource's 32-bit range
  const __amdgpu_buffer_rsrc_t ra = words_rsrc(a, FULL ? (uint32_t)(((size_t)(lin - 1) * ls + (1 << LOGN)) * 8) : 0u);
#pragma unroll
  for (int h = 0; h < E; h += G) {
    uint64_t acc[G];
#pragma unroll
    for (int c = 0; c < G; ++c) acc[c] = 0;
    for (int l = 0; l < lin; ++l) {
      const uint32_t w0 = (uint32_t)pw[2 * l], w1 = (uint32_t)pw[2 * l + 1];
      const uint64_t* al = a + (size_t)l * ls;
#pragma unroll
      for (int c = 0; c < G; ++c) {
        const int j = g + (h + c) * stride;
        const uint64_t v = FULL ? word_load(ra, (uint32_t)g * 8u, (uint32_t)(l * ls + (h + c) * stride) * 8u)
                                : j < ncoef ? ldg(al + j) : 0;
        // each product < 2^60; reducing every 4 limbs keeps the sum < 2^63
        acc[c] += (uint64_t)(uint32_t)v * w0 + (uint64_t)(uint32_t)(v >> 32) * w1;
      }
      if ((l & 3) == 3) {
#pragma unroll
        for (int c = 0; c < G; ++c) acc[c] = mod_small(acc[c], k);
      }
    }
#pragma unroll
    for (int c = h; c < h + G; ++c) {
      const int j = g + c * stride;
      const uint32_t r = FULL || j < ncoef ? mod_small(acc[c - h], k) : 0u;
      x[c] = MERGED ? twist_slot(r, k.p, c) : NEGA ? twist_centered(r, k.p, ldg(twist + j)) : center(r, k.p);
    }
    if (h + G < E) group_fence<G>(x + h);
  }
}

// Modulus 0 (P): exact residues mod P, twisted by psi^j, not centred
// (MERGED: by the slot factor 2^(78 c) only).
template <int LOGN, bool NEGA, bool MERGED = false, int G = 8, bool FULL = false>
__device__ __forceinline__ void load_crt_P(uint64_t* x, const uint64_t* a, int lin, int ncoef, size_t ls,
                                           const uint64_t* twist, int g) {
  constexpr int stride = 1 << (LOGN - 4);
  // FULL: the operand's limbs through a buffer resource (scalar base and
  // offsets, one vector offset per thread); the caller checks that its
  // extent fits the resource's 32-bit range
  const __amdgpu_buffer_rsrc_t ra = words_rsrc(a, FULL ? (uint32_t)(((size_t)(lin - 1) * ls + (1 << LOGN)) * 8) : 0u);
#pragma unroll
  for (int h = 0; h < E; h += G) {
    // crt_residue_P's Horner steps, limbs outer
    uint64_t r[G];
#pragma unroll
    for (int c = 0; c < G; ++c) r[c] = 0;
    for (int l = lin - 1; l >= 0; --l) {
      const uint64_t* al = a + (size_t)l * ls;
#pragma unroll
      for (int c = 0; c < G; ++c) {
        const int j = g + (h + c) * stride;
        r[c] = gl::reduce128(r[c], FULL ? word_load(ra, (uint32_t)g * 8u, (uint32_t)(l * ls + (h + c) * stride) * 8u)
                                        : j < ncoef ? ldg(al + j) : 0);
      }
    }
#pragma unroll
    for (int c = h; c < h + G; ++c) {
      const int j = g + c * stride;
      x[c] = MERGED ? gl::mul_pow2(r[c - h], (78 * c) % 192) : NEGA ? gl::mul(r[c - h], ldg(twist + j)) : r[c - h];
    }
    if (h + G < E) group_fence<G>(x + h);
  }
}

// y of a residue after the inverse transform (v already scaled/untwisted):
// modulus 0 (gold) keeps the exact residue mod P, the others reduce the
// signed convolution value mod p_i (finish_small); both are multiplied by
// (M/m_i)^-1 mod m_i. The constants come in registers (sk = the prime's row,
// inv0 = (M/P)^-1 mod P): the rows' stores may alias the tables, so reading
// the tables here would reload them after every store.
__device__ __forceinline__ uint64_t finish_any(bool gold, const SmallPrime& sk, uint64_t inv0, uint64_t v) {
  return gold ? gl::mul(v, inv0) : (uint64_t)finish_small(v, sk);
}

template <int LOGN, bool NEGA>
__global__ void __launch_bounds__((1 << LOGN) / 16, YH_PROD_WAVES) k_prod(RingDev rd, ProdArgs pa) {
  extern __shared__ uint64_t lds[];
  constexpr int N = 1 << LOGN;
  int pair, i;
  if (!xcd_map(blockIdx.x, rd.set_R[pa.set], pa.batch, pair, i)) return;
  const SmallPrime sk = rd.sp[pa.set][i];  // row 0 (modulus P) is unused
  const uint64_t* pw = rd.pow32[pa.set] + (size_t)i * 2 * rd.LW;
  uint64_t xa[E], xb[E];
  const uint64_t* A = pa.a_list ? pa.a_list[pair] : pa.A + pair * pa.ostride_a;
  const uint64_t* Bp = pa.b_list ? pa.b_list[pair] : pa.B + pair * pa.ostride_b;
  const bool gold = i == 0;
  // negacyclic products fold the twist into the first forward pass and the
  // untwist into the last inverse pass (ntt16.hpp, MERGED)
  constexpr bool M = NEGA;
  const int g1 = opaque_tid(pair);
  // operands of all n coefficients (every ciphertext) load without guards
  constexpr int GB = YH_LOAD_GROUP_B;
  const bool full_a = pa.ncoef_a >= N && full_extent(pa.lin_a, pa.lstride_a, N);
  const bool full_b = pa.ncoef_b >= N && full_extent(pa.lin_b, pa.lstride_b, N);
  if (gold) {
    if (full_a) load_crt_P<LOGN, NEGA, M, 8, true>(xa, A, pa.lin_a, N, pa.lstride_a, pa.tab.twist, g1);
    else load_crt_P<LOGN, NEGA, M>(xa, A, pa.lin_a, pa.ncoef_a, pa.lstride_a, pa.tab.twist, g1);
  } else {
    if (full_a) load_crt<LOGN, NEGA, M, 8, true>(xa, A, pa.lin_a, N, pa.lstride_a, pw, sk, pa.tab.twist, g1);
    else load_crt<LOGN, NEGA, M>(xa, A, pa.lin_a, pa.ncoef_a, pa.lstride_a, pw, sk, pa.tab.twist, g1);
  }
  ntt16_fwd<LOGN, M>(xa, lds, pa.tab.fwd, g1, pa.tab.fwd0);
  // The second operand re-reads the tables (L1 hits) rather than keeping the
  // first transform's twiddles live in registers (the empty asm hides that
  // the addresses repeat); fewer registers, more waves per SIMD.
  const uint64_t* twist2 = pa.tab.twist;
  const uint64_t* fwd2 = pa.tab.fwd;
  const uint64_t* fwd02 = pa.tab.fwd0;
  asm volatile("" : "+s"(twist2), "+s"(fwd2), "+s"(fwd02));
  const int g2 = opaque_tid(xa[E - 1]);
  if (gold) {
    if (full_b) load_crt_P<LOGN, NEGA, M, GB, true>(xb, Bp, pa.lin_b, N, pa.lstride_b, twist2, g2);
    else load_crt_P<LOGN, NEGA, M, GB>(xb, Bp, pa.lin_b, pa.ncoef_b, pa.lstride_b, twist2, g2);
  } else {
    if (full_b) load_crt<LOGN, NEGA, M, GB, true>(xb, Bp, pa.lin_b, N, pa.lstride_b, pw, sk, twist2, g2);
    else load_crt<LOGN, NEGA, M, GB>(xb, Bp, pa.lin_b, pa.ncoef_b, pa.lstride_b, pw, sk, twist2, g2);
  }
  __syncthreads();  // every thread has left the LDS of the first transform
  ntt16_fwd<LOGN, M>(xb, lds, fwd2, g2, fwd02);
#pragma unroll
  for (int c = 0; c < E; ++c) xa[c] = gl::mont_mul(xa[c], xb[c]);  // x 2^-64, cancelled by *_pw
  __syncthreads();
  const int g3 = opaque_tid(xa[0]);
  ntt16_inv<LOGN, M>(xa, lds, pa.tab.inv, g3, pa.tab.inv0_pw);
  const int nrow = y_rows(rd.set_R[pa.set]);
  // the pair's rows through a buffer resource (scalar base and row offset)
  const __amdgpu_buffer_rsrc_t y = rows_rsrc(pa.Y + (size_t)pair * nrow * N, (uint32_t)(nrow * N) * 4u);
  const uint32_t yv = (uint32_t)g3 * 4u;
  constexpr int stride = N >> 4;
  auto scaled = [&](int c, int j) {
    return M ? xa[c] : NEGA ? gl::mont_mul(xa[c], pa.tab.untwist_pw[j]) : gl::mont_mul(xa[c], pa.tab.n_inv_pw);
  };
  // one straight-line loop per modulus kind (i is uniform), with the
  // constants read once; modulus 0 keeps its 64-bit y in rows 0 and 1,
  // modulus i >= 1 in row i + 1 (store_y's layout)
  if (gold) {
    const uint64_t inv0 = rd.invMi[pa.set][0];
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const uint64_t v = gl::mul(scaled(c, g3 + c * stride), inv0);
      row_store(y, yv, (uint32_t)(c * stride) * 4u, (uint32_t)v);
      row_store(y, yv, (uint32_t)(N + c * stride) * 4u, (uint32_t)(v >> 32));
    }
  } else {
#pragma unroll
    for (int c = 0; c < E; ++c)
      row_store(y, yv, (uint32_t)((i + 1) * N + c * stride) * 4u, finish_small(scaled(c, g3 + c * stride), sk));
  }
}

// Small transforms (N < 64): one thread per (pair, prime) computes the same
// Goldilocks convolution of centred residues directly.
static __global__ void k_prod_small(RingDev rd, ProdArgs pa, int negacyclic) {
  const int i = blockIdx.x;
  const int pair = blockIdx.y;
  if (threadIdx.x != 0) return;
  const int N = 1 << pa.logN;
  const SmallPrime sk = rd.sp[pa.set][i];
  const uint64_t* pw = rd.pow32[pa.set] + (size_t)i * 2 * rd.LW;
  uint64_t ra[64], rb[64], rc[64];
  const uint64_t* A = pa.a_list ? pa.a_list[pair] : pa.A + pair * pa.ostride_a;
  const uint64_t* Bp = pa.b_list ? pa.b_list[pair] : pa.B + pair * pa.ostride_b;
  const bool gold = i == 0;
  for (int j = 0; j < N; ++j) {
    const uint64_t* a = A + j;
    const uint64_t* bb = Bp + j;
    if (gold) {
      ra[j] = j < pa.ncoef_a ? crt_residue_P(a, pa.lin_a, pa.lstride_a) : 0;
      rb[j] = j < pa.ncoef_b ? crt_residue_P(bb, pa.lin_b, pa.lstride_b) : 0;
    } else {
      ra[j] = j < pa.ncoef_a ? center(crt_residue(a, pa.lin_a, pa.lstride_a, pw, sk), sk.p) : 0;
      rb[j] = j < pa.ncoef_b ? center(crt_residue(bb, pa.lin_b, pa.lstride_b, pw, sk), sk.p) : 0;
    }
    rc[j] = 0;
  }
  for (int u = 0; u < N; ++u)
    for (int v = 0; v < N; ++v) {
      uint64_t prod = gl::mul(ra[u], rb[v]);
      int k = u + v;
      if (k >= N) {
        k -= N;
        if (negacyclic) prod = gl::neg(prod);
      }
      rc[k] = gl::add(rc[k], prod);
    }
  uint32_t* y = pa.Y + (size_t)pair * y_rows(rd.set_R[pa.set]) * N;
  const uint64_t inv0 = rd.invMi[pa.set][0];
  for (int j = 0; j < N; ++j) store_y(y, N, i, j, finish_any(gold, sk, inv0, rc[j]));
}

// ---------------------------------------------------------------- keyswitch
//
// Two launches per batch. k_ks_digits: one workgroup per (ciphertext, digit
// k); digit k of each coefficient (bits [k*w, (k+1)*w) of c in [0, q)) is
// read straight from the ciphertext limbs, centred (d - 2^(w-1)), twisted and
// transformed once, unreduced, and the transform is written to scratch.
// k_ks_mac: one workgroup per (ciphertext, modulus of the keyswitch set)
// multiplies the l_w digit transforms by the resident NTT(gamma_k mod m_i)
// and accumulates, then runs one inverse transform. Sharing the digit
// transforms between moduli is exact because the keyswitch primes are sized
// for centred unreduced digits (l_w n 2^(w-1) (p-1)/2 <= (P-1)/2): the
// accumulated value is the integer sum_k (d_k - 2^(w-1)) * (gamma_k mod p_i);
// the epilogue adds the per-key correction 2^(w-1) sum_k 1 * gamma_k mod p_i
// (k_ks_corr) and reduces mod p_i. Centring halves the digit bound, which
// admits primes twice as large and saves one keyswitch modulus on the common
// rings. The reference (src/ciphertext.cpp:81-130) runs one full CRT product
// per digit.
struct KsArgs {
  const uint64_t* C;        // [ct][LQ][n]
  size_t ostride;
  const uint64_t* evk;      // [lw][R_ks][n] transform domain
  const uint64_t* corr;     // [R_ks][n]: 2^(w_bits-1) sum_k (1 * gamma_k) mod m_i (digit centring)
  uint64_t* D;              // [ct][lw][n] digit transforms (scratch)
  uint32_t* Y;              // [ct][R_ks][n]
  int batch;
  // k_ks_mac takes the ciphertexts from the last to the first when the digit
  // transforms were written from the first to the last (the last ones are
  // the likeliest still cached), and in order otherwise
  int reverse;
  const uint32_t* P;        // k_ks_scale_digits: the product's rows [ct][R + 1][n]
};

__device__ __forceinline__ uint64_t digit_of(const uint64_t* a, int n, int lq, int k, int wb) {
  int bit = k * wb;
  int l = bit >> 6, off = bit & 63;
  uint64_t v = a[(size_t)l * n] >> off;
  if (off + wb > 64 && l + 1 < lq) v |= a[(size_t)(l + 1) * n] << (64 - off);
  return v & ((1ull << wb) - 1);
}

// W32: digits are the 32-bit words of the limbs (w = 2^32), read directly.
// Output layout (shared with the resident keys): slot c of thread g at
// c * n/16 + g, so each store or load instruction of a wavefront covers
// 512 contiguous bytes.
template <int LOGN, bool W32>
__global__ void __launch_bounds__((1 << LOGN) / 16, YH_KSDIG_WAVES) k_ks_digits(RingDev rd, KsArgs ka) {
  extern __shared__ uint64_t lds[];
  constexpr int N = 1 << LOGN;
  constexpr int stride = N >> 4;
  int ct, k;
  if (!xcd_map(blockIdx.x, rd.lw, ka.batch, ct, k)) return;
  const uint64_t* c0 = ka.C + ct * ka.ostride;
  const int g = threadIdx.x;
  const uint32_t* w32 = reinterpret_cast<const uint32_t*>(c0 + (size_t)(k >> 1) * N) + (k & 1);
  const uint64_t half = W32 ? (1ull << 31) : (1ull << (rd.w_bits - 1));
  uint64_t x[E];
#pragma unroll
  for (int h = 0; h < E; h += 8) {
#pragma unroll
    for (int c = h; c < h + 8; ++c) {
      const int j = g + c * stride;
      const uint64_t d = W32 ? (uint64_t)w32[2 * j] : digit_of(c0 + j, N, rd.LQ, k, rd.w_bits);
      // the centred digit d - 2^(w-1), twisted: |.| <= 2^31 times +-psi^j
      // (merged transforms: times the slot factor 2^(78 c) only)
      x[c] = twist_slot_digit(d, half, c);
    }
    if (h + 8 < E) group_fence(x + h);
  }
  ntt16_fwd<LOGN, true>(x, lds, rd.tab.fwd, g, rd.tab.fwd0);
  uint64_t* o = ka.D + ((size_t)ct * rd.lw + k) * N + g;
#pragma unroll
  for (int c = 0; c < E; ++c) o[c * stride] = x[c];
}

template <int LOGN>
__global__ void __launch_bounds__((1 << LOGN) / 16, YH_KSMAC_WAVES) k_ks_mac(RingDev rd, KsArgs ka) {
  extern __shared__ uint64_t lds[];
  constexpr int N = 1 << LOGN;
  constexpr int stride = N >> 4;
  const int Rks = rd.set_R[1];
  int ct, i;
  const bool mine = ka.reverse ? xcd_map_rev(blockIdx.x, Rks, ka.batch, ct, i) : xcd_map(blockIdx.x, Rks, ka.batch, ct, i);
  if (!mine) return;
  const int g = threadIdx.x;
  // the ciphertext's digit transforms and this modulus' keys through buffer
  // resources: uniform bases and offsets in scalar registers, one vector
  // offset per thread (no vector address arithmetic per load)
  const __amdgpu_buffer_rsrc_t dg = words_rsrc(ka.D + (size_t)ct * rd.lw * N, (uint32_t)(rd.lw * N) * 8u);
  const __amdgpu_buffer_rsrc_t ev = words_rsrc(ka.evk + (size_t)i * N, (uint32_t)(((rd.lw - 1) * Rks + 1) * N) * 8u);
  const uint32_t voff = (uint32_t)g * 8u;
  // The l_w products of a slot are summed exactly as 128-bit values (top
  // counts the carries out of bit 128, < l_w <= 16) and reduced once.
  uint64_t acc[E];
  constexpr int G = YH_KSMAC_GROUP;
#pragma unroll
  for (int h = 0; h < E; h += G) {
    uint64_t lo[G], hi[G];
    uint32_t top[G];
#pragma unroll
    for (int c = 0; c < G; ++c) {
      lo[c] = hi[c] = 0;
      top[c] = 0;
    }
    for (int k = 0; k < rd.lw; ++k) {
      const uint32_t dk = (uint32_t)(k * N) * 8u;
      const uint32_t ek = (uint32_t)(k * Rks * N) * 8u;
#pragma unroll
      for (int c = 0; c < G; ++c) {
        uint64_t ph, pl;
        const uint32_t co = (uint32_t)((h + c) * stride) * 8u;
        gl::mul128(word_load(dg, voff, dk + co), word_load(ev, voff, ek + co), ph, pl);
        lo[c] += pl;
        const uint64_t t = ph + (lo[c] < pl);  // ph <= 2^64 - 2
        hi[c] += t;
        top[c] += hi[c] < t;
      }
    }
#pragma unroll
    for (int c = 0; c < G; ++c) acc[h + c] = gl::reduce_wide(top[c], hi[c], lo[c]);
  }
  const int g2 = opaque_tid(acc[0]);
  constexpr bool M = true;  // untwist folded into the last inverse pass
  ntt16_inv<LOGN, M>(acc, lds, rd.tab.inv, g2, rd.tab.inv0_m);
  // the inverse gives sum_k d'_k * (gamma_k mod m_i) for the centred digits
  // d'_k = d_k - 2^(w-1); adding corr_i = 2^(w-1) sum_k (1 * gamma_k) mod m_i
  // restores sum_k d_k * gamma_k mod m_i
  uint32_t* y = ka.Y + (size_t)ct * y_rows(Rks) * N;
  const uint64_t* cr = ka.corr + (size_t)i * N;
  if (i == 0) {
    const uint64_t invM = rd.invMi[1][0];
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int j = g2 + c * stride;
      const uint64_t v = gl::add(M ? acc[c] : gl::mont_mul(acc[c], rd.tab.untwist_m[j]), cr[j]);
      store_y(y, N, 0, j, gl::mul(v, invM));
    }
  } else {
    const SmallPrime sp = rd.sp[1][i];
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int j = g2 + c * stride;
      store_y(y, N, i, j, finish_small(M ? acc[c] : gl::mont_mul(acc[c], rd.tab.untwist_m[j]), sp, (uint32_t)cr[j]));
    }
  }
}

// Keyswitch for small rings (n < 64): direct negacyclic convolution per thread.
static __global__ void k_keyswitch_small(RingDev rd, KsArgs ka, const uint64_t* gamma_res /*[lw][Rks][n] centred residues*/) {
  const int i = blockIdx.x;
  const int ct = blockIdx.y;
  if (threadIdx.x != 0) return;
  const int N = rd.n;
  const int Rks = rd.set_R[1];
  const uint64_t p = rd.primes[1][i];
  const uint64_t b = rd.barrett[1][i];
  const uint64_t* c0 = ka.C + ct * ka.ostride;
  const bool gold = i == 0;
  uint64_t rc[64];
  for (int j = 0; j < N; ++j) rc[j] = 0;
  for (int k = 0; k < rd.lw; ++k) {
    uint64_t d[64];
    for (int j = 0; j < N; ++j) {
      const uint64_t dg = digit_of(c0 + j, N, rd.LQ, k, rd.w_bits);
      d[j] = gold ? dg : center(mod_p(dg, p, b), p);
    }
    const uint64_t* gk = gamma_res + ((size_t)k * Rks + i) * N;
    for (int u = 0; u < N; ++u)
      for (int v = 0; v < N; ++v) {
        uint64_t prod = gl::mul(d[u], gk[v]);
        int kk = u + v;
        if (kk >= N) { kk -= N; prod = gl::neg(prod); }
        rc[kk] = gl::add(rc[kk], prod);
      }
  }
  uint32_t* y = ka.Y + (size_t)ct * y_rows(Rks) * N;
  const SmallPrime sk = rd.sp[1][i];
  const uint64_t inv0 = rd.invMi[1][0];
  for (int j = 0; j < N; ++j) store_y(y, N, i, j, finish_any(gold, sk, inv0, rc[j]));
}

// Pre-transform of the keyswitch keys: gamma_k (coefficients, [lw][LQ][n]) ->
// NTT(gamma_k mod p_i, centred) for every prime of the keyswitch set.
template <int LOGN>
__global__ void __launch_bounds__((1 << LOGN) / 16, YH_MIN_WAVES) k_evk_prep(RingDev rd, const uint64_t* G, uint64_t* out) {
  extern __shared__ uint64_t lds[];
  constexpr int N = 1 << LOGN;
  const int i = blockIdx.x;
  const int k = blockIdx.y;
  const int Rks = rd.set_R[1];
  const SmallPrime sk = rd.sp[1][i];
  const uint64_t* pw = rd.pow32[1] + (size_t)i * 2 * rd.LW;
  const uint64_t* g = G + (size_t)k * rd.LQ * N;
  uint64_t* o = out + ((size_t)k * Rks + i) * N;
  uint64_t x[E];
  if (i == 0) load_crt_P<LOGN, true, false, 8, true>(x, g, rd.LQ, N, N, rd.tab.twist, threadIdx.x);
  else load_crt<LOGN, true, false, 8, true>(x, g, rd.LQ, N, N, pw, sk, rd.tab.twist);
  ntt16_fwd<LOGN>(x, lds, rd.tab.fwd);
#pragma unroll
  for (int c = 0; c < E; ++c) o[c * (N >> 4) + threadIdx.x] = x[c];
}

// Correction of the centred keyswitch digits, once per key: corr_i =
// 2^(w-1) sum_k (1 * gamma_k) mod m_i, 1 the all-ones polynomial (its
// negacyclic product with g is 2 prefix_j(g) - total(g)). One workgroup per
// keyswitch modulus computes sum_k 1 * (gamma_k mod m_i) exactly as a
// keyswitch whose every digit transform is NTT(psi^j) (the twisted all-ones
// vector): |.| <= l_w n (p-1)/2, far inside the Goldilocks bound.
template <int LOGN>
__global__ void __launch_bounds__((1 << LOGN) / 16, YH_MIN_WAVES) k_ks_corr(RingDev rd, const uint64_t* evk,
                                                                             uint64_t* corr) {
  extern __shared__ uint64_t lds[];
  constexpr int N = 1 << LOGN;
  constexpr int stride = N >> 4;
  const int i = blockIdx.x;
  const int Rks = rd.set_R[1];
  const int g = threadIdx.x;
  uint64_t x[E];
#pragma unroll
  for (int c = 0; c < E; ++c) x[c] = rd.tab.twist[g + c * stride];
  ntt16_fwd<LOGN>(x, lds, rd.tab.fwd, g);
#pragma unroll
  for (int c = 0; c < E; ++c) {
    uint64_t s = 0;
    for (int k = 0; k < rd.lw; ++k) s = gl::add(s, evk[((size_t)k * Rks + i) * N + c * stride + g]);
    x[c] = gl::mul(x[c], s);
  }
  __syncthreads();
  ntt16_inv<LOGN>(x, lds, rd.tab.inv, g);
  const int wb = rd.w_bits;
  const uint64_t p = rd.primes[1][i], b = rd.barrett[1][i];
  const uint64_t half_p = i == 0 ? 0 : (1ull << (wb - 1)) % p;
#pragma unroll
  for (int c = 0; c < E; ++c) {
    const int j = g + c * stride;
    const uint64_t v = gl::mont_mul(x[c], rd.tab.untwist_m[j]);
    if (i == 0) {
      corr[j] = gl::mul_pow2(v, wb - 1);
    } else {
      const int64_t sv = gl::lift(v);
      uint64_t r = mod_p(sv < 0 ? (uint64_t)(-sv) : (uint64_t)sv, p, b);
      if (sv < 0 && r) r = p - r;
      corr[(size_t)i * N + j] = mod_p(r * half_p, p, b);
    }
  }
}

// Keyswitch keys of small rings (n < 16): centred residues, no transform.
static __global__ void k_evk_residues(RingDev rd, const uint64_t* G, uint64_t* out) {
  const int i = blockIdx.x;
  const int k = blockIdx.y;
  const int Rks = rd.set_R[1];
  const SmallPrime sk = rd.sp[1][i];
  const uint64_t* pw = rd.pow32[1] + (size_t)i * 2 * rd.LW;
  const int N = rd.n;
  const uint64_t* g = G + (size_t)k * rd.LQ * N;
  uint64_t* o = out + ((size_t)k * Rks + i) * N;
  for (int j = threadIdx.x; j < N; j += blockDim.x)
    o[j] = i == 0 ? crt_residue_P(g + j, rd.LQ, N) : center(crt_residue(g + j, rd.LQ, N, pw, sk), sk.p);
}

// ---------------------------------------------------------------- ICRT + epilogues

enum Epilogue { EPI_RAW = 0, EPI_MODQ = 1, EPI_SCALE = 2, EPI_DECRYPT = 3 };

template <int NL>
__device__ __forceinline__ void mac32(uint64_t* acc, const uint64_t* m, uint64_t y) {
  // acc += m * y (mod 2^(64 NL)), y < 2^32
  uint64_t carry = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    uint64_t lo = m[l] * y;
    uint64_t hi = gl::umulhi(m[l], y);
    uint64_t s = acc[l] + lo;
    hi += (s < lo);
    uint64_t s2 = s + carry;
    hi += (s2 < s);
    acc[l] = s2;
    carry = hi;
  }
}

template <int NL>
__device__ __forceinline__ void mac64(uint64_t* acc, const uint64_t* m, uint64_t y) {
  // acc += m * y (mod 2^(64 NL)), y any 64-bit value
  uint64_t carry = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    uint64_t lo = m[l] * y;
    uint64_t hi = gl::umulhi(m[l], y);
    uint64_t s = acc[l] + lo;
    hi += (s < lo);
    uint64_t s2 = s + carry;
    hi += (s2 < s);
    acc[l] = s2;
    carry = hi;
  }
}

template <int NL>
__device__ __forceinline__ void sub_n(uint64_t* a, const uint64_t* b) {
  uint64_t br = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    uint64_t d = a[l] - b[l];
    uint64_t b1 = a[l] < b[l];
    uint64_t d2 = d - br;
    uint64_t b2 = d < br;
    a[l] = d2;
    br = b1 | b2;
  }
}
template <int NL>
__device__ __forceinline__ void add_n(uint64_t* a, const uint64_t* b) {
  uint64_t c = 0;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    uint64_t s = a[l] + b[l];
    uint64_t c1 = s < a[l];
    uint64_t s2 = s + c;
    uint64_t c2 = s2 < s;
    a[l] = s2;
    c = c1 | c2;
  }
}
template <int NL>
__device__ __forceinline__ void neg_n(uint64_t* a) {
  uint64_t c = 1;
#pragma unroll
  for (int l = 0; l < NL; ++l) {
    uint64_t v = ~a[l] + c;
    c = (c && v == 0) ? 1 : 0;
    a[l] = v;
  }
}
// a >= b ?
template <int NL>
__device__ __forceinline__ bool geq_n(const uint64_t* a, const uint64_t* b) {
#pragma unroll
  for (int l = NL - 1; l >= 0; --l) {
    if (a[l] != b[l]) return a[l] > b[l];
  }
  return true;
}

// Quotient and remainder of a (LA = LW+1 limbs, only the low `lin` of them
// possibly nonzero) by q (LQ limbs) with the generalised Barrett reciprocal
// mu = floor(2^(64 LA) / q). `lin` is the same for every thread of a launch
// (it comes from the ring's bounds), so the skipped limbs are uniform
// branches. The partial products a_i mu_k with i + k < LA - 2 are left out:
// together they are below 2^(64 LA - 56), so the estimate
// floor(a*mu / 2^(64 LA)) falls short of the true quotient by at most 2 and
// at most two corrections follow.
template <int LQ>
__device__ void divmod_q(const uint64_t* a, const RingDev& rd, uint64_t* quo, uint64_t* rem, int lin) {
  constexpr int LW = 2 * LQ + 1;
  constexpr int LA = LW + 1;
  constexpr int LMU = LW + 2 - LQ;
  uint64_t prod[LA + LMU];
#pragma unroll
  for (int l = 0; l < LA + LMU; ++l) prod[l] = 0;
#pragma unroll
  for (int i = 0; i < LA; ++i) {
    if (i >= lin) break;
    uint64_t carry = 0;
#pragma unroll
    for (int k = 0; k < LMU; ++k) {
      if (i + k < LA - 2) continue;
      uint64_t lo = a[i] * rd.mu[k];
      uint64_t hi = gl::umulhi(a[i], rd.mu[k]);
      uint64_t s = prod[i + k] + lo;
      hi += (s < lo);
      uint64_t s2 = s + carry;
      hi += (s2 < s);
      prod[i + k] = s2;
      carry = hi;
    }
    prod[i + LMU] = carry;
  }
  // the quotient has at most lin - LQ + 1 nonzero limbs
  const int nq = lin - LQ + 1;
#pragma unroll
  for (int k = 0; k < LMU; ++k) quo[k] = k < nq ? prod[LA + k] : 0;
  // r = a - quo*q over LQ+1 limbs (the true remainder is < 3q)
  uint64_t qq[LQ + 1];
#pragma unroll
  for (int l = 0; l <= LQ; ++l) qq[l] = 0;
#pragma unroll
  for (int k = 0; k < LMU; ++k) {
    if (k >= nq) break;
    uint64_t carry = 0;
#pragma unroll
    for (int l = 0; l <= LQ; ++l) {
      if (k + l > LQ) break;
      uint64_t ql = l < LQ ? rd.q[l] : 0;
      uint64_t lo = quo[k] * ql;
      uint64_t hi = gl::umulhi(quo[k], ql);
      uint64_t s = qq[k + l] + lo;
      hi += (s < lo);
      uint64_t s2 = s + carry;
      hi += (s2 < s);
      qq[k + l] = s2;
      carry = hi;
    }
  }
  uint64_t r[LQ + 1];
#pragma unroll
  for (int l = 0; l <= LQ; ++l) r[l] = a[l];
  sub_n<LQ + 1>(r, qq);
  uint64_t qe[LQ + 1];
#pragma unroll
  for (int l = 0; l < LQ; ++l) qe[l] = rd.q[l];
  qe[LQ] = 0;
  uint64_t one[LMU];
#pragma unroll
  for (int k = 0; k < LMU; ++k) one[k] = k == 0;
#pragma unroll
  for (int fix = 0; fix < 2; ++fix) {
    if (geq_n<LQ + 1>(r, qe)) {
      sub_n<LQ + 1>(r, qe);
      add_n<LMU>(quo, one);
    }
  }
#pragma unroll
  for (int l = 0; l < LQ; ++l) rem[l] = r[l];
}

// 2r > q ?
template <int LQ>
__device__ __forceinline__ bool twice_gt_q(const uint64_t* r, const uint64_t* q) {
  uint64_t d[LQ + 1];
  uint64_t c = 0;
#pragma unroll
  for (int l = 0; l < LQ; ++l) {
    d[l] = (r[l] << 1) | c;
    c = r[l] >> 63;
  }
  d[LQ] = c;
  uint64_t qe[LQ + 1];
#pragma unroll
  for (int l = 0; l < LQ; ++l) qe[l] = q[l];
  qe[LQ] = 0;
  // d > q  <=>  !(q >= d)
  return !geq_n<LQ + 1>(qe, d);
}

// The pipelined reconstructions (rows held in registers: k_icrt's two
// coefficients per thread, k_icrt_pipe, k_ks_scale_digits): q of at most two
// limbs, at most ICRT_PIPE_ROWS scratch rows, mod-q and scale-and-round
// epilogues. The host sizes the grids and picks the kernels with it.
#ifndef YH_ICRT_PIPE
#define YH_ICRT_PIPE 1
#endif
constexpr int ICRT_PIPE_ROWS = 12;
__host__ __device__ inline bool icrt_pipelined(int LQ, int RS, int mode) {
  return YH_ICRT_PIPE && LQ <= 2 && RS + 1 <= ICRT_PIPE_ROWS && (mode == EPI_MODQ || mode == EPI_SCALE);
}

struct IcrtArgs {
  const uint32_t* Y;     // [pair][R_S][N]
  int N;                 // row length of Y
  int ncoef;             // coefficients reconstructed per pair
  int set;
  int mode;              // Epilogue
  uint64_t* out;         // [pair][out_limbs][out_stride]
  int out_limbs;
  size_t out_stride;     // elements between limbs
  size_t out_pstride;    // elements between pairs
  int positive;          // 1: result in [0, M) (plain product); 0: centred
  uint64_t* const* out_list;  // optional per-pair output addresses (replace out + pair*out_pstride)
  // 1: pairs are taken from the last to the first. A producer that wrote the
  // rows from the first pair to the last leaves the last ones in the
  // memory-side cache (the rows of a large batch exceed it), so reading them
  // first turns part of the reads into cache hits.
  int reverse;
};

// 32-bit word w of a little-endian limb array.
__device__ __forceinline__ uint32_t word_of(const uint64_t* a, int w) {
  return (uint32_t)(a[w >> 1] >> ((w & 1) * 32));
}

// A (NA 32-bit words) += y * X * 2^(32 OFF), X of NX words; the callers'
// bounds keep the sum below 2^(32 NA).
template <int NX, int NA, int OFF>
__device__ __forceinline__ void mac_words(uint32_t* A, const uint64_t* X, uint32_t y) {
  uint32_t c = 0;
#pragma unroll
  for (int w = 0; w < NX; ++w) {
    const uint64_t t = (uint64_t)y * word_of(X, w) + A[w + OFF] + c;  // < 2^64
    A[w + OFF] = (uint32_t)t;
    c = (uint32_t)(t >> 32);
  }
#pragma unroll
  for (int w = NX + OFF; w < NA; ++w) {
    const uint64_t t = (uint64_t)A[w] + c;
    A[w] = (uint32_t)t;
    c = (uint32_t)(t >> 32);
  }
}

// One pass over the moduli of set `set` for one coefficient: the 64-bit y of
// modulus 0, the rounded quotient k = round(sum_i y_i / m_i) of the
// reconstruction c = sum_i y_i M_i - k M (|c| <= M/4: round to nearest), and
// for each of the NF tables X[f] ([R][LQ] constants below 2^(64 LQ)) the
// column sums S[f][w] = sum_{i>=1} y_i X[f]_i[w] over the 32-bit words w
// (each < 2^64 because sum_i p_i < 2^32, which the prime choice guarantees).
// The y rows are loaded eight at a time so the loads overlap.
template <int LQ, int NF>
__device__ __forceinline__ uint32_t icrt_pass(const RingDev& rd, const uint32_t* y, size_t ystride, int set,
                                              const uint64_t* const* X, uint64_t (&S)[NF][2 * LQ], uint64_t& yP) {
  constexpr int NX = 2 * LQ;
  const int RS = rd.set_R[set];
  const uint64_t* bar = rd.barrett[set];
  // modulus 0 = P: floor(y 2^64 / P) = y + floor(y EPS / P) is y + umulhi(y,
  // EPS) to within two units of 2^-64 (far inside the quotient's error budget)
  yP = (uint64_t)y[0] | ((uint64_t)y[ystride] << 32);
  uint64_t flo = yP + gl::umulhi(yP, gl::EPS), fhi = 0;
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int w = 0; w < NX; ++w) S[f][w] = 0;
  for (int i0 = 1; i0 < RS; i0 += 8) {
    uint64_t yy[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) yy[k] = i0 + k < RS ? y[(size_t)(i0 + k + 1) * ystride] : 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      const int i = i0 + k;
      if (i >= RS) break;
      // fixed-point y_i/p_i with 64 fractional bits (y_i < p_i keeps it < 2^64)
      const uint64_t fr = yy[k] * bar[i];
      flo += fr;
      fhi += (flo < fr);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const uint64_t* Xi = X[f] + (size_t)i * LQ;
#pragma unroll
        for (int w = 0; w < NX; ++w) S[f][w] += yy[k] * word_of(Xi, w);
      }
    }
  }
  return (uint32_t)(fhi + (flo >> 63));
}

// V = sum_i y_i X_i + k Xk (< 2^(64 LQ + 34)) as LQ + 1 limbs from the column
// sums S of the small-prime terms (icrt_pass): normalise the columns, then
// add the two 32-bit halves of the y of modulus 0 (rows 0 and RS of X: the
// high half's row is 2^32 X_0 mod q, or the scale-and-round split of it) and
// k times Xk with carrying multiply-accumulates.
template <int LQ>
__device__ __forceinline__ void fold_finish(const uint64_t (&S)[2 * LQ], const uint64_t* X, int RS,
                                            const uint64_t* Xk, uint64_t yP, uint32_t k, uint64_t* V) {
  constexpr int NX = 2 * LQ, NA = NX + 2;
  uint32_t A[NA];
  uint64_t c = 0;
#pragma unroll
  for (int w = 0; w < NX; ++w) {
    const uint64_t cur = S[w] + c;  // S[w] < 2^64 - 2^33, c < 2^32
    A[w] = (uint32_t)cur;
    c = cur >> 32;
  }
  A[NX] = (uint32_t)c;
  A[NX + 1] = 0;
  mac_words<NX, NA, 0>(A, X, (uint32_t)yP);
  mac_words<NX, NA, 0>(A, X + (size_t)RS * LQ, (uint32_t)(yP >> 32));
  mac_words<NX, NA, 0>(A, Xk, k);
#pragma unroll
  for (int l = 0; l < LQ + 1; ++l) V[l] = (uint64_t)A[2 * l] | ((uint64_t)A[2 * l + 1] << 32);
}

// V (LQ + 1 limbs) = quo q + rem, rem in [0, q). On rings with fast_div
// (bits(q) >= 64 LQ - 26, so V / q < 2^62 for V < 2^(64 LQ + 35)) one 64-bit
// Barrett step: floor(floor(V / 2^s) qmu / 2^64) with s = bits(q) - 1 and
// qmu = floor(2^(s+64) / q) is the quotient or falls short by at most 2.
// Other rings take the general truncated Barrett.
template <int LQ>
__device__ __forceinline__ void divmod_v(const uint64_t* V, const RingDev& rd, uint64_t* quo, uint64_t* rem) {
  constexpr int LW = 2 * LQ + 1;
  constexpr int LA = LW + 1;
  constexpr int LMU = LW + 2 - LQ;
  if (rd.fast_div) {
    const int o = rd.q_shift;  // bit s of the value lies in limb LQ - 1
    uint64_t vhi = V[LQ - 1] >> o;
    if (o) vhi |= V[LQ] << (64 - o);
    uint64_t qe = gl::umulhi(vhi, rd.qmu);
    uint64_t qq[LQ + 1], r[LQ + 1], qext[LQ + 1];
    uint64_t carry = 0;
#pragma unroll
    for (int l = 0; l < LQ; ++l) {
      const uint64_t lo = qe * rd.q[l];
      uint64_t hi = gl::umulhi(qe, rd.q[l]);
      const uint64_t s = lo + carry;
      hi += (s < lo);
      qq[l] = s;
      carry = hi;
      qext[l] = rd.q[l];
      r[l] = V[l];
    }
    qq[LQ] = carry;
    qext[LQ] = 0;
    r[LQ] = V[LQ];
    sub_n<LQ + 1>(r, qq);  // V - qe q in [0, 3q)
#pragma unroll
    for (int fix = 0; fix < 2; ++fix) {
      if (geq_n<LQ + 1>(r, qext)) {
        sub_n<LQ + 1>(r, qext);
        ++qe;
      }
    }
    quo[0] = qe;
#pragma unroll
    for (int k = 1; k < LMU; ++k) quo[k] = 0;
#pragma unroll
    for (int l = 0; l < LQ; ++l) rem[l] = r[l];
  } else {
    uint64_t a[LA];
#pragma unroll
    for (int l = 0; l < LA; ++l) a[l] = l <= LQ ? V[l] : 0;
    divmod_q<LQ>(a, rd, quo, rem, LQ + 1);
  }
}

// c mod q in [0, q) for the centred reconstruction c (set `set`):
// c = sum y_i M_i - k M  ==  sum y_i (M_i mod q) + k (q - M mod q)  (mod q).
template <int LQ>
__device__ __forceinline__ void icrt_mod_q(const RingDev& rd, const uint32_t* y, size_t ystride, int set,
                                           uint64_t* rem) {
  constexpr int LMU = 2 * LQ + 3 - LQ;
  uint64_t yP, S[1][2 * LQ];
  const uint64_t* X[1] = {rd.Mq[set]};
  const uint32_t k = icrt_pass<LQ, 1>(rd, y, ystride, set, X, S, yP);
  uint64_t v[LQ + 1], quo[LMU];
  fold_finish<LQ>(S[0], rd.Mq[set], rd.set_R[set], rd.negMq[set], yP, k, v);
  divmod_v<LQ>(v, rd, quo, rem);
}

// Scale-and-round epilogue from the column sums of icrt_pass (NF = 2 over
// beta, alpha): res = round(t c / q) mod q (EPI_SCALE in icrt_coef).
template <int LQ>
__device__ __forceinline__ void scale_value(const RingDev& rd, int set, const uint64_t (&S)[2][2 * LQ], uint64_t yP,
                                            uint32_t k, uint64_t (&rem)[LQ]) {
  constexpr int LMU = 2 * LQ + 3 - LQ;
  const int RS = rd.set_R[set];
  uint64_t v[LQ + 1], quo[LMU];
  fold_finish<LQ>(S[0], rd.beta, RS, rd.negbetaM, yP, k, v);
  divmod_v<LQ>(v, rd, quo, rem);  // S' = quo q + rem
  const uint64_t up = twice_gt_q<LQ>(rem, rd.q) ? 1 : 0;
  // result = sum y_i alpha_i + k (q - (alpha_M + 1)) + quo + up  (mod q);
  // quo < 2^62 with fast_div, < 2^99 otherwise: the sum fits LQ + 1 limbs
  fold_finish<LQ>(S[1], rd.alpha, RS, rd.negalpha1, yP, k, v);
  uint64_t add[LQ + 1];
#pragma unroll
  for (int l = 0; l < LQ + 1; ++l) add[l] = quo[l];
  add_n<LQ + 1>(v, add);
#pragma unroll
  for (int l = 0; l < LQ + 1; ++l) add[l] = l == 0 ? up : 0;
  add_n<LQ + 1>(v, add);
  divmod_v<LQ>(v, rd, quo, rem);
}

template <int LQ>
__device__ __forceinline__ void scale_finish(const RingDev& rd, int set, const uint64_t (&S)[2][2 * LQ], uint64_t yP,
                                             uint32_t k, uint64_t* o, size_t out_stride) {
  uint64_t rem[LQ];
  scale_value<LQ>(rd, set, S, yP, k, rem);
#pragma unroll
  for (int l = 0; l < LQ; ++l) o[(size_t)l * out_stride] = rem[l];
}

// icrt_pass over y rows already in registers (yv[r] = row r, rows 0..RS,
// RS + 1 <= MR): the loop bound is the compile-time MR, so the rows stay in
// registers (the pipelined reconstructions).
template <int LQ, int NF, int MR>
__device__ __forceinline__ uint32_t icrt_pass_regs(const RingDev& rd, const uint32_t (&yv)[MR], int set,
                                                   const uint64_t* const* X, uint64_t (&S)[NF][2 * LQ],
                                                   uint64_t& yP) {
  constexpr int NX = 2 * LQ;
  const int RS = rd.set_R[set];
  const uint64_t* bar = rd.barrett[set];
  yP = (uint64_t)yv[0] | ((uint64_t)yv[1] << 32);
  uint64_t flo = yP + gl::umulhi(yP, gl::EPS), fhi = 0;
#pragma unroll
  for (int f = 0; f < NF; ++f)
#pragma unroll
    for (int w = 0; w < NX; ++w) S[f][w] = 0;
#pragma unroll
  for (int i = 1; i + 1 < MR; ++i) {
    if (i < RS) {
      const uint64_t yi = yv[i + 1];
      const uint64_t fr = yi * bar[i];
      flo += fr;
      fhi += (flo < fr);
#pragma unroll
      for (int f = 0; f < NF; ++f) {
        const uint64_t* Xi = X[f] + (size_t)i * LQ;
#pragma unroll
        for (int w = 0; w < NX; ++w) S[f][w] += yi * word_of(Xi, w);
      }
    }
  }
  return (uint32_t)(fhi + (flo >> 63));
}

// Decryption epilogue of one coefficient: m = round(t [g]_q / q) mod t with
// [g]_q in [0, q) the reconstruction mod q (src/yashe.cpp:167-197, applied to
// every coefficient); o receives m.
template <int LQ>
__device__ __forceinline__ void decrypt_coef(const RingDev& rd, const uint32_t* y, size_t ystride, int set,
                                             uint64_t* o) {
  constexpr int LW = 2 * LQ + 1;
  constexpr int LA = LW + 1;
  constexpr int LMU = LW + 2 - LQ;
  uint64_t rem[LQ], a[LA], quo[LMU];
  icrt_mod_q<LQ>(rd, y, ystride, set, rem);
#pragma unroll
  for (int l = 0; l < LA; ++l) a[l] = 0;
  uint64_t carry = 0;
#pragma unroll
  for (int l = 0; l < LQ; ++l) {
    uint64_t lo = rem[l] * rd.t, hi = gl::umulhi(rem[l], rd.t);
    uint64_t s = lo + carry;
    hi += (s < lo);
    a[l] = s;
    carry = hi;
  }
  a[LQ] = carry;
  divmod_q<LQ>(a, rd, quo, rem, LQ + 1);
  if (twice_gt_q<LQ>(rem, rd.q)) {
    uint64_t one[LMU];
#pragma unroll
    for (int k2 = 0; k2 < LMU; ++k2) one[k2] = k2 == 0;
    add_n<LMU>(quo, one);
  }
  // quo <= t here
  o[0] = quo[0] == rd.t ? 0 : quo[0];
}

// ICRT of one coefficient plus its epilogue. y points at the operand's first
// scratch row (rows ystride apart, layout of store_y), o at limb 0 of the
// output coefficient. EPI_RAW writes the reconstructed integer itself (in
// [0, M) when `positive`, else centred); the other epilogues need only
// residues mod q of it and never form it.
template <int LQ>
__device__ __forceinline__ void icrt_coef(const RingDev& rd, const uint32_t* y, size_t ystride, int set, int mode,
                                          int positive, uint64_t* o, int out_limbs, size_t out_stride) {
  constexpr int LW = 2 * LQ + 1;
  constexpr int LA = LW + 1;
  constexpr int LMU = LW + 2 - LQ;
  if (mode == EPI_MODQ) {
    uint64_t rem[LQ];
    icrt_mod_q<LQ>(rd, y, ystride, set, rem);
#pragma unroll
    for (int l = 0; l < LQ; ++l) o[(size_t)l * out_stride] = rem[l];
    return;
  }
  if (mode == EPI_DECRYPT) {
    decrypt_coef<LQ>(rd, y, ystride, set, o);
    return;
  }
  if (mode == EPI_SCALE) {
    // round(t c / q) mod q for the signed c (src/ciphertext.cpp:55-63
    // semantics applied to t*c: floor division, non-negative remainder, ties
    // down). With t M_i = q alpha_i + beta_i and t M = q alpha_M + beta_M:
    //   t c = q (sum y_i alpha_i - k alpha_M) + S,  S = sum y_i beta_i - k beta_M,
    // so floor(t c / q) = sum y_i alpha_i - k alpha_M + floor(S / q) and
    // (t c) mod q = S mod q. S' = S + k q = sum y_i beta_i + k (q - beta_M)
    // is non-negative: floor(S / q) = floor(S' / q) - k.
    uint64_t yP, S[2][2 * LQ];
    const uint64_t* X[2] = {rd.beta, rd.alpha};
    const uint32_t k = icrt_pass<LQ, 2>(rd, y, ystride, set, X, S, yP);
    scale_finish<LQ>(rd, set, S, yP, k, o, out_stride);
    return;
  }
  // EPI_RAW: the integer itself.
  const int RS = rd.set_R[set];
  const uint64_t* Mi = rd.Mi[set];
  uint64_t acc[LW];
#pragma unroll
  for (int l = 0; l < LW; ++l) acc[l] = 0;
  const uint64_t yP = (uint64_t)y[0] | ((uint64_t)y[ystride] << 32);
  uint64_t flo = yP + gl::umulhi(yP, gl::EPS), fhi = 0;
  mac64<LW>(acc, Mi, yP);
  for (int i = 1; i < RS; ++i) {
    uint64_t yi = y[(size_t)(i + 1) * ystride];
    uint64_t f = yi * rd.barrett[set][i];
    flo += f;
    fhi += (flo < f);
    mac32<LW>(acc, Mi + (size_t)i * LW, yi);
  }
  uint64_t k;
  if (positive) {
    // bias covers the truncation error so that floor() never undershoots;
    // a one-too-large k leaves a negative value that the check below fixes.
    uint64_t bias = (uint64_t)RS << 41;
    uint64_t nl = flo + bias;
    fhi += (nl < flo);
    k = fhi;
  } else {
    k = fhi + (flo >> 63);  // round to nearest: |g| <= M/4
  }
  uint64_t kM[LW];
#pragma unroll
  for (int l = 0; l < LW; ++l) kM[l] = 0;
  mac32<LW>(kM, rd.M[set], k);
  sub_n<LW>(acc, kM);
  bool neg = acc[LW - 1] >> 63;
  if (positive) {
    if (neg) add_n<LW>(acc, rd.M[set]);
  } else if (neg) {
    neg_n<LW>(acc);
  }
#pragma unroll
  for (int l = 0; l < LW; ++l)
    if (l < out_limbs) o[(size_t)l * out_stride] = acc[l];
  for (int l = LW; l < out_limbs; ++l) o[(size_t)l * out_stride] = 0;
}

// The raw integer of a reconstruction (plain products, EPI_RAW) and every
// other epilogue through icrt_coef (the host sends mod q, scale-and-round and
// decryption to k_icrt_pipe and k_icrt_epi, whose registers are those of one
// epilogue).
template <int LQ>
__global__ void __launch_bounds__(YH_ICRT_BLOCK) k_icrt(RingDev rd, IcrtArgs ia) {
  const int pair = ia.reverse ? gridDim.y - 1 - blockIdx.y : blockIdx.y;
  const int RS = rd.set_R[ia.set];
  uint64_t* o = ia.out_list ? ia.out_list[pair] : ia.out + pair * ia.out_pstride;
  const uint32_t* Y = ia.Y + (size_t)pair * y_rows(RS) * ia.N;
#pragma unroll
  for (int u = 0; u < YH_ICRT_COEFS; ++u) {
    const int j = (blockIdx.x * YH_ICRT_COEFS + u) * blockDim.x + threadIdx.x;
    if (j < ia.ncoef)
      icrt_coef<LQ>(rd, Y + j, ia.N, ia.set, ia.mode, ia.positive, o + j, ia.out_limbs, ia.out_stride);
  }
}

// Digit d of ciphertext ct (32-bit words w of its 16 coefficients) centred,
// twisted and transformed into the slot-major digit layout of k_ks_digits.
template <int LOGN>
__device__ __forceinline__ void ks_digit_transform(const RingDev& rd, const KsArgs& ka, int ct, int d,
                                                   const uint32_t (&w)[E], bool after_other = false) {
  extern __shared__ uint64_t lds[];
  constexpr int N = 1 << LOGN;
  constexpr int stride = N >> 4;
  uint64_t x[E];
#pragma unroll
  for (int c = 0; c < E; ++c) x[c] = twist_slot_digit(w[c], 1ull << 31, c);
  if (after_other) __syncthreads();  // every thread has left the previous digit's LDS
  const int g = opaque_tid(x[E - 1]);
  ntt16_fwd<LOGN, true>(x, lds, rd.tab.fwd, g, rd.tab.fwd0);
  // the digit's transform through a buffer resource over its n words
  const __amdgpu_buffer_rsrc_t o = words_rsrc(ka.D + ((size_t)ct * rd.lw + d) * N, (uint32_t)N * 8u);
#pragma unroll
  for (int c = 0; c < E; ++c) word_store(o, (uint32_t)g * 8u, (uint32_t)(c * stride) * 8u, x[c]);
}


// Coefficients C, C + 1, ... E - 1 of k_ks_scale_digits' reconstruction:
// round(t c / q) mod q from the rows in rows[C mod (AH + 1)] into the words
// w[.][C], after requesting the rows of coefficient C + AH. Recursion on a
// compile-time C keeps every index static, so w and rows stay in registers.
template <int LOGN, int LQ, int AH, int C>
__device__ __forceinline__ void scale_words(const RingDev& rd, __amdgpu_buffer_rsrc_t Y, uint32_t voff, int nrow,
                                            uint32_t (&rows)[AH + 1][ICRT_PIPE_ROWS], uint32_t (&w)[2 * LQ][E]) {
  constexpr int N = 1 << LOGN;
  constexpr int stride = N >> 4;
  constexpr int MR = ICRT_PIPE_ROWS;
  if constexpr (C + AH < E) {
#pragma unroll
    for (int r = 0; r < MR; ++r)
      rows[(C + AH) % (AH + 1)][r] = r < nrow ? row_load(Y, voff, (uint32_t)(r * N + (C + AH) * stride) * 4u) : 0u;
  }
  const uint64_t* X[2] = {rd.beta, rd.alpha};
  uint64_t yP, S[2][2 * LQ], rem[LQ];
  const uint32_t k = icrt_pass_regs<LQ, 2, MR>(rd, rows[C % (AH + 1)], 0, X, S, yP);
  scale_value<LQ>(rd, 0, S, yP, k, rem);
#pragma unroll
  for (int v = 0; v < 2 * LQ; ++v) w[v][C] = (uint32_t)(rem[v >> 1] >> (32 * (v & 1)));
  if constexpr (C + 1 < E) scale_words<LOGN, LQ, AH, C + 1>(rd, Y, voff, nrow, rows, w);
}

// The YASHE multiply's scale-and-round fused with the keyswitch digit
// transforms (mul+keyswitch, q of at most two limbs, w = 2^32): one
// workgroup per ciphertext. Thread g reconstructs round(t c / q) mod q for
// the coefficients g + c n/16 straight from the product's rows -- exactly
// the 16 elements it holds in every digit transform -- keeps their 32-bit
// words (the digits) in registers and transforms one digit after the other
// through the same LDS. The scaled product is never written to memory and
// never read back, and one reconstruction launch disappears. Ciphertexts go
// from the last to the first (k_prod wrote the rows from the first to the
// last) and each keeps its XCD, that of k_ks_mac's workgroups that read its
// transforms.
template <int LOGN, int LQ>
__global__ void __launch_bounds__((1 << LOGN) / 16, YH_KSSD_WAVES) k_ks_scale_digits(RingDev rd, KsArgs ka) {
  constexpr int N = 1 << LOGN;
  constexpr int stride = N >> 4;
  constexpr int MR = ICRT_PIPE_ROWS;
  constexpr int NW = 2 * LQ;  // 32-bit words of round(t c / q) mod q, at least l_w
  int ct, unit;
  if (!xcd_map_rev(blockIdx.x, 1, ka.batch, ct, unit)) return;
  const int g = threadIdx.x;
  const int nrow = y_rows(rd.set_R[0]);
  const __amdgpu_buffer_rsrc_t Y = rows_rsrc(ka.P + (size_t)ct * nrow * N, (uint32_t)(nrow * N) * 4u);
  const uint32_t voff = (uint32_t)g * 4u;
  uint32_t w[NW][E];
  // rows of coefficient g + c stride; the rows of the next AH coefficients
  // are requested before the current one is computed
  constexpr int AH = YH_KSSD_AHEAD;
  uint32_t rows[AH + 1][MR];
#pragma unroll
  for (int c = 0; c < AH; ++c)
#pragma unroll
    for (int r = 0; r < MR; ++r) rows[c][r] = r < nrow ? row_load(Y, voff, (uint32_t)(r * N + c * stride) * 4u) : 0u;
  scale_words<LOGN, LQ, AH, 0>(rd, Y, voff, nrow, rows, w);
  // one call per digit with a compile-time index, so w stays in registers
  ks_digit_transform<LOGN>(rd, ka, ct, 0, w[0]);
  if (rd.lw > 1) ks_digit_transform<LOGN>(rd, ka, ct, 1, w[1], true);
  if constexpr (NW > 2) {
    if (rd.lw > 2) ks_digit_transform<LOGN>(rd, ka, ct, 2, w[2], true);
    if (rd.lw > 3) ks_digit_transform<LOGN>(rd, ka, ct, 3, w[3], true);
  }
}

// Reconstruction mod q of the pipelined shapes (the keyswitch's last step),
// or with MODE = EPI_SCALE their scale-and-round (a YASHE multiply on its
// own, or a mul+keyswitch batch too small for k_ks_scale_digits): a kernel
// per epilogue, so its registers are those of that epilogue alone (k_icrt
// holds every epilogue and the raw integer) and more waves stay
// resident; YH_ICRT_PIPE_COEFS coefficients per thread, all their rows
// requested before the first is computed. Workgroups go from the last pair
// to the first when ia.reverse is set. Two coefficients at six requested
// waves take 66 VGPRs (seven resident) and ran 0.126 -> 0.087 ms per launch
// at n = 4096, batch 2048; eight waves spill 2 dwords and four coefficients
// (four waves, 118 VGPRs) ran 0.118 ms.
#ifndef YH_ICRT_PIPE_COEFS
#define YH_ICRT_PIPE_COEFS 2
#endif
#ifndef YH_ICRT_PIPE_WAVES
#define YH_ICRT_PIPE_WAVES 6
#endif
template <int LQ, int MODE = EPI_MODQ>
__global__ void __launch_bounds__(YH_ICRT_BLOCK) __attribute__((amdgpu_waves_per_eu(YH_ICRT_PIPE_WAVES)))
k_icrt_pipe(RingDev rd, IcrtArgs ia) {
  constexpr int MR = ICRT_PIPE_ROWS;
  constexpr int K = YH_ICRT_PIPE_COEFS;
  constexpr int LMU = 2 * LQ + 3 - LQ;
  const int pair = ia.reverse ? gridDim.y - 1 - blockIdx.y : blockIdx.y;
  const int RS = rd.set_R[ia.set];
  const int nrow = y_rows(RS);
  uint64_t* o = ia.out_list ? ia.out_list[pair] : ia.out + pair * ia.out_pstride;
  const __amdgpu_buffer_rsrc_t Y = rows_rsrc(ia.Y + (size_t)pair * nrow * ia.N, (uint32_t)(nrow * ia.N) * 4u);
  const int j0 = blockIdx.x * K * blockDim.x + threadIdx.x;
  uint32_t rows[K][MR];
  // a block whose coefficients all exist loads without the per-thread guard
  if ((blockIdx.x + 1) * K * (int)blockDim.x <= ia.ncoef) {
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const uint32_t voff = (uint32_t)(j0 + u * (int)blockDim.x) * 4u;
#pragma unroll
      for (int r = 0; r < MR; ++r) rows[u][r] = r < nrow ? row_load(Y, voff, (uint32_t)(r * ia.N) * 4u) : 0u;
    }
  } else {
#pragma unroll
    for (int u = 0; u < K; ++u) {
      const int j = j0 + u * blockDim.x;
#pragma unroll
      for (int r = 0; r < MR; ++r)
        rows[u][r] = (r < nrow && j < ia.ncoef) ? row_load(Y, (uint32_t)j * 4u, (uint32_t)(r * ia.N) * 4u) : 0u;
    }
  }
#pragma unroll
  for (int u = 0; u < K; ++u) {
    const int j = j0 + u * blockDim.x;
    if (j < ia.ncoef) {
      uint64_t yP, rem[LQ];
      if constexpr (MODE == EPI_MODQ) {
        uint64_t S[1][2 * LQ], v[LQ + 1], quo[LMU];
        const uint64_t* X[1] = {rd.Mq[ia.set]};
        const uint32_t k = icrt_pass_regs<LQ, 1, MR>(rd, rows[u], ia.set, X, S, yP);
        fold_finish<LQ>(S[0], rd.Mq[ia.set], RS, rd.negMq[ia.set], yP, k, v);
        divmod_v<LQ>(v, rd, quo, rem);
      } else {
        uint64_t S[2][2 * LQ];
        const uint64_t* X[2] = {rd.beta, rd.alpha};
        const uint32_t k = icrt_pass_regs<LQ, 2, MR>(rd, rows[u], ia.set, X, S, yP);
        scale_value<LQ>(rd, ia.set, S, yP, k, rem);
      }
#pragma unroll
      for (int l = 0; l < LQ; ++l) o[j + (size_t)l * ia.out_stride] = rem[l];
    }
  }
}

// Reconstructions of q of more than two limbs (or more rows than the
// pipelined path holds) with one epilogue each: mod q (EPI_MODQ),
// scale-and-round (EPI_SCALE) or decryption (EPI_DECRYPT, any q), one
// coefficient per thread, the rows loaded
// eight at a time (icrt_pass). Kernels of their own for the same reason as
// k_icrt_pipe: k_icrt compiles every epilogue and the raw integer into one
// register allocation.
template <int LQ, int MODE>
__global__ void __launch_bounds__(YH_ICRT_BLOCK) k_icrt_epi(RingDev rd, IcrtArgs ia) {
  const int pair = ia.reverse ? gridDim.y - 1 - blockIdx.y : blockIdx.y;
  const int RS = rd.set_R[ia.set];
  uint64_t* o = ia.out_list ? ia.out_list[pair] : ia.out + pair * ia.out_pstride;
  const uint32_t* Y = ia.Y + (size_t)pair * y_rows(RS) * ia.N;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ia.ncoef) return;
  if constexpr (MODE == EPI_MODQ) {
    uint64_t rem[LQ];
    icrt_mod_q<LQ>(rd, Y + j, ia.N, ia.set, rem);
#pragma unroll
    for (int l = 0; l < LQ; ++l) o[j + (size_t)l * ia.out_stride] = rem[l];
  } else if constexpr (MODE == EPI_DECRYPT) {
    decrypt_coef<LQ>(rd, Y + j, ia.N, ia.set, o + j);
  } else {
    uint64_t yP, S[2][2 * LQ];
    const uint64_t* X[2] = {rd.beta, rd.alpha};
    const uint32_t k = icrt_pass<LQ, 2>(rd, Y + j, ia.N, ia.set, X, S, yP);
    scale_finish<LQ>(rd, ia.set, S, yP, k, o + j, ia.out_stride);
  }
}

// ---------------------------------------------------------------- primitives

// CRT primitive: residues r[i][j] = coef_j mod p_i in [0, p_i), layout
// [R][spacing] like the reference's d_polyCRT (src/polynomial.h:1435).
static __global__ void k_crt(RingDev rd, const uint64_t* A, int lin, int ncoef, size_t lstride, uint64_t* out,
                      int spacing) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= spacing) return;
  uint64_t r = 0;
  if (j < ncoef)
    r = i == 0 ? crt_residue_P(A + j, lin, lstride)
               : crt_residue(A + j, lin, lstride, rd.pow32[0] + (size_t)i * 2 * rd.LW, rd.sp[0][i]);
  out[(size_t)i * spacing + j] = r;
}

// The reference's multi-word integer (bn_t, src/cuda_bn.h:18-23): words
// dp[0..used) in device memory, least significant first. Same layout as
// yh_bn_t of the C ABI.
struct BnT {
  int alloc;
  int used;
  int sign;
  uint64_t* dp;
};

// CRT of coefficients held as BnT (callCRT's input, src/cuda_bn.cu:970-1002):
// Horner steps over the words from the top, each r <- (r 2^64 + w) mod p_i
// with 2^64 and 2^32 mod p_i from the ring's table (any word count).
static __global__ void k_crt_bn(RingDev rd, const BnT* C, int ncoef, uint64_t* out, int spacing) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= spacing) return;
  uint64_t r = 0;
  if (j < ncoef) {
    const BnT c = C[j];
    if (i == 0) {
      for (int l = c.used - 1; l >= 0; --l) r = gl::reduce128(r, c.dp[l]);
    } else {
      const uint64_t* pw = rd.pow32[0] + (size_t)i * 2 * rd.LW;  // pw[k] = 2^(32k) mod p_i
      const SmallPrime& k = rd.sp[0][i];
      const uint32_t c32 = (uint32_t)pw[1], c64 = (uint32_t)pw[2];
      uint32_t acc = 0;
      for (int l = c.used - 1; l >= 0; --l) {
        const uint64_t w = c.dp[l];
        // acc c64 < 2^56, hi c32 < 2^60, lo < 2^32: the sum stays below 2^63
        acc = mod_small((uint64_t)acc * c64 + (uint64_t)(uint32_t)(w >> 32) * c32 + (uint32_t)w, k);
      }
      r = acc;
    }
  }
  out[(size_t)i * spacing + j] = r;
}

// Coefficients [LW][ncoef] (limb-major) into BnT (callICRT's output,
// src/cuda_bn.cu:1004-1045): the words that fit dp (alloc), used = the
// significant words, sign positive.
static __global__ void k_to_bn(const uint64_t* V, int lw, int ncoef, BnT* C) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncoef) return;
  BnT c = C[j];
  const int w = c.alloc < lw ? c.alloc : lw;
  int used = 0;
  for (int l = 0; l < w; ++l) {
    const uint64_t v = V[(size_t)l * ncoef + j];
    c.dp[l] = v;
    if (v) used = l + 1;
  }
  C[j].used = used;
  C[j].sign = 0;
}

// y_i = r_i * (M/p_i)^-1 mod p_i (the cuPreICRT inner step, src/cuda_bn.cu:875-879).
static __global__ void k_residue_to_y(RingDev rd, const uint64_t* res, int spacing, uint32_t* Y, int set) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= spacing) return;
  const uint64_t p = rd.primes[set][i];
  uint64_t r = res[(size_t)i * spacing + j];
  if (i == 0) {
    if (r >= gl::P) r -= gl::P;
    store_y(Y, spacing, 0, j, gl::mul(r, rd.invMi[set][0]));
  } else {
    r = mod_p(r, p, rd.barrett[set][i]);
    store_y(Y, spacing, i, j, mod_p(r * rd.invMi[set][i], p, rd.barrett[set][i]));
  }
}

// Residue-domain add/sub mod p_i (the reference's polynomialAddSub,
// src/cuda_functions.cu:99-126, with the mod p_i it omits).
static __global__ void k_residue_addsub(RingDev rd, uint64_t* C, const uint64_t* A, const uint64_t* B, int spacing, int sub) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= spacing) return;
  const uint64_t p = rd.primes[0][i];
  size_t o = (size_t)i * spacing + j;
  uint64_t a = A[o], b = B[o];
  uint64_t r;
  if (i == 0) r = sub ? gl::sub(a, b) : gl::add(a, b);  // residues mod P itself
  else if (sub) r = a >= b ? a - b : a + p - b;
  else { r = a + b; if (r >= p) r -= p; }
  C[o] = r;
}

// Pointwise Goldilocks product (polynomialNTTMul, src/cuda_functions.cu:460-474).
static __global__ void k_pointwise(uint64_t* A, const uint64_t* B, size_t count) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < count) A[t] = gl::mul(A[t], B[t]);
}

// Standalone cyclic transform over global data, natural order in and out
// (callNTT, src/cuda_functions.cu:447-458, with N^-1 scaling on the inverse):
// the radix-16 transforms of ntt16.hpp with the cyclic tables. The forward
// transform leaves X_k at position bitrev(k) (position 16 g + c in thread g's
// slot c) and the inverse takes that layout, so both reorder through the LDS.
__device__ __forceinline__ int bitrev(int v, int bits) { return (int)(__brev((unsigned)v) >> (32 - bits)); }

template <int LOGN>
__global__ void __launch_bounds__((1 << LOGN) / 16, YH_MIN_WAVES) k_ntt_cyclic(uint64_t* data, NttTab tab,
                                                                              uint64_t Ninv, int inverse) {
  extern __shared__ uint64_t lds[];
  constexpr int N = 1 << LOGN;
  constexpr int stride = N >> 4;
  uint64_t* d = data + (size_t)blockIdx.x * N;
  const int g = threadIdx.x;
  uint64_t x[E];
  if (!inverse) {
#pragma unroll
    for (int c = 0; c < E; ++c) x[c] = d[g + c * stride];
    ntt16_fwd<LOGN, false, false>(x, lds, tab.fwd, g);
    __syncthreads();
#pragma unroll
    for (int c = 0; c < E; ++c) lds[lds_pad(g * E + c)] = x[c];
    __syncthreads();
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int k = g + c * stride;
      d[k] = lds[lds_pad(bitrev(k, LOGN))];
    }
  } else {
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int k = g + c * stride;
      lds[lds_pad(bitrev(k, LOGN))] = d[k];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < E; ++c) x[c] = lds[lds_pad(g * E + c)];
    __syncthreads();
    ntt16_inv<LOGN, false, false>(x, lds, tab.inv, g);
#pragma unroll
    for (int c = 0; c < E; ++c) d[g + c * stride] = gl::mul(x[c], Ninv);
  }
}

// Coefficient-domain add/sub mod q for ring elements in [0, q) (homomorphic
// addition on resident ciphertexts).
// Optional per-element address lists (device arrays of device pointers)
// replace A/B/C + element*LQ*n.
struct AddsubLists {
  const uint64_t* const* a;
  const uint64_t* const* b;
  uint64_t* const* c;
};

template <int LQ>
__global__ void k_addsub_q(RingDev rd, const uint64_t* A, const uint64_t* B, uint64_t* C, int n, size_t count, int sub,
                           AddsubLists lists) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  size_t pr = t / n, j = t % n;
  const uint64_t* a = (lists.a ? lists.a[pr] : A + pr * LQ * n) + j;
  const uint64_t* b = (lists.b ? lists.b[pr] : B + pr * LQ * n) + j;
  uint64_t* c = (lists.c ? lists.c[pr] : C + pr * LQ * n) + j;
  uint64_t x[LQ + 1], y[LQ + 1], qe[LQ + 1];
#pragma unroll
  for (int l = 0; l < LQ; ++l) { x[l] = a[(size_t)l * n]; y[l] = b[(size_t)l * n]; qe[l] = rd.q[l]; }
  x[LQ] = y[LQ] = qe[LQ] = 0;
  if (!sub) {
    add_n<LQ + 1>(x, y);
    if (geq_n<LQ + 1>(x, qe)) sub_n<LQ + 1>(x, qe);
  } else {
    if (!geq_n<LQ + 1>(x, y)) add_n<LQ + 1>(x, qe);
    sub_n<LQ + 1>(x, y);
  }
#pragma unroll
  for (int l = 0; l < LQ; ++l) c[(size_t)l * n] = x[l];
}

// ---------------------------------------------------------------- scalar operations
// The reference's polynomialOPInteger (src/cuda_functions.cu:477-548): with
// op ADD / SUB the scalar is added to / subtracted from coefficient 0 only,
// with MUL every coefficient is multiplied by it (add_mode_t values,
// src/settings.h:36).
enum ScalarOp { SC_ADD = 0, SC_SUB = 1, SC_MUL = 2 };

// Scalar s mod q as LQ limbs (k_scalar_q) or its residues mod each CRT
// modulus (k_residue_scalar), passed by value.
struct ScalarArg {
  uint64_t v[64];
};

// Ring elements [batch][LQ][n] in [0, q): c = a (+,-) s on coefficient 0
// (the others copied), or c = a * s mod q on every coefficient; s in [0, q).
// The product of two LQ-limb values is reduced with the ring's Barrett
// reciprocal (divmod_q over its 2 LQ limbs). c may alias a.
template <int LQ>
__global__ void k_scalar_q(RingDev rd, const uint64_t* A, uint64_t* C, int n, size_t count, int op, ScalarArg sa) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= count) return;
  const size_t pr = t / n, j = t % n;
  const uint64_t* a = A + pr * LQ * n + j;
  uint64_t* c = C + pr * LQ * n + j;
  if (op == SC_MUL) {
    constexpr int LA = 2 * LQ + 2, LMU = LQ + 3;
    uint64_t x[LA], quo[LMU], rem[LQ];
#pragma unroll
    for (int l = 0; l < LA; ++l) x[l] = 0;
#pragma unroll
    for (int i = 0; i < LQ; ++i) {
      const uint64_t ai = a[(size_t)i * n];
      uint64_t carry = 0;
#pragma unroll
      for (int k = 0; k < LQ; ++k) {
        const uint64_t lo = ai * sa.v[k];
        uint64_t hi = gl::umulhi(ai, sa.v[k]);
        const uint64_t s1 = x[i + k] + lo;
        hi += (s1 < lo);
        const uint64_t s2 = s1 + carry;
        hi += (s2 < s1);
        x[i + k] = s2;
        carry = hi;
      }
      x[i + LQ] = carry;
    }
    divmod_q<LQ>(x, rd, quo, rem, 2 * LQ);
#pragma unroll
    for (int l = 0; l < LQ; ++l) c[(size_t)l * n] = rem[l];
    return;
  }
  uint64_t x[LQ + 1];
#pragma unroll
  for (int l = 0; l < LQ; ++l) x[l] = a[(size_t)l * n];
  if (j == 0) {
    uint64_t y[LQ + 1], qe[LQ + 1];
#pragma unroll
    for (int l = 0; l < LQ; ++l) { y[l] = sa.v[l]; qe[l] = rd.q[l]; }
    x[LQ] = y[LQ] = qe[LQ] = 0;
    if (op == SC_ADD) {
      add_n<LQ + 1>(x, y);
      if (geq_n<LQ + 1>(x, qe)) sub_n<LQ + 1>(x, qe);
    } else {
      if (!geq_n<LQ + 1>(x, y)) add_n<LQ + 1>(x, qe);
      sub_n<LQ + 1>(x, y);
    }
  }
#pragma unroll
  for (int l = 0; l < LQ; ++l) c[(size_t)l * n] = x[l];
}

// Residues [R][spacing] (modulus i in row i, values in [0, m_i)): the same
// operations mod each CRT modulus, with the scalar's residue
// s_i = S[i] mod m_i (S: device array of R words, the reference's
// integer_array). c may alias a.
static __global__ void k_residue_scalar(RingDev rd, uint64_t* C, const uint64_t* A, int spacing, int op,
                                        const uint64_t* S) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j >= spacing) return;
  const size_t o = (size_t)i * spacing + j;
  const uint64_t a = A[o];
  const uint64_t s = i == 0 ? (S[0] >= gl::P ? S[0] - gl::P : S[0]) : mod_p(S[i], rd.primes[0][i], rd.barrett[0][i]);
  uint64_t r = a;
  if (i == 0) {  // modulus P itself
    if (op == SC_MUL) r = gl::mul(a, s);
    else if (j == 0) r = op == SC_ADD ? gl::add(a, s) : gl::sub(a, s);
  } else {
    const uint64_t p = rd.primes[0][i];
    if (op == SC_MUL) r = mod_p(a * s, p, rd.barrett[0][i]);  // a, s < p < 2^28
    else if (j == 0) r = op == SC_ADD ? (a + s >= p ? a + s - p : a + s) : (a >= s ? a - s : a + p - s);
  }
  C[o] = r;
}

// Reduction mod (x^n + 1, q) of a length-2n coefficient vector
// (Polynomial::reduce, src/cuda_functions.cu:1061-1078, and the % q of
// src/polynomial.h:593-610): out_j = (c_j - c_{j+n}) mod q in [0, q).
// Input [pair][lin][2n], lin <= LW; output [pair][LQ][n].
template <int LQ>
__global__ void k_reduce(RingDev rd, const uint64_t* C, int lin, uint64_t* out) {
  constexpr int LW = 2 * LQ + 1;
  constexpr int LA = LW + 1;
  constexpr int LMU = LW + 2 - LQ;
  const int n = rd.n;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int pair = blockIdx.y;
  if (j >= n) return;
  const uint64_t* c = C + (size_t)pair * lin * 2 * n;
  uint64_t a[LA], b[LA];
#pragma unroll
  for (int l = 0; l < LA; ++l) {
    a[l] = l < lin ? c[(size_t)l * 2 * n + j] : 0;
    b[l] = l < lin ? c[(size_t)l * 2 * n + j + n] : 0;
  }
  sub_n<LA>(a, b);
  bool neg = a[LA - 1] >> 63;
  if (neg) neg_n<LA>(a);
  uint64_t quo[LMU], rem[LQ];
  divmod_q<LQ>(a, rd, quo, rem, lin);
  bool nz = false;
#pragma unroll
  for (int l = 0; l < LQ; ++l) nz |= rem[l] != 0;
  if (neg && nz) {
    uint64_t qv[LQ];
#pragma unroll
    for (int l = 0; l < LQ; ++l) qv[l] = rd.q[l];
    sub_n<LQ>(qv, rem);
#pragma unroll
    for (int l = 0; l < LQ; ++l) rem[l] = qv[l];
  }
  uint64_t* o = out + (size_t)pair * LQ * n + j;
#pragma unroll
  for (int l = 0; l < LQ; ++l) o[(size_t)l * n] = rem[l];
}

// Helpers of the norm-recursion inverse on [LQ][m] ring elements in [0, q).
// g_j = (-1)^j f_j mod q (f(-x))
template <int LQ>
__global__ void k_negate_odd(RingDev rd, const uint64_t* F, uint64_t* G, int m) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  uint64_t v[LQ];
  bool nz = false;
#pragma unroll
  for (int l = 0; l < LQ; ++l) { v[l] = F[(size_t)l * m + j]; nz |= v[l] != 0; }
  if ((j & 1) && nz) {
    uint64_t qv[LQ];
#pragma unroll
    for (int l = 0; l < LQ; ++l) qv[l] = rd.q[l];
    sub_n<LQ>(qv, v);
#pragma unroll
    for (int l = 0; l < LQ; ++l) v[l] = qv[l];
  }
#pragma unroll
  for (int l = 0; l < LQ; ++l) G[(size_t)l * m + j] = v[l];
}

// out_j = in_{2j}, j < m/2 (an even polynomial of degree < m as one in x^2)
static __global__ void k_take_even(const uint64_t* in, uint64_t* out, int m, int LQ) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int h = m / 2;
  if (j >= h) return;
  for (int l = 0; l < LQ; ++l) out[(size_t)l * h + j] = in[(size_t)l * m + 2 * j];
}

// x / 2 mod an odd modulus m for x in [0, m): x/2, or (x + m)/2 when x is odd
__device__ __forceinline__ uint64_t halve_mod(uint64_t x, uint64_t m) { return (x >> 1) + ((x & 1) ? (m >> 1) + 1 : 0); }

// Linear product c = ab of operands of at most h coefficients, assembled in
// the reconstruction rows from its cyclic half u = ab mod (x^h - 1) and its
// negacyclic half v = ab mod (x^h + 1): for j < h, c_j = (u_j + v_j)/2 and
// c_{j+h} = (u_j - v_j)/2 modulo every CRT modulus (each y row is a fixed
// multiple of the residue, so the halves combine row by row), and c_j = 0 for
// j >= 2h. Without v (Yv null) the product ends below h: c_j = u_j for j < h.
// Rows of Yu and Yv hold h values, rows of Yc N; grid.y runs over the moduli.
static __global__ void k_join_halves(RingDev rd, const uint32_t* Yu, const uint32_t* Yv, uint32_t* Yc, int h, int N) {
  const int i = blockIdx.y;
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= N) return;
  const bool low = j < h;
  const int k = low ? j : j - h;
  const bool zero = j >= 2 * h || (!Yv && !low);
  if (i == 0) {
    uint64_t c = 0;
    if (!zero) {
      const uint64_t u = Yu[k] | (uint64_t)Yu[h + k] << 32;
      if (Yv) {
        const uint64_t v = Yv[k] | (uint64_t)Yv[h + k] << 32;
        c = halve_mod(low ? gl::add(u, v) : gl::sub(u, v), gl::P);
      } else {
        c = u;
      }
    }
    Yc[j] = (uint32_t)c;
    Yc[(size_t)N + j] = (uint32_t)(c >> 32);
  } else {
    const uint32_t p = rd.sp[0][i].p;
    uint32_t c = 0;
    if (!zero) {
      const uint32_t u = Yu[(size_t)(i + 1) * h + k];
      if (Yv) {
        const uint32_t v = Yv[(size_t)(i + 1) * h + k];
        uint32_t s = low ? u + v : u + (p - v);  // < 2p < 2^29
        if (s >= p) s -= p;
        c = (uint32_t)halve_mod(s, p);
      } else {
        c = u;
      }
    }
    Yc[(size_t)(i + 1) * N + j] = c;
  }
}

// out_{2j} = in_j, out_{2j+1} = 0, j < m/2 (p(y) -> p(x^2))
static __global__ void k_spread_even(const uint64_t* in, uint64_t* out, int m, int LQ) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int h = m / 2;
  for (int l = 0; l < LQ; ++l) out[(size_t)l * m + j] = (j & 1) ? 0 : in[(size_t)l * h + j / 2];
}

}  // namespace yhk
