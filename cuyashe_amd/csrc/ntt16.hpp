// Radix-16 Goldilocks transforms for the fused product kernels.
//
// A length-N transform (N = 2^LOGN, 16 <= N <= 16384) is held by N/16 threads,
// 16 elements each. Forward = decimation in frequency in passes of four
// radix-2 stages:
//   pass p (b0 = LOGN - 4 - 4p, s = 2^b0): thread g takes the 16 elements
//   base(g) + c*s, runs a 16-point DFT in registers and multiplies output slot
//   c by w_{16s}^(o * brev4(c)), o = g mod s (the inter-pass twiddle).
// A short final pass (radix 2^(LOGN mod 4)) runs on 16 consecutive elements.
// Every root of unity of order <= 64 mod P = 2^64 - 2^32 + 1 is a power of two
// (7^((P-1)/64) = 2^39), so the twiddles inside a 16-, 8-, 4- or 2-point DFT
// are shifts with compile-time amounts; only the inter-pass twiddles are full
// 64x64 products. The inverse runs the passes backwards as decimation in time
// with inverse roots. The output order of the forward (and input order of the
// inverse) is a fixed digit-reversed permutation shared by every operand, so
// pointwise products and the resident keyswitch keys need no reordering.
//
// Negacyclic products (mod x^N + 1) twist the input by psi^j on load and
// untwist by psi^-j N^-1 on the way out (psi^2 = w_N); cyclic products
// (mod x^N - 1) skip the twist. The reference's transform
// (src/cuda_functions.cu:404-458) is a radix-2 Stockham stage per launch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "gl64.hpp"

namespace yhk {

// Device pointers of the tables of one transform length. The inter-pass
// twiddles are stored in Montgomery form (w 2^64 mod P) and applied with
// gl::mont_mul. A pointwise product of two transforms taken with mont_mul
// carries an extra 2^-64, which the *_pw constants cancel.
struct NttTab {
  const uint64_t* fwd;         // inter-pass twiddles, forward (Montgomery form)
  const uint64_t* inv;         // inter-pass twiddles, inverse (Montgomery form)
  const uint64_t* twist;       // psi^j
  const uint64_t* untwist_pw;  // psi^-j N^-1 2^128 (after a mont_mul pointwise product)
  const uint64_t* untwist_m;   // psi^-j N^-1 2^64 (Montgomery form)
  const uint64_t* fwd0;        // first forward pass with the twist folded in: w_N^(brev4(c) g) psi^g
  const uint64_t* inv0_pw;     // last inverse pass with the untwist folded in, 2^128
  const uint64_t* inv0_m;      // the same, 2^64 (Montgomery form)
  uint64_t n_inv_pw;           // N^-1 2^128 (cyclic products)
};

__host__ __device__ constexpr int brev4(int c) {
  return ((c & 1) << 3) | ((c & 2) << 1) | ((c & 4) >> 1) | ((c & 8) >> 3);
}

// Exponent of the primitive r-th root 7^((P-1)/r) as a power of two.
__host__ __device__ constexpr int root_exp(int r) { return (39 * (64 / r)) % 192; }

// Offset of pass p's table: passes with b0 > 0 store 15 * 2^b0 entries.
__host__ __device__ constexpr int pass_offset(int logn, int p) {
  int off = 0;
  for (int q = 0; q < p; ++q) {
    int b0 = logn - 4 - 4 * q;
    if (b0 > 0) off += 15 << b0;
  }
  return off;
}

// x * 2^e mod P for a compile-time e (gl::mul_pow2).
__device__ __forceinline__ uint64_t mul_pow2(uint64_t x, int e) { return gl::mul_pow2(x, e); }

// Loads and stores through pointers that the compiler cannot place in an
// address space (pointers read from memory, such as a ring's tables in the
// persistent batch kernel): the cast makes them global_* instead of flat_*
// instructions. A flat access also counts against the LDS counter, so every
// wait for an LDS exchange would wait for it as well. The pointers must
// address global memory.
template <class T>
__device__ __forceinline__ T ldg(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}
template <class T>
__device__ __forceinline__ void stg(T* p, T v) {
  *(__attribute__((address_space(1))) T*)p = v;
}

// Twiddle-table reads through a buffer resource: the table's base and the
// compile-time element offset sit in scalar registers and the thread's
// offset in one vector register, so the fifteen or sixteen reads of a pass
// take no vector address arithmetic. The resource's range is not used as a
// bound (every offset read lies inside the table).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t table_rsrc(const uint64_t* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint64_t*>(base), (short)0, 0x7FFFFFFF, 0x00020000);
}
__device__ __forceinline__ uint64_t table_load(__amdgpu_buffer_rsrc_t rs, uint32_t voff, uint32_t elem) {
  const auto v = __builtin_amdgcn_raw_buffer_load_b64(rs, voff, elem * 8u, 0);
  return (uint64_t)v[0] | ((uint64_t)v[1] << 32);
}

// a * b mod P for a < 2^32 (the high word of the 96-bit product is < 2^32).
__device__ __forceinline__ uint64_t mul_small(uint32_t a, uint64_t b) {
  const uint64_t p0 = (uint64_t)a * (uint32_t)b;
  const uint64_t p1 = (uint64_t)a * (uint32_t)(b >> 32) + (p0 >> 32);
  return gl::reduce128(p1 >> 32, (p1 << 32) | (uint32_t)p0);
}

// In-register DIF of R slots: natural order in, bit-reversed out.
template <int R>
__device__ __forceinline__ void dif_r(uint64_t* x) {
  constexpr int EW = root_exp(R);
#pragma unroll
  for (int ln = R; ln >= 2; ln >>= 1) {
    const int h = ln / 2;
#pragma unroll
    for (int i = 0; i < R; i += ln) {
#pragma unroll
      for (int j = 0; j < h; ++j) {
        const uint64_t u = x[i + j], v = x[i + j + h];
        x[i + j] = gl::add(u, v);
        // 2^e with e in [96, 192) is -2^(e-96): subtract the other way round
        // instead of negating
        const int e = (EW * (R / ln) * j) % 192;
        x[i + j + h] = e >= 96 ? mul_pow2(gl::sub_full(v, u), e - 96) : mul_pow2(gl::sub_full(u, v), e);
      }
    }
  }
}

// In-register DIT of R slots with inverse roots: bit-reversed in, natural
// out, unscaled.
template <int R>
__device__ __forceinline__ void dit_r(uint64_t* x) {
  constexpr int EWI = (192 - root_exp(R)) % 192;
#pragma unroll
  for (int ln = 2; ln <= R; ln <<= 1) {
    const int h = ln / 2;
#pragma unroll
    for (int i = 0; i < R; i += ln) {
#pragma unroll
      for (int j = 0; j < h; ++j) {
        const uint64_t u = x[i + j];
        const int e = (EWI * (R / ln) * j) % 192;
        if (e >= 96) {  // v = -x 2^(e-96): swap the add and the subtraction
          const uint64_t w = mul_pow2(x[i + j + h], e - 96);
          x[i + j] = gl::sub_full(u, w);
          x[i + j + h] = gl::add(u, w);
        } else {
          const uint64_t v = mul_pow2(x[i + j + h], e);
          x[i + j] = gl::add(u, v);
          x[i + j + h] = gl::sub_full(u, v);
        }
      }
    }
  }
}

__device__ __forceinline__ int lds_pad(int j) { return j + (j >> 4); }
__device__ __forceinline__ int pbase(int g, int b0) { return ((g >> b0) << (b0 + 4)) | (g & ((1 << b0) - 1)); }

__device__ __forceinline__ void st16(uint64_t* lds, const uint64_t* x, int base, int b0) {
#pragma unroll
  for (int c = 0; c < 16; ++c) lds[lds_pad(base + (c << b0))] = x[c];
}
__device__ __forceinline__ void ld16(const uint64_t* lds, uint64_t* x, int base, int b0) {
#pragma unroll
  for (int c = 0; c < 16; ++c) x[c] = lds[lds_pad(base + (c << b0))];
}

// Synchronisation of one transform's threads around its LDS exchanges: the
// whole workgroup, or (wave = true) only the wavefront, whose LDS operations
// execute in order. `wave` must be the same for the whole workgroup.
//
// Which exchanges need the workgroup: in the layout of pass b0 thread g owns
// the 16 positions base(g) + c 2^b0, so the layouts of b0 and b0 - 4 (and the
// final layout 16 g + c, that of b0 = 0) map a block of 2^(b0 + 4) positions
// onto the same 2^b0 threads g >> b0 (B = the larger b0 of the two layouts).
// Threads g are consecutive within a wavefront from a multiple of the
// transform's thread count, so an exchange with B <= 6 stays inside one
// wavefront. And a pass writes back exactly the positions it read, which no
// other thread touches in that pass: the load of a pass needs no barrier.
__device__ __forceinline__ void xsync(bool wave) {
  if (wave) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// Forward. Entry: x holds elements g + c*N/16 (natural positions). Exit:
// x holds the transform at positions 16g + c (permuted order). Leaves the LDS
// in use (other wavefronts may still be reading it): callers synchronise the
// workgroup before anything writes it again.
// g is the thread's index within the transform (threadIdx.x unless several
// transforms share a workgroup, each from a multiple of N/16 threads);
// __syncthreads spans the whole workgroup, so every thread of it calls the
// same transform.
//
// MERGED (negacyclic): the caller has multiplied input slot c by
// 2^(78 c) only, and the first pass multiplies every output slot c by
// tw0[c N/16 + g] = w_N^(brev4(c) g) psi^g (Montgomery form): psi^g is common
// to the thread's 16 inputs (elements g + c N/16, psi^(g + c N/16) = psi^g
// 2^(78 c)) and passes through the 16-point DFT, so the result is the twisted
// transform, bit for bit. (N = 16: one thread, psi^g = 1 and the table holds
// 1; psi itself is 2^78.)
// TBUF: twiddles through a buffer resource (table_load), the default for
// n <= 1024, where it frees registers (the n = 1024 product kernel stopped
// spilling); otherwise plain global loads, which the compiler may move
// across the LDS exchanges: measured faster for n >= 8192 products (2 %),
// the circuit of n = 8192 (4 %) and the stand-alone transform kernel (17 %),
// and even at n = 4096.
template <int LOGN, bool MERGED = false, bool TBUF = (LOGN <= 10)>
__device__ __forceinline__ void ntt16_fwd(uint64_t* x, uint64_t* lds, const uint64_t* __restrict__ tw,
                                          int g = threadIdx.x, const uint64_t* __restrict__ tw0 = nullptr) {
  constexpr int NFULL = LOGN / 4;
  constexpr int SHORT = LOGN % 4;
#pragma unroll
  for (int p = 0; p < NFULL; ++p) {
    const int b0 = LOGN - 4 - 4 * p;
    const int base = pbase(g, b0);
    if (p > 0) ld16(lds, x, base, b0);
    dif_r<16>(x);
    if (MERGED && p == 0) {
      const __amdgpu_buffer_rsrc_t t = table_rsrc(tw0);
      const uint32_t vo = (uint32_t)g * 8u;
#pragma unroll
      for (int c = 0; c < 16; ++c) x[c] = gl::mont_mul(x[c], table_load(t, vo, (uint32_t)(c << b0)));
    } else if (b0 > 0 && TBUF) {
      const int s = 1 << b0;
      const __amdgpu_buffer_rsrc_t t = table_rsrc(tw);
      const uint32_t vo = (uint32_t)(g & (s - 1)) * 8u;
#pragma unroll
      for (int c = 1; c < 16; ++c)
        x[c] = gl::mont_mul(x[c], table_load(t, vo, (uint32_t)(pass_offset(LOGN, p) + (brev4(c) - 1) * s)));
    } else if (b0 > 0) {
      const int s = 1 << b0;
      const uint64_t* t = tw + pass_offset(LOGN, p) + (g & (s - 1));
#pragma unroll
      for (int c = 1; c < 16; ++c) x[c] = gl::mont_mul(x[c], ldg(t + (brev4(c) - 1) * s));
    }
    if (p < NFULL - 1 || SHORT) {
      st16(lds, x, base, b0);
      xsync(b0 <= 6);
    }
  }
  if (SHORT) {
    ld16(lds, x, 16 * g, 0);
    constexpr int R = 1 << SHORT;
#pragma unroll
    for (int blk = 0; blk < 16; blk += R) dif_r<R>(x + blk);
  }
}

// Inverse (unscaled, no untwist). Entry: positions 16g + c as left by
// ntt16_fwd. Exit: x holds elements g + c*N/16. Callers synchronise the
// workgroup before calling if another transform used the LDS just before,
// and after it before anything writes the LDS again.
//
// MERGED (negacyclic): the last pass multiplies every input slot c
// by twi0[c N/16 + g] = w_N^(-brev4(c) g) psi^-g N^-1 (times the caller's
// Montgomery constant) and output slot c (element g + c N/16) by 2^(-78 c):
// the exit values are untwisted and scaled by N^-1.
template <int LOGN, bool MERGED = false, bool TBUF = (LOGN <= 10)>
__device__ __forceinline__ void ntt16_inv(uint64_t* x, uint64_t* lds, const uint64_t* __restrict__ twi,
                                          int g = threadIdx.x, const uint64_t* __restrict__ twi0 = nullptr) {
  constexpr int NFULL = LOGN / 4;
  constexpr int SHORT = LOGN % 4;
  if (SHORT) {
    constexpr int R = 1 << SHORT;
#pragma unroll
    for (int blk = 0; blk < 16; blk += R) dit_r<R>(x + blk);
    st16(lds, x, 16 * g, 0);
    xsync(true);
  }
#pragma unroll
  for (int p = NFULL - 1; p >= 0; --p) {
    const int b0 = LOGN - 4 - 4 * p;
    const int base = pbase(g, b0);
    if (p < NFULL - 1 || SHORT) ld16(lds, x, base, b0);
    if (MERGED && p == 0) {
      const __amdgpu_buffer_rsrc_t t = table_rsrc(twi0);
      const uint32_t vo = (uint32_t)g * 8u;
#pragma unroll
      for (int c = 0; c < 16; ++c) x[c] = gl::mont_mul(x[c], table_load(t, vo, (uint32_t)(c << b0)));
    } else if (b0 > 0 && TBUF) {
      const int s = 1 << b0;
      const __amdgpu_buffer_rsrc_t t = table_rsrc(twi);
      const uint32_t vo = (uint32_t)(g & (s - 1)) * 8u;
#pragma unroll
      for (int c = 1; c < 16; ++c)
        x[c] = gl::mont_mul(x[c], table_load(t, vo, (uint32_t)(pass_offset(LOGN, p) + (brev4(c) - 1) * s)));
    } else if (b0 > 0) {
      const int s = 1 << b0;
      const uint64_t* t = twi + pass_offset(LOGN, p) + (g & (s - 1));
#pragma unroll
      for (int c = 1; c < 16; ++c) x[c] = gl::mont_mul(x[c], ldg(t + (brev4(c) - 1) * s));
    }
    dit_r<16>(x);
    if (p > 0) {
      st16(lds, x, base, b0);
      xsync(b0 + 4 <= 6);
    }
  }
  if (MERGED) {
#pragma unroll
    for (int c = 1; c < 16; ++c) x[c] = gl::mul_pow2(x[c], (192 - (78 * c) % 192) % 192);
  }
}

// Runtime-length variants (4 <= logn <= 14) for kernels that run transforms
// of several lengths: the 16-point butterflies are the same compile-time
// code, only pass count, strides and table offsets are runtime values, so one
// copy of the code serves every length.
__device__ __forceinline__ void short_dif(uint64_t* x, int r) {
  if (r == 8) {
#pragma unroll
    for (int blk = 0; blk < 16; blk += 8) dif_r<8>(x + blk);
  } else if (r == 4) {
#pragma unroll
    for (int blk = 0; blk < 16; blk += 4) dif_r<4>(x + blk);
  } else if (r == 2) {
#pragma unroll
    for (int blk = 0; blk < 16; blk += 2) dif_r<2>(x + blk);
  }
}
__device__ __forceinline__ void short_dit(uint64_t* x, int r) {
  if (r == 8) {
#pragma unroll
    for (int blk = 0; blk < 16; blk += 8) dit_r<8>(x + blk);
  } else if (r == 4) {
#pragma unroll
    for (int blk = 0; blk < 16; blk += 4) dit_r<4>(x + blk);
  } else if (r == 2) {
#pragma unroll
    for (int blk = 0; blk < 16; blk += 2) dit_r<2>(x + blk);
  }
}

__device__ __forceinline__ void ntt16_fwd_rt(uint64_t* x, uint64_t* lds, const uint64_t* __restrict__ tw, int g,
                                             int logn, bool wave = false) {
  const int nfull = logn >> 2;
  const int sh = logn & 3;
  int off = 0;
  for (int p = 0; p < nfull; ++p) {
    const int b0 = logn - 4 - 4 * p;
    const int base = pbase(g, b0);
    if (p > 0) ld16(lds, x, base, b0);
    dif_r<16>(x);
    if (b0 > 0) {
      const int s = 1 << b0;
      const uint64_t* t = tw + off + (g & (s - 1));
#pragma unroll
      for (int c = 1; c < 16; ++c) x[c] = gl::mont_mul(x[c], ldg(t + (brev4(c) - 1) * s));
      off += 15 * s;
    }
    if (p < nfull - 1 || sh) {
      st16(lds, x, base, b0);
      xsync(wave || b0 <= 6);
    }
  }
  if (sh) {
    ld16(lds, x, 16 * g, 0);
    short_dif(x, 1 << sh);
  }
}

__device__ __forceinline__ void ntt16_inv_rt(uint64_t* x, uint64_t* lds, const uint64_t* __restrict__ twi, int g,
                                             int logn, bool wave = false) {
  const int nfull = logn >> 2;
  const int sh = logn & 3;
  if (sh) {
    short_dit(x, 1 << sh);
    st16(lds, x, 16 * g, 0);
    xsync(true);
  }
  // table offset of pass p = sum of 15 * 2^b0 over the passes before it
  int off = 0;
  for (int p = 0; p < nfull; ++p) {
    const int b0 = logn - 4 - 4 * p;
    if (b0 > 0) off += 15 << b0;
  }
  for (int p = nfull - 1; p >= 0; --p) {
    const int b0 = logn - 4 - 4 * p;
    const int base = pbase(g, b0);
    if (p < nfull - 1 || sh) ld16(lds, x, base, b0);
    if (b0 > 0) {
      const int s = 1 << b0;
      off -= 15 * s;
      const uint64_t* t = twi + off + (g & (s - 1));
#pragma unroll
      for (int c = 1; c < 16; ++c) x[c] = gl::mont_mul(x[c], ldg(t + (brev4(c) - 1) * s));
    }
    dit_r<16>(x);
    if (p > 0) {
      st16(lds, x, base, b0);
      xsync(wave || b0 + 4 <= 6);
    }
  }
}

}  // namespace yhk
