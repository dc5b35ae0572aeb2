// Host-side launchers of the library's kernels, one group per translation
// unit (k_prod.hip, k_keyswitch.hip, k_recon.hip, k_misc.hip, k_batch.hip),
// so the kernel families compile in parallel and yashe_hip.hip holds only the
// C ABI's host logic. Each launcher enqueues one kernel on `s` and returns
// hipGetLastError() (hipErrorInvalidValue for a shape with no instantiation).
#pragma once
#include <hip/hip_runtime.h>

#include "batch.hpp"
#include "kernels.hpp"

namespace yhl {

using yhk::IcrtArgs;
using yhk::KsArgs;
using yhk::ProdArgs;
using yhk::RingDev;

// Every transform kernel is instantiated per log2(N) in [4, 14].
constexpr int MIN_LOGN = 4;
constexpr int MAX_LOGN = 14;

// LDS of one length-N transform (one padding element per 16).
inline size_t lds_bytes(int N) { return (size_t)(N + N / 16) * sizeof(uint64_t); }

// ---------------------------------------------------------------- k_prod.hip
// Fused product of `pa` (length 2^logN, negacyclic or cyclic); logN < 4 runs
// the direct small-transform kernel.
hipError_t prod(int logN, bool nega, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const ProdArgs& pa);
void prod_attrs();

// ---------------------------------------------------------------- k_keyswitch.hip
hipError_t ks_digits(int logn, bool w32, dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka);
hipError_t ks_mac(int logn, dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka);
// the multiply's scale-and-round fused with the digit transforms (q of one or
// two limbs, w = 2^32, n from 1024 to 16384): one workgroup per ciphertext,
// reads ka.P. The instantiations live in k_ks_fused_lq1.hip and
// k_ks_fused_lq2.hip (compiled in parallel: each is a large unrolled kernel).
hipError_t ks_scale_digits(int logn, int LQ, dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka);
hipError_t ks_scale_digits_lq1(int logn, dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka);
hipError_t ks_scale_digits_lq2(int logn, dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka);
void ks_fused_attrs_lq1();
void ks_fused_attrs_lq2();
constexpr int FUSED_MIN_LOGN = 10;
hipError_t ks_small(dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka, const uint64_t* gamma_res);
// resident key: pre-transform of gamma and the digit-centring correction
hipError_t evk_prepare(int logn, hipStream_t s, const RingDev& rd, const uint64_t* gamma, uint64_t* evk,
                       uint64_t* corr);
hipError_t evk_residues(hipStream_t s, const RingDev& rd, const uint64_t* gamma, uint64_t* evk);
void ks_attrs();

// ---------------------------------------------------------------- k_recon.hip
// Kernels templated on the limb count LQ of q (1..8).
hipError_t icrt(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const IcrtArgs& ia);
// mod-q or scale-and-round reconstruction of the other shapes (any LQ), one
// coefficient per thread, a kernel per epilogue
hipError_t icrt_epi(int LQ, int mode, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const IcrtArgs& ia);
// mod-q or scale-and-round reconstruction of the pipelined shapes (LQ <= 2),
// YH_ICRT_PIPE_COEFS per thread
hipError_t icrt_pipe(int LQ, int mode, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const IcrtArgs& ia);
hipError_t addsub_q(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* a,
                    const uint64_t* b, uint64_t* c, int n, size_t count, int sub, yhk::AddsubLists lists);
hipError_t scalar_q(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* a, uint64_t* c,
                    int n, size_t count, int op, const yhk::ScalarArg& sa);
hipError_t reduce(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* c, int lin,
                  uint64_t* out);
hipError_t negate_odd(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* F,
                      uint64_t* G, int m);

// ---------------------------------------------------------------- k_misc.hip
hipError_t ntt_cyclic(int logN, int count, hipStream_t s, uint64_t* data, const yhk::NttTab& tab, uint64_t Ninv,
                      int inverse);
hipError_t crt(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* A, int lin, int ncoef,
               size_t lstride, uint64_t* out, int spacing);
hipError_t residue_to_y(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* res, int spacing,
                        uint32_t* Y, int set);
hipError_t residue_addsub(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, uint64_t* C, const uint64_t* A,
                          const uint64_t* B, int spacing, int sub);
hipError_t residue_scalar(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, uint64_t* C, const uint64_t* A,
                          int spacing, int op, const uint64_t* S);
hipError_t crt_bn(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const yhk::BnT* C, int ncoef, uint64_t* out,
                  int spacing);
hipError_t to_bn(dim3 grid, dim3 block, hipStream_t s, const uint64_t* V, int lw, int ncoef, yhk::BnT* C);
hipError_t pointwise(dim3 grid, dim3 block, hipStream_t s, uint64_t* A, const uint64_t* B, size_t count);
// linear product rows from its cyclic and negacyclic halves of length h
hipError_t join_halves(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint32_t* Yu,
                       const uint32_t* Yv, uint32_t* Yc, int h, int N);
hipError_t take_even(dim3 grid, dim3 block, hipStream_t s, const uint64_t* in, uint64_t* out, int m, int LQ);
hipError_t spread_even(dim3 grid, dim3 block, hipStream_t s, const uint64_t* in, uint64_t* out, int m, int LQ);
void misc_attrs();

// ---------------------------------------------------------------- k_batch.hip
hipError_t batch_reset(dim3 grid, dim3 block, hipStream_t s, int* head, int nq, int* done, int ninst);
// maxlq > 4: the two-wave wide-modulus variant; else waves 3 or 4
hipError_t batch_prod(int maxlq, int waves, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                      const yhk::BatchArgs& args);
void batch_attrs();

}  // namespace yhl
