// Keyswitch kernels: digit transforms, the multiply-accumulate with the
// resident key and its inverse transform, the key's pre-transform and the
// digit-centring correction, and the direct kernels of rings below n = 16.
#include "launch.hpp"

namespace yhl {

#define YH_KS_CASES(X) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14)

hipError_t ks_digits(int logn, bool w32, dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka) {
  const dim3 block((1 << logn) / 16);
  const size_t lds = lds_bytes(1 << logn);
  switch (logn) {
#define X(L)                                                                              \
  case L:                                                                                 \
    if (w32) hipLaunchKernelGGL((yhk::k_ks_digits<L, true>), grid, block, lds, s, rd, ka); \
    else hipLaunchKernelGGL((yhk::k_ks_digits<L, false>), grid, block, lds, s, rd, ka);    \
    break;
    YH_KS_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t ks_mac(int logn, dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka) {
  const dim3 block((1 << logn) / 16);
  const size_t lds = lds_bytes(1 << logn);
  switch (logn) {
#define X(L) case L: hipLaunchKernelGGL((yhk::k_ks_mac<L>), grid, block, lds, s, rd, ka); break;
    YH_KS_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t ks_small(dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka, const uint64_t* gamma_res) {
  hipLaunchKernelGGL(yhk::k_keyswitch_small, grid, dim3(64), 0, s, rd, ka, gamma_res);
  return hipGetLastError();
}

hipError_t evk_prepare(int logn, hipStream_t s, const RingDev& rd, const uint64_t* gamma, uint64_t* evk,
                       uint64_t* corr) {
  const dim3 block((1 << logn) / 16);
  const size_t lds = lds_bytes(1 << logn);
  const dim3 grid(rd.set_R[1], rd.lw);
  switch (logn) {
#define X(L)                                                                                   \
  case L:                                                                                      \
    hipLaunchKernelGGL(yhk::k_evk_prep<L>, grid, block, lds, s, rd, gamma, evk);               \
    hipLaunchKernelGGL(yhk::k_ks_corr<L>, dim3(rd.set_R[1]), block, lds, s, rd, evk, corr);    \
    break;
    YH_KS_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t evk_residues(hipStream_t s, const RingDev& rd, const uint64_t* gamma, uint64_t* evk) {
  hipLaunchKernelGGL(yhk::k_evk_residues, dim3(rd.set_R[1], rd.lw), dim3(64), 0, s, rd, gamma, evk);
  return hipGetLastError();
}

template <int L>
static void ks_attr() {
  const int bytes = (int)lds_bytes(1 << L);
  (void)hipFuncSetAttribute((const void*)yhk::k_ks_digits<L, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)yhk::k_ks_digits<L, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)yhk::k_ks_mac<L>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)yhk::k_evk_prep<L>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)yhk::k_ks_corr<L>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

void ks_attrs() {
#define X(L) ks_attr<L>();
  YH_KS_CASES(X)
#undef X
  ks_fused_attrs_lq1();
  ks_fused_attrs_lq2();
}

hipError_t ks_scale_digits(int logn, int LQ, dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka) {
  if (LQ == 1) return ks_scale_digits_lq1(logn, grid, s, rd, ka);
  if (LQ == 2) return ks_scale_digits_lq2(logn, grid, s, rd, ka);
  return hipErrorInvalidValue;
}

}  // namespace yhl
