// Instantiations of k_ks_scale_digits for q of 2 limbs (n = 1024 ... 16384),
// a translation unit of their own so they compile in parallel with the rest.
#include "launch.hpp"

namespace yhl {

#define YH_FUSED_CASES(X) X(10) X(11) X(12) X(13) X(14)

hipError_t ks_scale_digits_lq2(int logn, dim3 grid, hipStream_t s, const RingDev& rd, const KsArgs& ka) {
  const dim3 block((1 << logn) / 16);
  const size_t lds = lds_bytes(1 << logn);
  switch (logn) {
#define X(L) case L: hipLaunchKernelGGL((yhk::k_ks_scale_digits<L, 2>), grid, block, lds, s, rd, ka); break;
    YH_FUSED_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

void ks_fused_attrs_lq2() {
#define X(L)                                                                                                \
  (void)hipFuncSetAttribute((const void*)yhk::k_ks_scale_digits<L, 2>,                                  \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes(1 << L));
  YH_FUSED_CASES(X)
#undef X
}

}  // namespace yhl
