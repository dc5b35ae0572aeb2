// Fused product kernels (k_prod: CRT, forward transforms, pointwise,
// inverse transform, residue -> y) for every transform length.
#include "launch.hpp"

namespace yhl {

#define YH_PROD_CASES(X) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14)

hipError_t prod(int logN, bool nega, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const ProdArgs& pa) {
  if (logN < MIN_LOGN) {
    hipLaunchKernelGGL(yhk::k_prod_small, grid, block, 0, s, rd, pa, nega ? 1 : 0);
    return hipGetLastError();
  }
  const size_t lds = lds_bytes(1 << logN);
  switch (logN) {
#define X(L)                                                                      \
  case L:                                                                         \
    if (nega) hipLaunchKernelGGL((yhk::k_prod<L, true>), grid, block, lds, s, rd, pa); \
    else hipLaunchKernelGGL((yhk::k_prod<L, false>), grid, block, lds, s, rd, pa);     \
    break;
    YH_PROD_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int L>
static void prod_attr() {
  const int bytes = (int)lds_bytes(1 << L);
  (void)hipFuncSetAttribute((const void*)yhk::k_prod<L, true>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  (void)hipFuncSetAttribute((const void*)yhk::k_prod<L, false>, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

void prod_attrs() {
#define X(L) prod_attr<L>();
  YH_PROD_CASES(X)
#undef X
}

}  // namespace yhl
