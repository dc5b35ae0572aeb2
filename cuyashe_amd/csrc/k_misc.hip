// The primitive kernels of the C ABI (standalone cyclic NTT, CRT, residue
// operations, pointwise product) and the helpers of the inverse recursion.
#include "launch.hpp"

namespace yhl {

#define YH_NTT_CASES(X) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14)

hipError_t ntt_cyclic(int logN, int count, hipStream_t s, uint64_t* data, const yhk::NttTab& tab, uint64_t Ninv,
                      int inverse) {
  const dim3 grid(count), block((1 << logN) / 16);
  const size_t lds = lds_bytes(1 << logN);
  switch (logN) {
#define X(L) case L: hipLaunchKernelGGL(yhk::k_ntt_cyclic<L>, grid, block, lds, s, data, tab, Ninv, inverse); break;
    YH_NTT_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t crt(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* A, int lin, int ncoef,
               size_t lstride, uint64_t* out, int spacing) {
  hipLaunchKernelGGL(yhk::k_crt, grid, block, 0, s, rd, A, lin, ncoef, lstride, out, spacing);
  return hipGetLastError();
}

hipError_t residue_to_y(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* res, int spacing,
                        uint32_t* Y, int set) {
  hipLaunchKernelGGL(yhk::k_residue_to_y, grid, block, 0, s, rd, res, spacing, Y, set);
  return hipGetLastError();
}

hipError_t residue_addsub(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, uint64_t* C, const uint64_t* A,
                          const uint64_t* B, int spacing, int sub) {
  hipLaunchKernelGGL(yhk::k_residue_addsub, grid, block, 0, s, rd, C, A, B, spacing, sub);
  return hipGetLastError();
}

hipError_t residue_scalar(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, uint64_t* C, const uint64_t* A,
                          int spacing, int op, const uint64_t* S) {
  hipLaunchKernelGGL(yhk::k_residue_scalar, grid, block, 0, s, rd, C, A, spacing, op, S);
  return hipGetLastError();
}

hipError_t crt_bn(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const yhk::BnT* C, int ncoef, uint64_t* out,
                  int spacing) {
  hipLaunchKernelGGL(yhk::k_crt_bn, grid, block, 0, s, rd, C, ncoef, out, spacing);
  return hipGetLastError();
}

hipError_t to_bn(dim3 grid, dim3 block, hipStream_t s, const uint64_t* V, int lw, int ncoef, yhk::BnT* C) {
  hipLaunchKernelGGL(yhk::k_to_bn, grid, block, 0, s, V, lw, ncoef, C);
  return hipGetLastError();
}

hipError_t pointwise(dim3 grid, dim3 block, hipStream_t s, uint64_t* A, const uint64_t* B, size_t count) {
  hipLaunchKernelGGL(yhk::k_pointwise, grid, block, 0, s, A, B, count);
  return hipGetLastError();
}

hipError_t take_even(dim3 grid, dim3 block, hipStream_t s, const uint64_t* in, uint64_t* out, int m, int LQ) {
  hipLaunchKernelGGL(yhk::k_take_even, grid, block, 0, s, in, out, m, LQ);
  return hipGetLastError();
}

hipError_t spread_even(dim3 grid, dim3 block, hipStream_t s, const uint64_t* in, uint64_t* out, int m, int LQ) {
  hipLaunchKernelGGL(yhk::k_spread_even, grid, block, 0, s, in, out, m, LQ);
  return hipGetLastError();
}

hipError_t join_halves(dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint32_t* Yu,
                        const uint32_t* Yv, uint32_t* Yc, int h, int N) {
  hipLaunchKernelGGL(yhk::k_join_halves, grid, block, 0, s, rd, Yu, Yv, Yc, h, N);
  return hipGetLastError();
}

template <int L>
static void ntt_attr() {
  (void)hipFuncSetAttribute((const void*)yhk::k_ntt_cyclic<L>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_bytes(1 << L));
}

void misc_attrs() {
#define X(L) ntt_attr<L>();
  YH_NTT_CASES(X)
#undef X
}

}  // namespace yhl
