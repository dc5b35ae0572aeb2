// Kernels templated on the limb count of q: the reconstruction (ICRT and
// its epilogues), coefficient-domain add/sub and scalar operations mod q,
// the reduction mod (x^n + 1, q) and the odd-coefficient negation of the
// inverse recursion.
#include "launch.hpp"

namespace yhl {

#define YH_LQ_SWITCH(LQ, CALL)                 \
  switch (LQ) {                                \
    case 1: CALL(1); break;                    \
    case 2: CALL(2); break;                    \
    case 3: CALL(3); break;                    \
    case 4: CALL(4); break;                    \
    case 5: CALL(5); break;                    \
    case 6: CALL(6); break;                    \
    case 7: CALL(7); break;                    \
    case 8: CALL(8); break;                    \
    default: return hipErrorInvalidValue;      \
  }                                            \
  return hipGetLastError();

hipError_t icrt(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const IcrtArgs& ia) {
#define C(L) hipLaunchKernelGGL(yhk::k_icrt<L>, grid, block, 0, s, rd, ia)
  YH_LQ_SWITCH(LQ, C)
#undef C
}

hipError_t icrt_pipe(int LQ, int mode, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const IcrtArgs& ia) {
  const bool scale = mode == yhk::EPI_SCALE;
  switch (LQ) {
    case 1:
      if (scale) hipLaunchKernelGGL((yhk::k_icrt_pipe<1, yhk::EPI_SCALE>), grid, block, 0, s, rd, ia);
      else hipLaunchKernelGGL((yhk::k_icrt_pipe<1, yhk::EPI_MODQ>), grid, block, 0, s, rd, ia);
      break;
    case 2:
      if (scale) hipLaunchKernelGGL((yhk::k_icrt_pipe<2, yhk::EPI_SCALE>), grid, block, 0, s, rd, ia);
      else hipLaunchKernelGGL((yhk::k_icrt_pipe<2, yhk::EPI_MODQ>), grid, block, 0, s, rd, ia);
      break;
    default: return hipErrorInvalidValue;  // the pipelined shapes have q of one or two limbs
  }
  return hipGetLastError();
}

hipError_t icrt_epi(int LQ, int mode, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const IcrtArgs& ia) {
#define C(L)                                                                                            \
  if (mode == yhk::EPI_MODQ) hipLaunchKernelGGL((yhk::k_icrt_epi<L, yhk::EPI_MODQ>), grid, block, 0, s, rd, ia);     \
  else if (mode == yhk::EPI_DECRYPT)                                                                       \
    hipLaunchKernelGGL((yhk::k_icrt_epi<L, yhk::EPI_DECRYPT>), grid, block, 0, s, rd, ia);               \
  else hipLaunchKernelGGL((yhk::k_icrt_epi<L, yhk::EPI_SCALE>), grid, block, 0, s, rd, ia)
  YH_LQ_SWITCH(LQ, C)
#undef C
}

hipError_t addsub_q(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* a,
                    const uint64_t* b, uint64_t* c, int n, size_t count, int sub, yhk::AddsubLists lists) {
#define C(L) hipLaunchKernelGGL(yhk::k_addsub_q<L>, grid, block, 0, s, rd, a, b, c, n, count, sub, lists)
  YH_LQ_SWITCH(LQ, C)
#undef C
}

hipError_t scalar_q(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* a, uint64_t* c,
                    int n, size_t count, int op, const yhk::ScalarArg& sa) {
#define C(L) hipLaunchKernelGGL(yhk::k_scalar_q<L>, grid, block, 0, s, rd, a, c, n, count, op, sa)
  YH_LQ_SWITCH(LQ, C)
#undef C
}

hipError_t reduce(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* c, int lin,
                  uint64_t* out) {
#define C(L) hipLaunchKernelGGL(yhk::k_reduce<L>, grid, block, 0, s, rd, c, lin, out)
  YH_LQ_SWITCH(LQ, C)
#undef C
}

hipError_t negate_odd(int LQ, dim3 grid, dim3 block, hipStream_t s, const RingDev& rd, const uint64_t* F,
                      uint64_t* G, int m) {
#define C(L) hipLaunchKernelGGL(yhk::k_negate_odd<L>, grid, block, 0, s, rd, F, G, m)
  YH_LQ_SWITCH(LQ, C)
#undef C
}

}  // namespace yhl
