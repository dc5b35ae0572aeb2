// The persistent kernel of uneven batches (k_batch_prod) and its counter
// reset.
#include "launch.hpp"

namespace yhl {

hipError_t batch_reset(dim3 grid, dim3 block, hipStream_t s, int* head, int nq, int* done, int ninst) {
  hipLaunchKernelGGL(yhk::k_batch_reset, grid, block, 0, s, head, nq, done, ninst);
  return hipGetLastError();
}

hipError_t batch_prod(int maxlq, int waves, dim3 grid, dim3 block, size_t lds, hipStream_t s,
                      const yhk::BatchArgs& args) {
  if (block.x > (unsigned)yhk::BATCH_TUNED_THREADS) {
    if (maxlq > 4)
      hipLaunchKernelGGL((yhk::k_batch_prod<8, 4, yhk::BATCH_MAX_THREADS>), grid, block, lds, s, args);
    else
      hipLaunchKernelGGL((yhk::k_batch_prod<4, 4, yhk::BATCH_MAX_THREADS>), grid, block, lds, s, args);
  } else if (maxlq > 4) {
    hipLaunchKernelGGL((yhk::k_batch_prod<8, 2, yhk::BATCH_TUNED_THREADS>), grid, block, lds, s, args);
  } else if (waves >= 4) {
    hipLaunchKernelGGL((yhk::k_batch_prod<4, 4, yhk::BATCH_TUNED_THREADS>), grid, block, lds, s, args);
  } else {
    hipLaunchKernelGGL((yhk::k_batch_prod<4, 3, yhk::BATCH_TUNED_THREADS>), grid, block, lds, s, args);
  }
  return hipGetLastError();
}

void batch_attrs() {
  using namespace yhk;
  for (const void* f : {(const void*)k_batch_prod<4, 3, BATCH_TUNED_THREADS>,
                        (const void*)k_batch_prod<4, 4, BATCH_TUNED_THREADS>,
                        (const void*)k_batch_prod<8, 2, BATCH_TUNED_THREADS>})
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)batch_lds_bytes(BATCH_TUNED_THREADS));
  for (const void* f : {(const void*)k_batch_prod<4, 4, BATCH_MAX_THREADS>,
                        (const void*)k_batch_prod<8, 4, BATCH_MAX_THREADS>})
    (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)batch_lds_bytes(BATCH_MAX_THREADS));
}

}  // namespace yhl
