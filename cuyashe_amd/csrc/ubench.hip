// Register-resident microbenchmark of the Goldilocks transform arithmetic
// (libyashe_ubench.so, used by bench.py to state the arithmetic floor of the
// product kernel beside its measured roofline). Every wave runs a loop of one
// building block on 16 values per lane held in registers, with no memory
// traffic inside the loop:
//   0: forward radix-16 pass (dif_r<16>: 32 shift-only butterflies, then 15
//      inter-pass twiddle products, gl::mont_mul)
//   1: forward radix-16 DFT alone (dif_r<16>)
//   2: inverse radix-16 DFT alone (dit_r<16>)
//   3: 16 Montgomery products (the pointwise product)
// and reports SIMD cycles per wave per iteration:
//   ms * clock * (CUs * 4 SIMDs) / (waves * iterations).
// This is a separate library: no product kernel carries a switch for it.
#include <hip/hip_runtime.h>

#include "ntt16.hpp"

namespace {

using namespace yhk;

template <int V>
__global__ void __launch_bounds__(256) k_ubench(uint64_t* out, const uint64_t* tw, uint64_t seed, int iters) {
  uint64_t x[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) x[c] = (seed * (threadIdx.x + 7) * (c + 3) + blockIdx.x) % gl::P;
  uint64_t w[16];
#pragma unroll
  for (int c = 0; c < 16; ++c) w[c] = tw[c];
  for (int it = 0; it < iters; ++it) {
    if constexpr (V == 0) {
      dif_r<16>(x);
#pragma unroll
      for (int c = 1; c < 16; ++c) x[c] = gl::mont_mul(x[c], w[c]);
    } else if constexpr (V == 1) {
      dif_r<16>(x);
    } else if constexpr (V == 2) {
      dit_r<16>(x);
    } else {
#pragma unroll
      for (int c = 0; c < 16; ++c) x[c] = gl::mont_mul(x[c], w[c]);
    }
    __builtin_amdgcn_sched_barrier(0);
  }
  uint64_t acc = 0;
#pragma unroll
  for (int c = 0; c < 16; ++c) acc ^= x[c];
  out[blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

template <int V>
hipError_t time_one(uint64_t* out, const uint64_t* tw, int blocks, int iters, float* ms) {
  hipEvent_t a, b;
  hipError_t e = hipEventCreate(&a);
  if (e != hipSuccess) return e;
  if ((e = hipEventCreate(&b)) != hipSuccess) return e;
  hipLaunchKernelGGL(k_ubench<V>, dim3(blocks), dim3(256), 0, 0, out, tw, 12345ull, 8);  // warm-up
  (void)hipEventRecord(a, 0);
  hipLaunchKernelGGL(k_ubench<V>, dim3(blocks), dim3(256), 0, 0, out, tw, 12345ull, iters);
  (void)hipEventRecord(b, 0);
  e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  return e;
}

}  // namespace

extern "C" {

// Fills cycles[4] (SIMD cycles per wave per iteration of blocks 0..3) and
// info[3] = {CUs, clock in kHz, resident waves launched}. Returns 0 or the
// HIP error code.
int yhu_transform_cycles(int device, int iters, double* cycles, double* info) {
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return (int)e;
  int cus = 0, clk = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  (void)hipDeviceGetAttribute(&clk, hipDeviceAttributeClockRate, device);
  const int blocks = cus * 8;  // eight 256-thread workgroups per CU: 8 waves per SIMD
  uint64_t *out = nullptr, *tw = nullptr;
  if ((e = hipMalloc(&out, (size_t)blocks * 256 * sizeof(uint64_t))) != hipSuccess) return (int)e;
  if ((e = hipMalloc(&tw, 16 * sizeof(uint64_t))) != hipSuccess) return (int)e;
  uint64_t h[16];
  for (int c = 0; c < 16; ++c) h[c] = gl::to_mont(gl::pow(7, (gl::P - 1) / 4096 * (uint64_t)(c + 1)));
  if ((e = hipMemcpy(tw, h, sizeof(h), hipMemcpyHostToDevice)) != hipSuccess) return (int)e;
  float ms[4] = {};
  if (!e) e = time_one<0>(out, tw, blocks, iters, &ms[0]);
  if (!e) e = time_one<1>(out, tw, blocks, iters, &ms[1]);
  if (!e) e = time_one<2>(out, tw, blocks, iters, &ms[2]);
  if (!e) e = time_one<3>(out, tw, blocks, iters, &ms[3]);
  (void)hipFree(out);
  (void)hipFree(tw);
  if (e != hipSuccess) return (int)e;
  const double waves = (double)blocks * 256 / 64;
  const double simd_cycles_per_ms = (double)clk * cus * 4;  // kHz = cycles per ms
  for (int v = 0; v < 4; ++v) cycles[v] = ms[v] * simd_cycles_per_ms / (waves * iters);
  info[0] = cus;
  info[1] = clk;
  info[2] = waves;
  return 0;
}

}  // extern "C"
