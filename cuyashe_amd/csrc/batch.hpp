// Persistent batched ring products over uneven rings in one launch.
//
// A batch is a list of independent products c = a*b mod (x^n + 1, q) whose
// rings differ (n from 16 to 16384, q of any supported width, so the prime
// count R differs too). The host splits every product into R tasks (one per
// CRT prime) and packs tasks of equal n into bundles of W / (n/16) tasks, so
// one workgroup of W threads runs a bundle as W / (n/16) side-by-side
// transforms (CRT -> forward NTT of both operands -> pointwise -> inverse NTT
// -> residue y_i). Bundles sit in per-XCD queues (all tasks of one product in
// the same queue, so its operands stay in one L2); a workgroup drains its own
// queue first and then takes bundles from the others.
//
// The workgroup whose task completes the last prime of a product (a
// per-product counter, agent-scope release/acquire) runs that product's
// ICRT + mod-q epilogue with all its threads, in the same launch. Nothing in the kernel waits on another workgroup: every
// loop either consumes a bundle index or ends, so the grid always drains.
//
// Replaces the reference's per-product call chain callCRT -> callNTT ->
// polynomialNTTMul -> callNTT -> callICRT (src/cuda_bn.cu:970-1045,
// src/cuda_functions.cu:447-474), which it runs one product at a time.
#pragma once
#include "kernels.hpp"

namespace yhk {

struct BatchInst {
  const uint64_t* A;  // [LQ][n], coefficients in [0, q) (any value below 2^(64 LQ) is reduced)
  const uint64_t* B;
  uint64_t* out;      // [LQ][n]
  uint32_t* Y;        // [R + 1][n] residue scratch (layout of store_y)
  int ring;           // index into BatchArgs::rings
  int pad;
};

struct BatchTask {
  int inst;
  int prime;
};

// Every task of a bundle belongs to the same ring and either all or none
// are modulus 0, so the ring's tables and the residue path are
// workgroup-uniform (scalar loads and branches); only instance and prime
// vary.
struct BatchBundle {
  int logn;
  int count;  // tasks in this bundle (<= W / (n/16))
  int task0;  // first task index
  int ring;   // index into BatchArgs::rings
  int gold;   // 1: every task is modulus 0 (P); 0: none is
  int pad;
};

struct BatchArgs {
  const RingDev* rings;
  const BatchInst* inst;
  const BatchTask* tasks;
  const BatchBundle* bundles;
  const int* qstart;  // [nqueues + 1]: queue k holds bundles [qstart[k], qstart[k+1])
  int* head;          // [nqueues] next unclaimed bundle of each queue
  int* done;          // [ninst] primes finished per product
  int nqueues;
};

// Workgroups of up to 512 threads run two or more per CU; a batch with an
// n = 16384 ring runs 1024-thread workgroups (one length-16384 transform
// each, 139 KB of LDS, one workgroup per CU).
constexpr int BATCH_MAX_THREADS = 1024;
constexpr int BATCH_TUNED_THREADS = 512;  // largest workgroup whose waves per SIMD are tunable
constexpr int BATCH_MIN_LOGN = 4;
constexpr int BATCH_MAX_LOGN = 14;

// Dynamic LDS of a W-thread batch workgroup: W/(n/16) transforms of n
// elements plus padding = 17 W words whatever n is.
__host__ __device__ constexpr size_t batch_lds_bytes(int W) { return (size_t)W * 17 * sizeof(uint64_t); }

static __global__ void k_batch_reset(int* head, int nq, int* done, int ninst) {
  for (int i = threadIdx.x; i < nq; i += blockDim.x) head[i] = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < ninst; i += gridDim.x * blockDim.x) done[i] = 0;
}

// s_waitcnt vmcnt(0) (gfx9 encoding: expcnt and lgkmcnt left at their
// maxima): this wave's global loads and stores have completed.
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// CRT residues of a task's 16 coefficients j = g + c T (limb stride N),
// twisted for the negacyclic transform: in groups of G coefficients, limbs
// outer, so a group's loads of one limb are in flight together.
template <bool GOLD, int G>
__device__ __forceinline__ void batch_load(uint64_t* x, const uint64_t* a, int LQ, int N, int T, int g,
                                           const uint64_t* pw, const SmallPrime& sk, const uint64_t* twist) {
#pragma unroll
  for (int h = 0; h < E; h += G) {
    uint64_t acc[G];
#pragma unroll
    for (int c = 0; c < G; ++c) acc[c] = 0;
    if (GOLD) {
      for (int l = LQ - 1; l >= 0; --l) {
        const uint64_t* al = a + (size_t)l * N + g;
#pragma unroll
        for (int c = 0; c < G; ++c) acc[c] = gl::reduce128(acc[c], ldg(al + (h + c) * T));
      }
    } else {
      for (int l = 0; l < LQ; ++l) {
        const uint32_t w0 = (uint32_t)ldg(pw + 2 * l), w1 = (uint32_t)ldg(pw + 2 * l + 1);
        const uint64_t* al = a + (size_t)l * N + g;
#pragma unroll
        for (int c = 0; c < G; ++c) {
          const uint64_t v = ldg(al + (h + c) * T);
          acc[c] += (uint64_t)(uint32_t)v * w0 + (uint64_t)(uint32_t)(v >> 32) * w1;
        }
        if ((l & 3) == 3) {
#pragma unroll
          for (int c = 0; c < G; ++c) acc[c] = mod_small(acc[c], sk);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < G; ++c) {
      const int j = g + (h + c) * T;
      x[h + c] = GOLD ? gl::mul(acc[c], ldg(twist + j)) : twist_centered(mod_small(acc[c], sk), sk.p, ldg(twist + j));
    }
    if (h + G < E) group_fence<G>(x + h);
  }
}

// One bundle of tasks of length-2^logn rings (runtime length: one copy of the
// transform code serves every n). Sets s_last[sub] for the sub-groups whose
// task finished the last prime of its product.
// Transforms of n <= 1024 lie inside one wavefront each, so their LDS
// exchanges synchronise the wavefront only.
template <bool GOLD>
__device__ __forceinline__ void batch_bundle(const BatchArgs& ba, const BatchBundle& bd, const RingDev& rd,
                                             uint64_t* lds, int* s_last) {
  const int logn = bd.logn;
  const int N = 1 << logn;
  const int T = N >> 4;
  const int sub = threadIdx.x >> (logn - 4);
  const int g = threadIdx.x & (T - 1);
  // sub-groups without a task repeat task 0's work (every thread must take
  // part in the transform's barriers) and store nothing
  const bool active = sub < bd.count;
  const BatchTask tk = ba.tasks[bd.task0 + (active ? sub : 0)];
  const bool wave = logn <= 10;
  const BatchInst in = ba.inst[tk.inst];
  const int i = tk.prime;
  const SmallPrime sk = rd.sp[0][i];  // row 0 (modulus P) is unused
  const int LQ = rd.LQ;
  const uint64_t* pw = rd.pow32[0] + (size_t)i * 2 * rd.LW;
  const NttTab tab = rd.tab;
  uint64_t* my = lds + (size_t)sub * (N + T);
  uint64_t xa[E], xb[E];
  constexpr bool gold = GOLD;  // modulus 0 is P itself (exact residues, not centred)
  batch_load<GOLD, 8>(xa, in.A, LQ, N, T, g, pw, sk, tab.twist);
  ntt16_fwd_rt(xa, my, tab.fwd, g, logn, wave);
  // second operand in groups of four (group_fence): fewer loads in flight
  // while the first operand's transform is held in registers
  batch_load<GOLD, 4>(xb, in.B, LQ, N, T, g, pw, sk, tab.twist);
  xsync(wave);
  ntt16_fwd_rt(xb, my, tab.fwd, g, logn, wave);
#pragma unroll
  for (int c = 0; c < E; ++c) xa[c] = gl::mont_mul(xa[c], xb[c]);  // x 2^-64, cancelled by untwist_pw
  xsync(wave);
  ntt16_inv_rt(xa, my, tab.inv, g, logn, wave);
  if (active) {
    const uint64_t inv0 = rd.invMi[0][0];
    // The untwist factors are read before the first row is stored: the
    // stores may alias the table, so a read issued after a store would wait
    // for it (one memory round trip per coefficient).
    uint64_t uw[E];
#pragma unroll
    for (int c = 0; c < E; ++c) uw[c] = ldg(tab.untwist_pw + g + c * T);
#pragma unroll
    for (int c = 0; c < E; ++c) {
      const int j = g + c * T;
      store_y(in.Y, N, i, j, finish_any(gold, sk, inv0, gl::mont_mul(xa[c], uw[c])));
    }
  }
  // Publish the residues: every wave waits until its stores have reached L2
  // (vmcnt = 0), then after the barrier one thread writes the L2 back at
  // agent scope (release) and moves the counters.
  wait_vmem();
  __syncthreads();
  if (threadIdx.x == 0) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (active && g == 0) {
    const int prev = atomicAdd(&ba.done[tk.inst], 1);
    s_last[sub] = prev == rd.set_R[0] - 1;
  }
  __syncthreads();
}

// ICRT + mod q of one product by the whole workgroup.
template <int LQ>
__device__ __forceinline__ void batch_icrt(const RingDev& rd, const BatchInst& in, int n) {
  for (int j = threadIdx.x; j < n; j += blockDim.x)
    icrt_coef<LQ>(rd, in.Y + j, n, 0, EPI_MODQ, 0, in.out + j, LQ, n);
}

// MAXLQ bounds the limb count of q in the batch: with the in-kernel ICRT for
// LQ <= 4 the kernel needs ~150 VGPRs, so WAVES = 3 waves per SIMD fit
// without spills and WAVES = 4 (128 VGPRs) spills a little; LQ up to 8 needs
// the full 256 (WAVES = 2). WAVES is a tuning knob of the launch shape for
// WT <= 512 threads per workgroup; WT = 1024 is always four waves per SIMD.
template <int MAXLQ, int WAVES, int WT>
__global__ void __launch_bounds__(WT, WAVES) k_batch_prod(BatchArgs ba) {
  extern __shared__ uint64_t lds[];
  __shared__ int s_b;
  __shared__ int s_last[WT];
  const int q0 = blockIdx.x % ba.nqueues;  // consecutive workgroups go to consecutive XCDs
  int qi = 0;
  while (true) {
    if (threadIdx.x == 0) {
      int b = -1;
      while (qi < ba.nqueues) {
        const int q = (q0 + qi) % ba.nqueues;
        const int k = atomicAdd(&ba.head[q], 1);
        if (k < ba.qstart[q + 1] - ba.qstart[q]) {
          b = ba.qstart[q] + k;
          break;
        }
        ++qi;  // queue q is drained: move on to the next one
      }
      s_b = b;
    }
    __syncthreads();
    const int bi = __builtin_amdgcn_readfirstlane(s_b);
    if (bi < 0) break;
    const BatchBundle bd = ba.bundles[bi];
    const RingDev& rd = ba.rings[bd.ring];
    if (bd.gold) batch_bundle<true>(ba, bd, rd, lds, s_last);
    else batch_bundle<false>(ba, bd, rd, lds, s_last);
    // products completed by this bundle: ICRT by the whole workgroup (the
    // bundle's ring is workgroup-uniform, so is its limb count)
    const int n = 1 << bd.logn;
    bool acquired = false;
    for (int sub = 0; sub < bd.count; ++sub) {
      if (!s_last[sub]) continue;
      if (!acquired) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");  // see every prime's residues
        acquired = true;
      }
      const BatchInst in = ba.inst[ba.tasks[bd.task0 + sub].inst];
      switch (rd.LQ) {
        case 1: batch_icrt<1>(rd, in, n); break;
        case 2: batch_icrt<2>(rd, in, n); break;
        case 3: batch_icrt<3>(rd, in, n); break;
        case 4: batch_icrt<4>(rd, in, n); break;
        case 5: if (MAXLQ >= 5) batch_icrt<MAXLQ >= 5 ? 5 : 1>(rd, in, n); break;
        case 6: if (MAXLQ >= 6) batch_icrt<MAXLQ >= 6 ? 6 : 1>(rd, in, n); break;
        case 7: if (MAXLQ >= 7) batch_icrt<MAXLQ >= 7 ? 7 : 1>(rd, in, n); break;
        default: if (MAXLQ >= 8) batch_icrt<MAXLQ >= 8 ? 8 : 1>(rd, in, n); break;
      }
    }
    __syncthreads();  // s_b, s_last and the LDS are reused by the next bundle
  }
}

}  // namespace yhk
