/*
 * yashe_hip.h - C ABI of the MI355X (gfx950) YASHE polynomial-ring hot path.
 *
 * This is the lower boundary the host data model (Polynomial / Ciphertext /
 * Yashe) calls. It replaces the reference's CUDAFunctions statics
 * (src/cuda_functions.h:19-84) and the free CRT/ICRT functions
 * (src/cuda_bn.h:76-87). Differences in convention, applied to every call:
 *  - every entry point returns an int status (YH_OK == 0); yh_last_error()
 *    gives a message for the calling thread. The reference asserts.
 *  - device buffers are owned by the caller (yh_malloc / yh_free); no call
 *    returns a freshly allocated buffer (the reference's callPolynomialMul
 *    and callPolynomialOPInteger do, and leak them).
 *  - all work is ordered on the context's stream (yh_ctx_stream), and calls
 *    return without synchronising unless documented. The reference uses the
 *    legacy default stream with cudaDeviceSynchronize after each operator.
 *  - ring parameters live in a yh_ring handle instead of global statics
 *    (Polynomial::CRTPrimes, CUDAFunctions::N, Yashe::q ...).
 *  - coefficient polynomials are limb-major: limb l of coefficient j of
 *    polynomial b sits at ptr[(b*L + l)*n + j] (L 64-bit limbs, least
 *    significant first). The reference's bn_t {alloc, used, sign, dp*}
 *    per coefficient is not used.
 *  - CRT residues are u64 [R][spacing], as the reference's d_polyCRT.
 */
#ifndef YASHE_HIP_H
#define YASHE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define YH_OK 0
#define YH_ERR_ARG 1
#define YH_ERR_HIP 2
#define YH_ERR_UNSUPPORTED 3
#define YH_ERR_NOMEM 4
#define YH_ERR_INTERNAL 5

#define YH_FORWARD 1 /* ntt_mode_t FORWARD, src/settings.h:38 */
#define YH_INVERSE 0 /* ntt_mode_t INVERSE */
#define YH_ADD 0     /* add_mode_t ADD, src/settings.h:36 */
#define YH_SUB 1     /* add_mode_t SUB */
#define YH_MUL 2     /* add_mode_t MUL (scalar operations only) */

typedef struct yh_ctx yh_ctx;
typedef struct yh_ring yh_ring;
typedef struct yh_evk yh_evk;

typedef struct yh_ring_info {
  int n;          /* ring degree: x^n + 1 */
  int q_limbs;    /* 64-bit limbs of q */
  int q_bits;
  int R;          /* CRT primes of ring products */
  int R_ks;       /* leading primes used by the keyswitch */
  int w_bits;     /* keyswitch base w = 2^w_bits (0: none) */
  int lw;         /* keyswitch digits */
  int work_limbs; /* limbs of the reconstruction (ICRT) output */
  int m_bits;     /* bits of the CRT modulus M (all R primes) */
  int m_ks_bits;  /* bits of the keyswitch CRT modulus */
  uint64_t t;     /* plaintext modulus */
  uint64_t prime_max;
} yh_ring_info;

/* ------------------------------------------------------------ errors */
const char* yh_last_error(void);
int yh_version(void);

/* ------------------------------------------------------------ context
 * Replaces the implicit CUDA context + cudaSetDevice(0)
 * (src/benchmark_polynomial.cpp:25) and the per-object streams of
 * src/polynomial.h:48. */
int yh_ctx_create(int device, yh_ctx** out);
int yh_ctx_destroy(yh_ctx* ctx);
/* hipStream_t of the context, as void*. */
void* yh_ctx_stream(yh_ctx* ctx);
/* Blocks until the context stream is idle (cudaDeviceSynchronize in the
 * reference, src/common.h:223). */
int yh_ctx_sync(yh_ctx* ctx);

/* ------------------------------------------------------------ memory
 * Replaces the cudaMalloc / cudaMemcpy calls scattered through
 * src/polynomial.h:1127-1314 (update_crt_spacing) and
 * src/polynomial.cpp:265-417 (update_device_data / update_host_data). */
int yh_malloc(yh_ctx* ctx, size_t bytes, void** dptr);
int yh_free(yh_ctx* ctx, void* dptr);
/* Stream-ordered copies; they return after the copy completed. */
int yh_memcpy_htod(yh_ctx* ctx, void* dst, const void* src, size_t bytes);
int yh_memcpy_dtoh(yh_ctx* ctx, void* dst, const void* src, size_t bytes);
int yh_memcpy_dtod(yh_ctx* ctx, void* dst, const void* src, size_t bytes);
int yh_memset(yh_ctx* ctx, void* dst, int value, size_t bytes);

/* ------------------------------------------------------------ ring set-up
 * Replaces Polynomial::gen_crt_primes (src/polynomial.h:898-955),
 * CUDAFunctions::write_crt_primes (src/cuda_bn.cu:1047-1147) and
 * CUDAFunctions::init (src/cuda_functions.cu:878-956).
 * n: power of two in [2, 16384]; q: q_limbs little-endian limbs (q > 1,
 * at most 8 limbs); t: plaintext modulus (0 if unused); w_bits: keyswitch
 * digit width in bits, 0..32 (0 = no keyswitch). Primes are chosen so that
 * every product of the ring is exact (see the ring_params.hpp header). */
int yh_ring_create(yh_ctx* ctx, int n, const uint64_t* q, int q_limbs, uint64_t t, int w_bits,
                   yh_ring** out);
int yh_ring_destroy(yh_ring* ring);
/* Host-only planning of the same parameters (no device needed): fills info
 * and the moduli as yh_ring_get_primes does. Used for admission control and
 * CPU-side checks. */
int yh_ring_plan(int n, const uint64_t* q, int q_limbs, uint64_t t, int w_bits, yh_ring_info* info,
                 uint64_t* primes, int cap);
int yh_ring_get_info(const yh_ring* ring, yh_ring_info* info);
/* The R CRT moduli of ring products (Polynomial::CRTPrimes): modulus 0 is
 * P = 2^64 - 2^32 + 1 itself, then small primes descending. When cap is at
 * least R + R_ks the R_ks keyswitch moduli follow (P, then smaller primes
 * sized for the l_w-digit multiply-accumulate). */
int yh_ring_get_primes(const yh_ring* ring, uint64_t* out, int cap);
/* M as limbs: set 0 = the product moduli (Polynomial::CRTProduct), 1 = the
 * keyswitch moduli. */
int yh_ring_get_modulus(const yh_ring* ring, int set, uint64_t* out, int cap);

/* ------------------------------------------------------------ primitives */
/* CRT: res[i][j] = coef_j mod p_i for j < ncoef, 0 for ncoef <= j < spacing.
 * coefs: limb-major [limbs][ncoef] (limbs <= 2*q_limbs+1).
 * Replaces callCRT / cuCRT (src/cuda_bn.cu:970-1002, :803-835). */
int yh_crt(yh_ring* ring, const uint64_t* d_coefs, int limbs, int ncoef, int spacing, uint64_t* d_res);
/* ICRT: coefs_j = sum_i M_i ((r_ij (M_i^-1 mod p_i)) mod p_i) mod M, in
 * [0, M), written as out_limbs limbs [out_limbs][ncoef]. r_ij < p_i.
 * Replaces callICRT / cuPreICRT / cuICRT (src/cuda_bn.cu:1004-1045). */
int yh_icrt(yh_ring* ring, const uint64_t* d_res, int spacing, int ncoef, uint64_t* d_coefs, int out_limbs);
/* Cyclic NTT over P = 2^64 - 2^32 + 1 of `count` contiguous vectors of
 * length N (power of two, 16 <= N <= 16384), in place, natural order:
 * X_k = sum_j x_j w^(jk), w = 7^((P-1)/N); YH_INVERSE uses w^-1 and scales
 * by N^-1 mod P. Replaces CUDAFunctions::callNTT
 * (src/cuda_functions.cu:447-458, inverse scaling NTTScale :397-402). */
int yh_ntt(yh_ctx* ctx, uint64_t* d_data, int N, int count, int direction);
/* a[k] = a[k] * b[k] mod P. Replaces polynomialNTTMul
 * (src/cuda_functions.cu:460-474). */
int yh_pointwise_mulmod(yh_ctx* ctx, uint64_t* d_a, const uint64_t* d_b, size_t count);
/* c = a +- b mod p_i on residues [R][spacing] (op YH_ADD / YH_SUB).
 * Replaces callPolynomialAddSub (src/cuda_functions.cu:115-126), adding
 * the mod p_i that the reference omits. c may alias a. */
int yh_residue_addsub(yh_ring* ring, uint64_t* d_c, const uint64_t* d_a, const uint64_t* d_b, int spacing,
                      int op);
/* Scalar operation on residues [R][spacing] (values in [0, p_i)): op YH_ADD /
 * YH_SUB adds / subtracts the scalar's residue to coefficient 0 of every
 * residue row, YH_MUL multiplies every residue by it, all mod p_i. The
 * scalar is given by its residues: d_scalar_res is a DEVICE array of R
 * words, word i taken mod p_i (the reference's integer_array, Integer's
 * CRT residues). d_c may alias d_a; other coefficients are copied.
 * Replaces CUDAFunctions::callPolynomialOPInteger
 * (src/cuda_functions.cu:477-548), which returns a freshly allocated buffer. */
int yh_residue_scalar(yh_ring* ring, uint64_t* d_c, const uint64_t* d_a, int spacing, int op,
                      const uint64_t* d_scalar_res);
/* The reference's multi-word integer, bn_t (src/cuda_bn.h:18-23), byte for
 * byte: `used` 64-bit words at dp (device memory), least significant first,
 * `alloc` words allocated there; sign is ignored (values are non-negative). */
typedef struct yh_bn_t {
  int alloc;
  int used;
  int sign;
  uint64_t* dp;
} yh_bn_t;
/* yh_crt over the reference's own coefficient layout: d_coefs is a DEVICE
 * array of ncoef yh_bn_t (any word count). Replaces callCRT
 * (src/cuda_bn.cu:970-1002) without changing the caller's data. */
int yh_crt_bn(yh_ring* ring, const yh_bn_t* d_coefs, int ncoef, int spacing, uint64_t* d_res);
/* yh_icrt into the reference's layout: coefficient j (in [0, M)) is written
 * to d_coefs[j].dp (the words that fit its alloc; alloc >= the ring's
 * work_limbs holds every value), used = its significant words, sign = 0.
 * d_coefs is a DEVICE array of ncoef yh_bn_t. Replaces callICRT
 * (src/cuda_bn.cu:1004-1045). */
int yh_icrt_bn(yh_ring* ring, const uint64_t* d_res, int spacing, int ncoef, yh_bn_t* d_coefs);

/* ------------------------------------------------------------ fused operations
 * All take `batch` independent operands laid out back to back. */

/* Plain product in Z[x]/(x^N - 1), N = max(min_N, 2^ceil(log2(deg_a+deg_b)))
 * (the NTT-length rule of common_multiplication, src/common.h:136-226;
 * power-of-two degree sums wrap cyclically exactly like the reference).
 * a: [q_limbs][deg_a+1], b: [q_limbs][deg_b+1], coefficients in [0, q),
 * deg_a, deg_b < n. out: [out_limbs][N] with the exact coefficients in
 * [0, M). *N_out receives N (at most 2^20; above 16384 the product cannot
 * wrap and is the linear one, zero above degree deg_a + deg_b). Replaces Polynomial::operator*
 * (src/polynomial.cpp:52-60) -> callPolynomialMul
 * (src/cuda_functions.cu:617-738) -> icrt. */
int yh_poly_mul_plain(yh_ring* ring, const uint64_t* d_a, int deg_a, const uint64_t* d_b, int deg_b,
                      int min_N, uint64_t* d_out, int out_limbs, int* N_out);
/* The N that yh_poly_mul_plain uses for these degrees (size d_out from it). */
int yh_poly_mul_plain_length(int deg_a, int deg_b, int min_N);
/* c = a * b mod (x^n + 1, q), operands [batch][q_limbs][n] in [0, q).
 * Replaces the sequence c = a*b; c.reduce(); c %= q
 * (src/polynomial.cpp:52-60, src/cuda_functions.cu:1040-1078,
 * src/polynomial.h:593-610). */
int yh_ring_mul(yh_ring* ring, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out, int batch);
/* c = a +- b mod q (op YH_ADD / YH_SUB) on [batch][q_limbs][n] in [0, q).
 * Replaces common_addition (src/common.h:24-80) followed by % q. */
int yh_ring_addsub(yh_ring* ring, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out, int batch, int op);
/* Scalar operation on ring elements [batch][q_limbs][n] in [0, q): op YH_ADD
 * / YH_SUB gives (a_0 +- s) mod q on coefficient 0 (the other coefficients
 * copied), YH_MUL gives a_j * s mod q on every coefficient, s = scalar mod q
 * (scalar: host limbs, any value). d_out may alias d_a. Replaces the scalar
 * operators of Polynomial (src/polynomial.h:459-573, the +, +=, -, -=, * and *= operators taking a ZZ)
 * and Integer::operator* (src/integer.cpp:167-222) on ring elements, which
 * the reference runs as polynomialOPInteger on residues followed by % q. */
int yh_ring_scalar(yh_ring* ring, const uint64_t* d_a, const uint64_t* scalar, int scalar_limbs, uint64_t* d_out,
                   int batch, int op);
/* Reduction of [batch][in_limbs][2n] coefficients (any value < 2^(64 in_limbs))
 * mod (x^n + 1) and q: out_j = (c_j - c_{j+n}) mod q in [0, q).
 * Replaces Polynomial::reduce (src/cuda_functions.cu:1040-1078). */
int yh_ring_reduce(yh_ring* ring, const uint64_t* d_c, int in_limbs, uint64_t* d_out, int batch);
/* YASHE multiply: c = [round(t * c1*c2 / q)]_q, the product taken mod
 * x^n + 1 over Z (no reduction mod q before rounding); ties round down.
 * Replaces Ciphertext::operator* (src/ciphertext.cpp:38-73) with the
 * t factor and rounding order that the reference intends. */
int yh_yashe_mul(yh_ring* ring, const uint64_t* d_c1, const uint64_t* d_c2, uint64_t* d_out, int batch);
/* Keyswitch key: gamma [lw][q_limbs][n] (device, coefficients in [0, q))
 * pre-transformed and kept resident. Replaces the per-call
 * P[k] * gamma[k] products of Ciphertext::keyswitch (src/ciphertext.cpp:81-107). */
int yh_evk_create(yh_ring* ring, const uint64_t* d_gamma, yh_evk** out);
int yh_evk_destroy(yh_evk* evk);
/* c' = [sum_k D_k(c) gamma_k]_(x^n+1, q), D_k(c) = floor(c / w^k) mod w.
 * Replaces Ciphertext::convert/keyswitch/worddecomp
 * (src/ciphertext.cpp:75-130) with the base, mask and digit count fixed
 * as the reference intends (w a base 2^w_bits, l_w = ceil(bits(q)/w_bits)). */
int yh_keyswitch(yh_ring* ring, const yh_evk* evk, const uint64_t* d_c, uint64_t* d_out, int batch);
/* yh_yashe_mul followed by yh_keyswitch (the lazy convert() of the
 * reference, src/ciphertext.cpp:44-47, made eager). */
int yh_yashe_mul_keyswitch(yh_ring* ring, const yh_evk* evk, const uint64_t* d_c1, const uint64_t* d_c2,
                           uint64_t* d_out, int batch);
/* yh_yashe_mul_keyswitch without any device allocation: its scratch
 * (reconstruction rows, the intermediate product and the digit transforms,
 * batch * (4 (max(R, R_ks) + 1) n + 8 q_limbs n + 8 l_w n) bytes, each part
 * rounded up to 256) lives in one static device array of
 * YH_STATIC_SCRATCH_BYTES per GPU, held until the process ends. A larger
 * batch is refused with YH_ERR_ARG. Every call on one GPU shares that array:
 * calls that overlap on one GPU (different contexts, or rings with their own
 * streams) race and must be serialised by the caller. Results equal
 * yh_yashe_mul_keyswitch. Requires n >= 16. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). */
#define YH_STATIC_SCRATCH_BYTES (64u << 20)
int yh_yashe_mul_keyswitch_static(yh_ring* ring, const yh_evk* evk, const uint64_t* d_c1, const uint64_t* d_c2,
                                  uint64_t* d_out, int batch);
/* Decryption core: m_j = round(t * [f*c]_q,j / q) mod t for every coefficient
 * ([.]_q in [0, q)). f: [q_limbs][n] (one key), c: [batch][q_limbs][n];
 * m: [batch][n] u64. Replaces Yashe::decrypt (src/yashe.cpp:143-201), which
 * decodes coefficient 0 only. */
int yh_yashe_decrypt(yh_ring* ring, const uint64_t* d_f, const uint64_t* d_c, uint64_t* d_m, int batch);

/* c_b = a * b_b mod (x^n + 1, q) for one a and a batch of b: the h*s of
 * Yashe::encrypt (src/yashe.cpp:104-142) for many messages under one key. */
int yh_ring_mul_bcast(yh_ring* ring, const uint64_t* d_a, const uint64_t* d_b, uint64_t* d_out, int batch);
/* f^-1 mod (x^n + 1, q) on the device, by the norm recursion
 * f^-1(x) = f(-x) * N(f)^-1(x^2), N(f)(x^2) = f(x) f(-x): log2(n) ring
 * products and one scalar inverse. Returns YH_ERR_ARG if f is not
 * invertible. Replaces Polynomial::InvMod (src/polynomial.h:1338-1367, NTL
 * on the host) for key generation (src/yashe.cpp:37-55). */
int yh_ring_inverse(yh_ring* ring, const uint64_t* d_f, uint64_t* d_out);

/* ------------------------------------------------------------ pointer lists
 * The same operations over ciphertexts scattered in device memory: d_a,
 * d_b, d_c1, d_c2, d_c, d_out and d_m are DEVICE arrays of `batch` device
 * pointers (each to one [L][n] ring element, or one [n] message for d_m),
 * so a serving loop batches resident ciphertexts without gathering them.
 * Results equal the contiguous variants element by element. The reference
 * runs each operation on one Polynomial/Ciphertext object at a time
 * (src/ciphertext.cpp:38-107, src/yashe.cpp:167-197). */
int yh_ring_addsub_list(yh_ring* ring, const uint64_t* const* d_a, const uint64_t* const* d_b, uint64_t* const* d_out,
                        int batch, int op);
int yh_yashe_mul_keyswitch_list(yh_ring* ring, const yh_evk* evk, const uint64_t* const* d_c1,
                                const uint64_t* const* d_c2, uint64_t* const* d_out, int batch);
int yh_yashe_decrypt_list(yh_ring* ring, const uint64_t* d_f, const uint64_t* const* d_c, uint64_t* const* d_m,
                          int batch);

/* ------------------------------------------------------------ streams and events
 * A ring's operations run on the context stream unless the ring is given a
 * stream of its own (default priority), after which work on different rings
 * overlaps on the device. Such a ring is outside yh_graph_begin/end captures
 * and yh_ctx_sync; wait for it with yh_ring_sync or an event. Events mark a
 * point in a ring's (or, with ring NULL, the context's) stream. No
 * counterpart in the reference (every call there ends in
 * cudaDeviceSynchronize). */
int yh_ring_use_own_stream(yh_ring* ring);
int yh_ring_sync(yh_ring* ring);
typedef struct yh_event yh_event;
int yh_event_create(yh_ctx* ctx, yh_event** out);
int yh_event_record(yh_event* ev, yh_ctx* ctx, yh_ring* ring);
/* 1 when the work before the event has finished, 0 while it runs, < 0 on error */
int yh_event_query(yh_event* ev);
int yh_event_sync(yh_event* ev);
int yh_event_elapsed_ms(yh_event* start, yh_event* stop, float* ms);
int yh_event_destroy(yh_event* ev);

/* ------------------------------------------------------------ graphs
 * Captures the launches issued on the context stream into a HIP graph
 * (hipStreamBeginCapture), for replay of a fixed homomorphic circuit. Run
 * the same sequence once before capturing so scratch buffers are sized:
 * a call that would allocate during capture fails with YH_ERR_ARG. No
 * counterpart in the reference (it launches everything eagerly). */
typedef struct yh_graph yh_graph;
int yh_graph_begin(yh_ctx* ctx);
int yh_graph_end(yh_ctx* ctx, yh_graph** out);
int yh_graph_launch(yh_graph* graph);
int yh_graph_destroy(yh_graph* graph);

/* ------------------------------------------------------------ uneven batches
 * Many independent ring products c_k = a_k * b_k mod (x^n_k + 1, q_k) over
 * rings of different n (16..16384) and q, in ONE persistent launch: every
 * product becomes R_k per-prime tasks, packed into workgroup-sized bundles
 * and drained from per-XCD work queues; the task that finishes a product's
 * last prime runs its ICRT in the same launch. Operands and outputs are
 * [L][n] device buffers of each item's ring (values in [0, q)). Results are
 * identical to yh_ring_mul on each item alone. The reference multiplies one
 * product at a time (Polynomial::operator*, src/common.h:136-226, followed by
 * reduce() and %= q); there is no batched counterpart. */
typedef struct yh_batch_item {
  yh_ring* ring;
  const uint64_t* a;
  const uint64_t* b;
  uint64_t* out;
} yh_batch_item;

/* Launch shape of a batch: threads per workgroup (64..1024, a power of two,
 * at least the largest n/16), workgroups per CU in the persistent grid, the
 * number of work queues (1, or 8 = one per XCD) and the waves per SIMD the
 * kernel is compiled for (3 or 4 when every q has at most 4 limbs, else 2;
 * always 4 for 1024-thread workgroups, which a batch with n = 16384 needs). */
typedef struct yh_batch_config {
  int wg_threads;
  int wgs_per_cu;
  int queues;
  int waves;
} yh_batch_config;

typedef struct yh_batch yh_batch;
/* The default shape for these items (no measurement). */
int yh_batch_default_config(const yh_batch_item* items, int count, yh_batch_config* cfg);
/* Plans a batch: uploads the descriptors and sizes scratch once; cfg NULL =
 * the default shape. The plan keeps the item pointers: run it again after
 * writing new operands into the same buffers. */
int yh_batch_create(yh_ctx* ctx, const yh_batch_item* items, int count, const yh_batch_config* cfg, yh_batch** out);
/* Enqueues the batch on the context stream (capturable into a graph). */
int yh_batch_run(yh_batch* batch);
int yh_batch_destroy(yh_batch* batch);
/* Picks the fastest launch shape for this batch composition by timing every
 * candidate `reps` times (HIP events); with cache_path non-NULL the result is
 * read from / appended to that JSON file under a key naming the device and
 * the composition (histogram of n and limb counts), so a second call does
 * not measure. Operands are only read and outputs written as by
 * yh_batch_run. */
int yh_batch_autotune(yh_ctx* ctx, const yh_batch_item* items, int count, const char* cache_path, int reps,
                      yh_batch_config* out);

/* ------------------------------------------------------------ instrumentation
 * Per-kernel timing with HIP events recorded on the context stream around
 * every launch of the fused kernels (no counterpart in the reference, which
 * times whole operators with clock_gettime, src/benchmark_yashe.cpp:152-291).
 * kernel ids: 0 product (CRT + NTTs + pointwise + INTT), 1 keyswitch
 * multiply-accumulate + inverse transform, 2 ICRT epilogue, 3 persistent
 * batch, 4 keyswitch digit transforms (in a fused mul+keyswitch together
 * with the multiply's scale-and-round, see yh_ctx_set_fuse_min_batch). */
#define YH_KERNEL_PROD 0
#define YH_KERNEL_KEYSWITCH 1
#define YH_KERNEL_ICRT 2
#define YH_KERNEL_BATCH 3
#define YH_KERNEL_KS_DIGITS 4
#define YH_KERNEL_COUNT 5
int yh_ctx_set_timing(yh_ctx* ctx, int enable);
/* Restricts timing to the kernels whose bit (1 << kernel_id) is set in
 * `mask` (default: all), so a measured loop pays for event pairs around the
 * launches it reports only. */
int yh_ctx_set_timing_mask(yh_ctx* ctx, unsigned mask);
/* Synchronises, then returns the summed milliseconds and launch count of
 * kernel `kernel_id` since timing was enabled (or last reset). */
int yh_ctx_kernel_time(yh_ctx* ctx, int kernel_id, double* total_ms, long* launches);
int yh_ctx_reset_timing(yh_ctx* ctx);

/* Batch size from which yh_yashe_mul_keyswitch (and its _list and _static
 * forms) runs the multiply's scale-and-round inside the keyswitch digit
 * transforms, one workgroup per ciphertext (n >= 1024, q of one or two limbs,
 * w = 2^32); smaller batches take the separate reconstruction and digit
 * kernels, which give one ciphertext more parallelism. 0 (the default): eight
 * per compute unit (2048 on MI355X). Results are identical either way. A
 * tuning knob with no counterpart in the reference. */
int yh_ctx_set_fuse_min_batch(yh_ctx* ctx, int min_batch);

/* ------------------------------------------------------------ wire format
 * Ring elements as bytes, for clients of a serving process and for keeping
 * keys across runs. No counterpart in the reference, whose only serialisation
 * is the Python 2 to_JSON of experiments/modq/polynomial.py:158-165.
 * Little-endian: "YHW1", u16 version (1), u16 kind, u32 n, u32 q_limbs,
 * u64 t, u32 w_bits, u32 count, q (q_limbs x u64), then count x q_limbs x n
 * u64 coefficients in the device layout ([count][L][n]), then the CRC-32
 * (IEEE, as zlib.crc32) of every preceding byte. Host memory on both sides:
 * move data to and from the device with yh_memcpy_*. Encoding refuses
 * coefficients outside [0, q) ([0, t) for plaintexts); decoding checks the
 * ring parameters (n, q, t, w_bits) against the caller's, the kind (unless
 * 0), the exact length, the checksum and every coefficient's range. The
 * resident keyswitch key is stored as its coefficient-domain gamma_k
 * (YH_WIRE_KEYSWITCH_KEY, l_w polynomials) and rebuilt with yh_evk_create. */
#define YH_WIRE_CIPHERTEXT 1
#define YH_WIRE_KEYSWITCH_KEY 2
#define YH_WIRE_PUBLIC_KEY 3
#define YH_WIRE_SECRET_KEY 4
#define YH_WIRE_PLAINTEXT 5
typedef struct yh_wire_header {
  int version;
  int kind;
  int n;
  int q_limbs;
  uint64_t t;
  int w_bits;
  int count;
  uint64_t q[8];
} yh_wire_header;
/* Bytes of a blob holding `count` polynomials. */
size_t yh_wire_size(int n, int q_limbs, int count);
/* Writes `count` polynomials from h_src ([count][q_limbs][n]) into out;
 * *len receives the blob size (out NULL: size query only). */
int yh_wire_encode(int n, const uint64_t* q, int q_limbs, uint64_t t, int w_bits, int kind, const uint64_t* h_src,
                   int count, void* out, size_t cap, size_t* len);
/* Reads a blob into h_dst ([cap][q_limbs][n]); *count receives the number of
 * polynomials (h_dst NULL: validate and count only). */
int yh_wire_decode(int n, const uint64_t* q, int q_limbs, uint64_t t, int w_bits, int kind, const void* in,
                   size_t len, uint64_t* h_dst, int cap, int* count);
/* Header of a blob (to pick the ring or size buffers before decoding). */
int yh_wire_peek(const void* in, size_t len, yh_wire_header* hdr);

/* ------------------------------------------------------------ batching helpers */
/* Device bytes one operation of `op` (0 ring_mul, 1 yashe_mul, 2 keyswitch,
 * 3 mul+keyswitch, 4 add) needs per batch element, inputs and outputs
 * included: the admission unit of a serving loop. */
size_t yh_mem_footprint(const yh_ring* ring, int op);

/* ------------------------------------------------------------ request loop
 * Continuous batching of homomorphic requests over resident keys and
 * ciphertexts of several rings. One host thread queues arrived requests per
 * (ring, kind); a queue with fewer than `depth` groups in flight forms its
 * next group from up to max_batch of its oldest requests (queues taken in
 * the order of their oldest request) while the HBM reserved for work in
 * flight (per request: yh_mem_footprint of its operation, or 3 ring elements
 * for an encryption, plus its result element and its decryption) stays
 * within hbm_budget and the ring has a free result slot, so batches grow
 * with the load; it launches each group as one batched list operation on
 * the ring's own stream; decrypts
 * every result on the device and copies the decryption to the request's
 * host buffer. Encryption noise (ternary s, D_{Z,8} on [-48, 48]) is drawn
 * on the device from a counter-based generator keyed by (seed, rid,
 * coefficient), so results do not depend on how requests were batched.
 * Calls that overlap on one server are not allowed. No counterpart in the
 * reference: it runs one Ciphertext operator at a time
 * (src/ciphertext.cpp:38-107; Yashe::encrypt / decrypt, src/yashe.cpp:104-197),
 * each followed by cudaDeviceSynchronize. */
#define YH_REQ_ADD 0    /* c_a + c_b (Ciphertext::operator+) */
#define YH_REQ_MULKS 1  /* c_a * c_b then keyswitch (Ciphertext::operator*, convert) */
#define YH_REQ_ENCDEC 2 /* encrypt msg under the ring's public key h */

typedef struct yh_serve_ring {
  yh_ring* ring;          /* with t set; the server gives it its own stream */
  const yh_evk* evk;      /* keyswitch key (NULL: no YH_REQ_MULKS on this ring) */
  const uint64_t* d_f;    /* secret key f [L][n] (decryption) */
  const uint64_t* d_h;    /* public key h [L][n] (encryption) */
  const uint64_t* d_pool; /* resident ciphertexts [pool][L][n] */
  int pool;
  int slots;              /* result ciphertexts in flight at once */
} yh_serve_ring;

typedef struct yh_request {
  double arrival;         /* seconds after the start of the run */
  uint64_t rid;           /* request id: keys the encryption noise */
  int kind;               /* YH_REQ_* */
  int ring;               /* index into the server's rings */
  int a, b;               /* pool indices (YH_REQ_ADD, YH_REQ_MULKS) */
  const uint64_t* msg;    /* YH_REQ_ENCDEC: host message [n], values in [0, t) */
  uint64_t* result;       /* host [n]: receives the decryption of the result */
  uint64_t* output;       /* host [L][n] or NULL: receives the result ciphertext */
  double done;            /* out: seconds (device clock, same origin) when the result was complete */
} yh_request;

typedef struct yh_serve_stats {
  double wall_s;          /* from the start until the last result was on the host */
  long groups;            /* batched launches */
  double mean_batch;      /* requests per launch */
  size_t peak_reserved;   /* most HBM reserved at once */
  int max_inflight;       /* most groups in flight at once */
} yh_serve_stats;

typedef struct yh_server yh_server;
/* rings: nrings descriptors (copied; the rings and keys are held until
 * yh_server_destroy, and refuse destruction meanwhile). max_batch: most
 * requests per launch, 1..4096; depth: groups in flight per (ring, kind),
 * 1..64. */
int yh_server_create(yh_ctx* ctx, const yh_serve_ring* rings, int nrings, size_t hbm_budget, int max_batch,
                     int depth, uint64_t seed, yh_server** out);
int yh_server_set_budget(yh_server* server, size_t hbm_budget);
/* Admission footprint of one request of `kind` on ring index `ring` (bytes). */
size_t yh_server_footprint(const yh_server* server, int ring, int kind);
/* Serves `count` requests (any order; arrivals relative to this call) and
 * returns when every result is on the host. Refuses (YH_ERR_ARG, nothing
 * run) a request with a bad kind, ring or pool index, a message value >= t,
 * or a footprint above the whole budget. */
int yh_server_run(yh_server* server, yh_request* reqs, int count, yh_serve_stats* stats);
int yh_server_destroy(yh_server* server);

#ifdef __cplusplus
}
#endif
#endif
