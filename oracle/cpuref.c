/*
 * cpuref.c - CPU restatement of the reference's polynomial-ring hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This is the oracle the GPU library is checked
 * against: only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py load it. Nothing in the product links or calls it.
 *
 * It follows the reference's algorithm step by step, in plain C, with its own
 * bignum code (no code shared with the product):
 *   CRT           bn_mod1_low            src/cuda_bn.cu:290-317
 *   NTT           radix-2 Stockham stage src/cuda_functions.cu:404-458, twiddle
 *                 tables W[j] = w^j, WInv[j] = W[j]^-1, w = 7^((P-1)/N)
 *                 (src/cuda_functions.cu:878-918)
 *   Goldilocks    s_mul host branch: (uint128)a*b % P (src/cuda_functions.cu:351),
 *                 s_add/s_sub (:355-383) as exact arithmetic mod P
 *   pointwise     polynomialNTTMul (src/cuda_functions.cu:460-474)
 *   inverse scale NTTScale (src/cuda_functions.cu:397-402) - here a multiply by
 *                 N^-1 mod P; the reference's integer /N equals it whenever
 *                 N c < P, i.e. whenever the reference itself is exact.
 *   ICRT          sum_i M_i * ((r_i * (M_i^-1 mod p_i)) mod p_i) mod M
 *                 (cuPreICRT / cuICRT, src/cuda_bn.cu:854-968; Integer::icrt,
 *                 src/integer.cpp:84-117)
 *   product       callPolynomialMul (src/cuda_functions.cu:617-738): CRT both
 *                 operands, forward NTT, pointwise, inverse NTT, ICRT; NTT
 *                 length rule of common_multiplication (src/common.h:139-145).
 *   primes        gen_crt_primes (src/polynomial.h:898-955): distinct primes in
 *                 descending order until M > 2 * degree * q^2; the size is the
 *                 largest keeping the residue convolution below P (the
 *                 reference's 29-bit bucket is not exact for n > 64).
 * Composite operations (fold mod x^n + 1, mod q, scale-and-round, keyswitch
 * digits) are composed in oracle/cpuref.py on Python integers, and again in C
 * at the end of this file (or_yashe_mul, or_keyswitch) for the CPU baseline;
 * the tests check the two against each other.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

static const uint64_t GP = 0xFFFFFFFF00000001ull; /* PRIMEP, src/cuda_functions.cu:18 */

uint64_t or_gmul(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a * b) % GP); }
uint64_t or_gadd(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a + b) % GP); }
uint64_t or_gsub(uint64_t a, uint64_t b) { return (uint64_t)(((u128)a + GP - b) % GP); }

uint64_t or_gpow(uint64_t b, uint64_t e) {
  uint64_t r = 1;
  b %= GP;
  while (e) {
    if (e & 1) r = or_gmul(r, b);
    b = or_gmul(b, b);
    e >>= 1;
  }
  return r;
}

/* ----------------------------------------------------------------- NTT */

/* One radix-2 Stockham stage, NTTIteration (src/cuda_functions.cu:404-433)
 * for every j in [0, N/2). */
static void stockham_stage(const uint64_t* W, int N, int Ns, const uint64_t* in, uint64_t* out) {
  const int R = 2;
  for (int j = 0; j < N / R; ++j) {
    uint64_t v[2];
    int w_index = ((j % Ns) * N) / (Ns * R);
    for (int r = 0; r < R; ++r) v[r] = or_gmul(in[j + r * N / R], W[w_index * r]);
    /* butterfly (src/cuda_functions.cu:385-391) */
    uint64_t v0 = v[0], v1 = v[1];
    v[0] = or_gadd(v0, v1);
    v[1] = or_gsub(v0, v1);
    /* expand (src/cuda_functions.cu:393-395) */
    int idxD = (j / Ns) * Ns * R + (j % Ns);
    for (int r = 0; r < R; ++r) out[idxD + r * Ns] = v[r];
  }
}

/* In-place natural-order transform of length N (power of two, N | P-1).
 * inverse != 0: uses WInv and multiplies by N^-1 mod P. */
/* Twiddle tables are cached per length (init in src/cuda_functions.cu:878-918
 * rebuilds them whenever N changes). */
/* (per thread: the CPU baseline runs products on several threads) */
static __thread int tw_N = 0;
static __thread uint64_t* tw_W = NULL;
static __thread uint64_t* tw_WInv = NULL;

static int twiddles(int N) {
  if (tw_N == N) return 0;
  free(tw_W);
  free(tw_WInv);
  tw_W = (uint64_t*)malloc(sizeof(uint64_t) * N);
  tw_WInv = (uint64_t*)malloc(sizeof(uint64_t) * N);
  if (!tw_W || !tw_WInv) { tw_N = 0; return -2; }
  uint64_t w = or_gpow(7, (GP - 1) / (uint64_t)N);
  uint64_t wi = or_gpow(w, GP - 2);
  uint64_t x = 1, xi = 1;
  for (int j = 0; j < N; ++j) {
    tw_W[j] = x;      /* PowerMod(wN, j, P)                  */
    tw_WInv[j] = xi;  /* InvMod(W[j], P) = wN^-j            */
    x = or_gmul(x, w);
    xi = or_gmul(xi, wi);
  }
  tw_N = N;
  return 0;
}

int or_ntt(uint64_t* data, int N, int inverse) {
  if (N < 2 || (N & (N - 1))) return -1;
  if (twiddles(N)) return -2;
  const uint64_t* W = inverse ? tw_WInv : tw_W;
  uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * N);
  if (!tmp) return -2;
  uint64_t* a = data;
  uint64_t* b = tmp;
  for (int Ns = 1; Ns < N; Ns *= 2) {
    stockham_stage(W, N, Ns, a, b);
    uint64_t* t = a; a = b; b = t;
  }
  if (a != data) memcpy(data, a, sizeof(uint64_t) * N);
  if (inverse) {
    uint64_t ninv = or_gpow((uint64_t)N, GP - 2);
    for (int j = 0; j < N; ++j) data[j] = or_gmul(data[j], ninv);
  }
  free(tmp);
  return 0;
}

/* ----------------------------------------------------------------- bignum
 * Non-negative integers as little-endian 64-bit limbs of fixed capacity. */

#define BNMAX 48
typedef struct { int n; uint64_t d[BNMAX]; } bn;

static void bn_trim(bn* a) { while (a->n > 0 && a->d[a->n - 1] == 0) a->n--; }
static void bn_set_u64(bn* a, uint64_t v) { memset(a, 0, sizeof(*a)); a->d[0] = v; a->n = v ? 1 : 0; }
static void bn_mul_small(bn* a, uint64_t m) {
  u128 carry = 0;
  for (int i = 0; i < a->n; ++i) {
    u128 c = (u128)a->d[i] * m + carry;
    a->d[i] = (uint64_t)c;
    carry = c >> 64;
  }
  if (carry) a->d[a->n++] = (uint64_t)carry;
  bn_trim(a); /* m = 0 leaves zero limbs */
}
static void bn_add(bn* a, const bn* b) {
  int n = a->n > b->n ? a->n : b->n;
  uint64_t carry = 0;
  for (int i = 0; i < n; ++i) {
    uint64_t x = i < a->n ? a->d[i] : 0, y = i < b->n ? b->d[i] : 0;
    u128 s = (u128)x + y + carry;
    a->d[i] = (uint64_t)s;
    carry = (uint64_t)(s >> 64);
  }
  a->n = n;
  if (carry) a->d[a->n++] = carry;
}
static int bn_cmp(const bn* a, const bn* b) {
  if (a->n != b->n) return a->n < b->n ? -1 : 1;
  for (int i = a->n - 1; i >= 0; --i)
    if (a->d[i] != b->d[i]) return a->d[i] < b->d[i] ? -1 : 1;
  return 0;
}
static void bn_sub(bn* a, const bn* b) { /* a >= b */
  uint64_t br = 0;
  for (int i = 0; i < a->n; ++i) {
    uint64_t y = i < b->n ? b->d[i] : 0;
    u128 d = (u128)a->d[i] - y - br;
    a->d[i] = (uint64_t)d;
    br = (uint64_t)(d >> 64) ? 1 : 0;
  }
  bn_trim(a);
}
/* (skips zero top limbs: never counts the leading zeros of 0) */
static int bn_bits(const bn* a) {
  int n = a->n;
  while (n > 0 && a->d[n - 1] == 0) --n;
  return n ? 64 * (n - 1) + (64 - __builtin_clzll(a->d[n - 1])) : 0;
}
static void bn_shl1(bn* a) {
  uint64_t c = 0;
  for (int i = 0; i < a->n; ++i) {
    uint64_t nc = a->d[i] >> 63;
    a->d[i] = (a->d[i] << 1) | c;
    c = nc;
  }
  if (c) a->d[a->n++] = c;
}
/* r = a mod m by shift-and-subtract (bit-serial long division). */
static void bn_mod(const bn* a, const bn* m, bn* r) {
  bn_set_u64(r, 0);
  for (int i = bn_bits(a) - 1; i >= 0; --i) {
    bn_shl1(r);
    if ((a->d[i / 64] >> (i % 64)) & 1) {
      if (r->n == 0) r->n = 1;
      r->d[0] |= 1;
    }
    if (bn_cmp(r, m) >= 0) bn_sub(r, m);
  }
}
/* a mod p for a word-size p (bn_mod1_low semantics, src/cuda_bn.cu:290-317). */
static uint64_t bn_mod_word(const uint64_t* a, int n, uint64_t p) {
  u128 w = 0;
  for (int i = n - 1; i >= 0; --i) w = ((w << 64) | a[i]) % p;
  return (uint64_t)w;
}
static uint64_t inv_word(uint64_t a, uint64_t m) {
  /* signed 128-bit extended Euclid: m may use all 64 bits */
  __int128 t = 0, nt = 1, r = (__int128)m, nr = (__int128)(a % m);
  while (nr) {
    __int128 q = r / nr, tmp;
    tmp = t - q * nt; t = nt; nt = tmp;
    tmp = r - q * nr; r = nr; nr = tmp;
  }
  if (t < 0) t += (__int128)m;
  return (uint64_t)t;
}
static int is_prime_word(uint64_t x) {
  if (x < 2) return 0;
  for (uint64_t d = 2; d * d <= x; ++d)
    if (x % d == 0) return 0;
  return 1;
}

/* ----------------------------------------------------------------- CRT / ICRT */

/* res[i][j] = coef_j mod p_i; coefs limb-major [L][ncoef]. */
void or_crt(const uint64_t* coefs, int L, int ncoef, const uint64_t* primes, int R, uint64_t* res) {
  uint64_t* tmp = (uint64_t*)malloc(sizeof(uint64_t) * (L > 0 ? L : 1));
  for (int j = 0; j < ncoef; ++j) {
    for (int l = 0; l < L; ++l) tmp[l] = coefs[(size_t)l * ncoef + j];
    for (int i = 0; i < R; ++i) res[(size_t)i * ncoef + j] = bn_mod_word(tmp, L, primes[i]);
  }
  free(tmp);
}

/* out_j = sum_i M_i ((r_ij * invM_i) mod p_i) mod M, as LO limbs [LO][ncoef].
 * r_ij may be any value below 2^64 (the reference feeds the raw inverse-NTT
 * output, src/cuda_bn.cu:875-879). Returns -1 if M needs more than LO limbs. */
int or_icrt(const uint64_t* res, int ncoef, const uint64_t* primes, int R, uint64_t* out, int LO) {
  bn M, *Mi = (bn*)malloc(sizeof(bn) * R);
  uint64_t* inv = (uint64_t*)malloc(sizeof(uint64_t) * R);
  bn_set_u64(&M, 1);
  for (int i = 0; i < R; ++i) bn_mul_small(&M, primes[i]);
  for (int i = 0; i < R; ++i) {
    bn_set_u64(&Mi[i], 1);
    for (int k = 0; k < R; ++k)
      if (k != i) bn_mul_small(&Mi[i], primes[k]);
    inv[i] = inv_word(bn_mod_word(Mi[i].d, Mi[i].n, primes[i]), primes[i]);
  }
  int rc = 0;
  for (int j = 0; j < ncoef && rc == 0; ++j) {
    bn acc, r;
    bn_set_u64(&acc, 0);
    for (int i = 0; i < R; ++i) {
      uint64_t x = (uint64_t)(((u128)(res[(size_t)i * ncoef + j] % primes[i]) * inv[i]) % primes[i]);
      bn t = Mi[i];
      bn_mul_small(&t, x);
      bn_add(&acc, &t);
    }
    bn_mod(&acc, &M, &r);
    if (r.n > LO) rc = -1;
    for (int l = 0; l < LO; ++l) out[(size_t)l * ncoef + j] = l < r.n ? r.d[l] : 0;
  }
  free(Mi);
  free(inv);
  return rc;
}

/* ----------------------------------------------------------------- primes */

/* Descending primes p with terms*(p-1)^2 < P until their product exceeds
 * 2 * terms * 2^(2*qbits) (the M > 2*degree*q^2 rule with q < 2^qbits).
 * Returns the count, or -1 if cap is too small. */
int or_gen_primes(int terms, int qbits, uint64_t* out, int cap) {
  /* largest p with terms*(p-1)^2 < P */
  uint64_t lo = 3, hi = (1ull << 32);
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo + 1) / 2;
    u128 v = (u128)terms * (mid - 1) * (mid - 1);
    if (v < GP) lo = mid; else hi = mid - 1;
  }
  /* M >= 2^need_bits >= 2 * 2^ceil(log2 terms) * 2^(2 qbits) > 2 * terms * q^2 */
  int lt = 0;
  while ((1 << lt) < terms) ++lt;
  int need_bits = 1 + lt + 2 * qbits;
  uint64_t c = lo | 1;
  if (c > lo) c -= 2;
  int R = 0;
  bn M;
  bn_set_u64(&M, 1);
  while (bn_bits(&M) <= need_bits) {
    while (!is_prime_word(c)) c -= 2;
    if (R >= cap || M.n >= BNMAX - 1) return -1;
    out[R++] = c;
    bn_mul_small(&M, c);
    c -= 2;
  }
  return R;
}

/* ----------------------------------------------------------------- product */

/* Plain product mod x^N - 1 of a (na coefficients) and b (nb coefficients),
 * both limb-major with L limbs, over the given primes. Output [LO][N] exact
 * in [0, M). The residue convolution must stay below P (caller picks primes
 * with or_gen_primes). Returns 0 on success. */
int or_poly_mul_plain(const uint64_t* a, int na, const uint64_t* b, int nb, int L, int N, const uint64_t* primes,
                      int R, uint64_t* out, int LO) {
  if (na > N || nb > N) return -3;
  uint64_t* ra = (uint64_t*)calloc((size_t)R * na, sizeof(uint64_t));
  uint64_t* rb = (uint64_t*)calloc((size_t)R * nb, sizeof(uint64_t));
  uint64_t* res = (uint64_t*)calloc((size_t)R * N, sizeof(uint64_t));
  uint64_t* x = (uint64_t*)malloc(sizeof(uint64_t) * N);
  uint64_t* y = (uint64_t*)malloc(sizeof(uint64_t) * N);
  or_crt(a, L, na, primes, R, ra);
  or_crt(b, L, nb, primes, R, rb);
  for (int i = 0; i < R; ++i) {
    memset(x, 0, sizeof(uint64_t) * N);
    memset(y, 0, sizeof(uint64_t) * N);
    memcpy(x, ra + (size_t)i * na, sizeof(uint64_t) * na);
    memcpy(y, rb + (size_t)i * nb, sizeof(uint64_t) * nb);
    or_ntt(x, N, 0);
    or_ntt(y, N, 0);
    for (int j = 0; j < N; ++j) x[j] = or_gmul(x[j], y[j]);
    or_ntt(x, N, 1);
    memcpy(res + (size_t)i * N, x, sizeof(uint64_t) * N);
  }
  int rc = or_icrt(res, N, primes, R, out, LO);
  free(ra); free(rb); free(res); free(x); free(y);
  return rc;
}

/* ----------------------------------------------------------------- Z_q helpers
 * Polynomial inverse in Z_q[x]/(x^n + 1) for test key generation
 * (Polynomial::InvMod, src/polynomial.h:1338-1367, NTL InvMod). q < 2^(64 QL). */

#define QL_MAX 8
typedef struct { int L; uint64_t q[QL_MAX]; uint64_t mu[QL_MAX + 2]; } zq;

/* mu = floor(2^(128 L) / q), by bit-serial long division (once per modulus). */
static void zq_init(zq* Z) {
  bn num, m, quo, rem;
  memset(&num, 0, sizeof(num));
  num.d[2 * Z->L] = 1;
  num.n = 2 * Z->L + 1;
  memset(&m, 0, sizeof(m));
  memcpy(m.d, Z->q, sizeof(uint64_t) * Z->L);
  m.n = Z->L;
  bn_trim(&m);
  bn_set_u64(&rem, 0);
  memset(&quo, 0, sizeof(quo));
  for (int i = bn_bits(&num) - 1; i >= 0; --i) {
    bn_shl1(&rem);
    if ((num.d[i / 64] >> (i % 64)) & 1) {
      if (rem.n == 0) rem.n = 1;
      rem.d[0] |= 1;
    }
    if (bn_cmp(&rem, &m) >= 0) {
      bn_sub(&rem, &m);
      quo.d[i / 64] |= 1ull << (i % 64);
    }
  }
  for (int l = 0; l < Z->L + 2; ++l) Z->mu[l] = quo.d[l];
}
static int zq_geq(const zq* Z, const uint64_t* a, const uint64_t* b) {
  for (int l = Z->L - 1; l >= 0; --l)
    if (a[l] != b[l]) return a[l] > b[l];
  return 1;
}
static void zq_sub(const zq* Z, uint64_t* r, const uint64_t* a, const uint64_t* b) {
  uint64_t t[QL_MAX];
  uint64_t br = 0;
  for (int l = 0; l < Z->L; ++l) {
    u128 d = (u128)a[l] - b[l] - br;
    t[l] = (uint64_t)d;
    br = (uint64_t)(d >> 64) ? 1 : 0;
  }
  if (br) {
    uint64_t c = 0;
    for (int l = 0; l < Z->L; ++l) {
      u128 s = (u128)t[l] + Z->q[l] + c;
      t[l] = (uint64_t)s;
      c = (uint64_t)(s >> 64);
    }
  }
  memcpy(r, t, sizeof(uint64_t) * Z->L);
}
/* r = a*b mod q (a, b < q): schoolbook product, Barrett quotient
 * qhat = floor(x*mu / 2^(128 L)) in {q_true - 2 .. q_true}, then correction. */
static void zq_mul(const zq* Z, uint64_t* r, const uint64_t* a, const uint64_t* b) {
  const int L = Z->L;
  uint64_t x[2 * QL_MAX] = {0}, xm[3 * QL_MAX + 3] = {0}, qq[QL_MAX + 2] = {0}, t[QL_MAX + 1];
  for (int i = 0; i < L; ++i) {
    u128 carry = 0;
    for (int k = 0; k < L; ++k) {
      u128 c = (u128)a[i] * b[k] + x[i + k] + carry;
      x[i + k] = (uint64_t)c;
      carry = c >> 64;
    }
    x[i + L] = (uint64_t)carry;
  }
  for (int i = 0; i < 2 * L; ++i) {
    u128 carry = 0;
    for (int k = 0; k < L + 2; ++k) {
      u128 c = (u128)x[i] * Z->mu[k] + xm[i + k] + carry;
      xm[i + k] = (uint64_t)c;
      carry = c >> 64;
    }
    xm[i + L + 2] = (uint64_t)carry;
  }
  const uint64_t* qh = xm + 2 * L; /* L + 2 limbs */
  /* t = x - qh*q over L+1 limbs */
  for (int i = 0; i < L + 1; ++i) {
    u128 carry = 0;
    for (int k = 0; k < L && i + k < L + 1; ++k) {
      u128 c = (u128)qh[i] * Z->q[k] + qq[i + k] + carry;
      qq[i + k] = (uint64_t)c;
      carry = c >> 64;
    }
    if (i + L < L + 1) qq[i + L] += (uint64_t)carry;
  }
  uint64_t br = 0;
  for (int l = 0; l < L + 1; ++l) {
    u128 d = (u128)x[l] - qq[l] - br;
    t[l] = (uint64_t)d;
    br = (uint64_t)(d >> 64) ? 1 : 0;
  }
  for (int it = 0; it < 4; ++it) {
    /* t >= q ? */
    int ge = t[L] != 0 || zq_geq(Z, t, Z->q);
    if (!ge) break;
    br = 0;
    for (int l = 0; l < L + 1; ++l) {
      uint64_t ql = l < L ? Z->q[l] : 0;
      u128 d = (u128)t[l] - ql - br;
      t[l] = (uint64_t)d;
      br = (uint64_t)(d >> 64) ? 1 : 0;
    }
  }
  memcpy(r, t, sizeof(uint64_t) * L);
}
static int zq_is_zero(const zq* Z, const uint64_t* a) {
  for (int l = 0; l < Z->L; ++l)
    if (a[l]) return 0;
  return 1;
}
static void zq_pow(const zq* Z, uint64_t* r, const uint64_t* b, const uint64_t* e) {
  uint64_t acc[QL_MAX] = {1}, base[QL_MAX];
  memcpy(base, b, sizeof(uint64_t) * Z->L);
  for (int l = 1; l < Z->L; ++l) acc[l] = 0;
  for (int l = 0; l < Z->L; ++l)
    for (int k = 0; k < 64; ++k) {
      if ((e[l] >> k) & 1) zq_mul(Z, acc, acc, base);
      zq_mul(Z, base, base, base);
    }
  memcpy(r, acc, sizeof(uint64_t) * Z->L);
}
/* inverse of a mod prime q by Fermat */
static void zq_inv(const zq* Z, uint64_t* r, const uint64_t* a) {
  uint64_t e[QL_MAX];
  memcpy(e, Z->q, sizeof(uint64_t) * Z->L);
  /* e = q - 2 */
  uint64_t br = 2;
  for (int l = 0; l < Z->L && br; ++l) {
    uint64_t old = e[l];
    e[l] = old - br;
    br = old < br ? 1 : 0;
  }
  zq_pow(Z, r, a, e);
}

/* f (coefficient-major: f[j*QL + l]) inverse mod (x^n + 1, q), q prime.
 * Extended Euclid over Z_q[x]. Returns 0 on success, 1 if not invertible. */
int or_poly_inv(const uint64_t* f, int n, const uint64_t* q, int QL, uint64_t* out) {
  zq Z;
  Z.L = QL;
  memset(Z.q, 0, sizeof(Z.q));
  memcpy(Z.q, q, sizeof(uint64_t) * QL);
  zq_init(&Z);
  size_t W = (size_t)QL;
  /* r0 = x^n + 1, r1 = f; s0 = 0, s1 = 1 (track only the f-coefficient) */
  int cap = n + 2;
  uint64_t* r0 = (uint64_t*)calloc((size_t)cap * W, 8);
  uint64_t* r1 = (uint64_t*)calloc((size_t)cap * W, 8);
  uint64_t* s0 = (uint64_t*)calloc((size_t)cap * W, 8);
  uint64_t* s1 = (uint64_t*)calloc((size_t)cap * W, 8);
  uint64_t* tmp = (uint64_t*)calloc((size_t)cap * W, 8);
  r0[0] = 1;
  r0[(size_t)n * W] = 1;
  int d0 = n, d1 = -1;
  for (int j = 0; j < n; ++j) {
    memcpy(r1 + j * W, f + j * W, 8 * W);
    if (!zq_is_zero(&Z, r1 + j * W)) d1 = j;
  }
  s1[0] = 1;
  int ds0 = -1, ds1 = 0;
  int rc = 1;
  uint64_t inv[QL_MAX], c[QL_MAX], t[QL_MAX];
  while (d1 >= 0) {
    if (d1 == 0) {
      /* r1 is a unit: inverse = s1 / r1[0] */
      zq_inv(&Z, inv, r1);
      for (int j = 0; j < n; ++j) {
        if (j <= ds1) zq_mul(&Z, out + j * W, s1 + j * W, inv);
        else memset(out + j * W, 0, 8 * W);
      }
      rc = 0;
      break;
    }
    /* r0 = r0 - c x^k r1 repeatedly (division), s0 likewise */
    zq_inv(&Z, inv, r1 + (size_t)d1 * W);
    while (d0 >= d1) {
      int k = d0 - d1;
      zq_mul(&Z, c, r0 + (size_t)d0 * W, inv);
      for (int j = 0; j <= d1; ++j) {
        zq_mul(&Z, t, c, r1 + (size_t)j * W);
        zq_sub(&Z, r0 + (size_t)(j + k) * W, r0 + (size_t)(j + k) * W, t);
      }
      for (int j = 0; j <= ds1; ++j) {
        zq_mul(&Z, t, c, s1 + (size_t)j * W);
        zq_sub(&Z, s0 + (size_t)(j + k) * W, s0 + (size_t)(j + k) * W, t);
      }
      if (ds1 + k > ds0) ds0 = ds1 + k;
      while (d0 >= 0 && zq_is_zero(&Z, r0 + (size_t)d0 * W)) --d0;
      while (ds0 >= 0 && zq_is_zero(&Z, s0 + (size_t)ds0 * W)) --ds0;
    }
    /* swap (r0,s0) <-> (r1,s1) */
    uint64_t* p;
    p = r0; r0 = r1; r1 = p;
    p = s0; s0 = s1; s1 = p;
    int d;
    d = d0; d0 = d1; d1 = d;
    d = ds0; ds0 = ds1; ds1 = d;
    if (ds1 >= cap - 1 || ds0 >= cap - 1) break; /* cannot happen for n-degree moduli */
  }
  free(r0); free(r1); free(s0); free(s1); free(tmp);
  return rc;
}

/* ----------------------------------------------------------------- composite operations
 * The YASHE multiply and keyswitch composed in C (the same steps as
 * oracle/cpuref.py's yashe_mul / keyswitch, which the tests pin), so a CPU
 * baseline spends no time in an interpreter. Operands are limb-major
 * [QL][n] with coefficients in [0, q). */

/* Signed bignum: magnitude + sign. */
typedef struct { int neg; bn m; } sbn;

static void bn_from_col(bn* a, const uint64_t* limbs, int L, int n, int j) {
  memset(a, 0, sizeof(*a));
  for (int l = 0; l < L; ++l) a->d[l] = limbs[(size_t)l * n + j];
  a->n = L;
  bn_trim(a);
}
static void bn_to_col(const bn* a, uint64_t* limbs, int L, int n, int j) {
  for (int l = 0; l < L; ++l) limbs[(size_t)l * n + j] = l < a->n ? a->d[l] : 0;
}
/* q, r with a = q m + r, 0 <= r < m (bit-serial long division). */
static void bn_divmod(const bn* a, const bn* m, bn* q, bn* r) {
  memset(q, 0, sizeof(*q));
  bn_set_u64(r, 0);
  const int nb = bn_bits(a);
  for (int i = nb - 1; i >= 0; --i) {
    bn_shl1(r);
    if ((a->d[i / 64] >> (i % 64)) & 1) {
      if (r->n == 0) r->n = 1;
      r->d[0] |= 1;
    }
    if (bn_cmp(r, m) >= 0) {
      bn_sub(r, m);
      q->d[i / 64] |= 1ull << (i % 64);
      if (q->n < i / 64 + 1) q->n = i / 64 + 1;
    }
  }
  bn_trim(q);
}
/* a - b as a signed value (a, b >= 0). */
static void sbn_sub(const bn* a, const bn* b, sbn* out) {
  if (bn_cmp(a, b) >= 0) {
    out->neg = 0;
    out->m = *a;
    bn_sub(&out->m, b);
  } else {
    out->neg = 1;
    out->m = *b;
    bn_sub(&out->m, a);
  }
}
/* v mod q in [0, q) for a signed v. */
static void sbn_mod(const sbn* v, const bn* q, bn* out) {
  bn quo;
  bn_divmod(&v->m, q, &quo, out);
  if (v->neg && out->n) {
    bn t = *q;
    bn_sub(&t, out);
    *out = t;
  }
}

/* c = [round(t * c1*c2 / q)]_q, the product mod x^n + 1 over Z
 * (Ciphertext::operator*, src/ciphertext.cpp:38-73, with the rounding of
 * :55-63 applied to t*g: floor division, non-negative remainder, round up
 * when 2*rem > q). Returns 0 on success. */
int or_yashe_mul(const uint64_t* c1, const uint64_t* c2, int n, const uint64_t* q, int QL, uint64_t t, uint64_t* out) {
  bn Q;
  memset(&Q, 0, sizeof(Q));
  memcpy(Q.d, q, sizeof(uint64_t) * QL);
  Q.n = QL;
  bn_trim(&Q);
  uint64_t pr[256];
  const int R = or_gen_primes(n, bn_bits(&Q), pr, 256);
  if (R < 0) return -1;
  bn M;
  bn_set_u64(&M, 1);
  for (int i = 0; i < R; ++i) bn_mul_small(&M, pr[i]);
  const int LO = M.n, N = 2 * n;
  uint64_t* prod = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)LO * N);
  if (!prod) return -2;
  int rc = or_poly_mul_plain(c1, n, c2, n, QL, N, pr, R, prod, LO);
  for (int j = 0; j < n && rc == 0; ++j) {
    bn lo, hi, x, quo, rem;
    sbn g;
    bn_from_col(&lo, prod, LO, N, j);
    bn_from_col(&hi, prod, LO, N, j + n);
    sbn_sub(&lo, &hi, &g); /* fold mod x^n + 1 (src/cuda_functions.cu:1066-1072) */
    x = g.m;
    bn_mul_small(&x, t);
    bn_divmod(&x, &Q, &quo, &rem);
    /* floor division of the signed t*g: quotient -(quo + [rem != 0]) and
     * remainder q - rem when g < 0 and rem != 0 */
    sbn v;
    if (g.neg && rem.n) {
      bn one;
      bn_set_u64(&one, 1);
      bn_add(&quo, &one);
      bn r2 = Q;
      bn_sub(&r2, &rem);
      rem = r2;
    }
    v.neg = g.neg && quo.n;
    v.m = quo;
    bn twice = rem;
    bn_shl1(&twice);
    if (bn_cmp(&twice, &Q) > 0) { /* round up: v + 1 */
      bn one;
      bn_set_u64(&one, 1);
      if (!v.neg) {
        bn_add(&v.m, &one);
      } else {
        bn_sub(&v.m, &one);
        if (!v.m.n) v.neg = 0;
      }
    }
    bn r;
    sbn_mod(&v, &Q, &r);
    bn_to_col(&r, out, QL, n, j);
  }
  free(prod);
  return rc;
}

/* c' = [sum_k D_k(c) gamma_k]_(x^n + 1, q) with D_k(c) = floor(c / w^k) mod w,
 * w = 2^w_bits, for k < lw (Ciphertext::keyswitch / worddecomp,
 * src/ciphertext.cpp:81-130 with the base, mask and shift the reference
 * intends). gammas: [lw][QL][n]. Returns 0 on success. */
int or_keyswitch(const uint64_t* c, const uint64_t* gammas, int n, const uint64_t* q, int QL, int w_bits, int lw,
                 uint64_t* out) {
  bn Q;
  memset(&Q, 0, sizeof(Q));
  memcpy(Q.d, q, sizeof(uint64_t) * QL);
  Q.n = QL;
  bn_trim(&Q);
  const int qbits = bn_bits(&Q) > w_bits ? bn_bits(&Q) : w_bits;
  uint64_t pr[256];
  const int R = or_gen_primes(n, qbits, pr, 256);
  if (R < 0) return -1;
  bn M;
  bn_set_u64(&M, 1);
  for (int i = 0; i < R; ++i) bn_mul_small(&M, pr[i]);
  const int LO = M.n, N = 2 * n;
  uint64_t* prod = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)LO * N);
  uint64_t* dig = (uint64_t*)calloc((size_t)QL * n, sizeof(uint64_t)); /* [QL][n], digits in limb 0 */
  bn* acc = (bn*)calloc((size_t)N, sizeof(bn));
  if (!prod || !dig || !acc) { free(prod); free(dig); free(acc); return -2; }
  const uint64_t mask = w_bits >= 64 ? ~0ull : ((1ull << w_bits) - 1);
  int rc = 0;
  for (int k = 0; k < lw && rc == 0; ++k) {
    int any = 0;
    for (int j = 0; j < n; ++j) {
      const int bit = k * w_bits, l = bit / 64, off = bit % 64;
      uint64_t v = l < QL ? c[(size_t)l * n + j] >> off : 0;
      if (off && off + w_bits > 64 && l + 1 < QL) v |= c[(size_t)(l + 1) * n + j] << (64 - off);
      dig[j] = v & mask;
      any |= dig[j] != 0;
    }
    if (!any) continue;
    rc = or_poly_mul_plain(dig, n, gammas + (size_t)k * QL * n, n, QL, N, pr, R, prod, LO);
    for (int j = 0; j < N && rc == 0; ++j) {
      bn p;
      bn_from_col(&p, prod, LO, N, j);
      bn_add(&acc[j], &p);
    }
  }
  for (int j = 0; j < n && rc == 0; ++j) {
    sbn g;
    bn r;
    sbn_sub(&acc[j], &acc[j + n], &g);
    sbn_mod(&g, &Q, &r);
    bn_to_col(&r, out, QL, n, j);
  }
  free(prod);
  free(dig);
  free(acc);
  return rc;
}
