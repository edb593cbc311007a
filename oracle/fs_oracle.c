/*
 * fs_oracle.c -- CPU oracle for the FactSieve n!+-1 factor sieve.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithm, used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg as the checker.  The product path (the HIP
 * library under factsieve_amd/) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * exact: unsigned __int128 products and Montgomery REDC with R = 2^64, so
 * every residue is the canonical value in [0, p).
 *
 * Results this oracle must reproduce (reference self-test, cl_sieve.cpp:1282-1375):
 *   p in [1e8, 1.01e8),  n in [1e6, 2e6):  1071 factors, 54211 candidates, 000004F4F744CA97
 *   p in [127, 1e5),     n in [127, 1e6):  42770 factors, 9566 candidates,  0000000065C074F0
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <omp.h>

typedef unsigned __int128 u128;

/* 16-byte factor record, same layout as `factor` (cl_sieve.h:4-7, iterate.cl:13-16). */
typedef struct {
  uint64_t p;
  int32_t nc; /* -n: p | n!-1 ; +n: p | n!+1 */
  int32_t pad;
} ora_factor;

/* ---------------------------------------------------------------- arithmetic */

/* p^-1 mod 2^64 by Newton iteration to a fixed point (getsegprimes.cl:73-77,
 * verifyprime.c:19-24). */
uint64_t ora_invert(uint64_t p) {
  uint64_t x = 1, prev = 0;
  while (x != prev) { prev = x; x *= 2 - p * x; }
  return x;
}

/* Montgomery product a*b*2^-64 mod p, subtractive REDC with q = p^-1 mod 2^64
 * (iterate.cl:41-47, verifyprime.c:27-43).  Canonical output for a*b < p*2^64. */
uint64_t ora_mmul(uint64_t a, uint64_t b, uint64_t p, uint64_t q) {
  u128 ab = (u128)a * b;
  uint64_t lo = (uint64_t)ab, hi = (uint64_t)(ab >> 64);
  uint64_t m = lo * q;
  uint64_t mp = (uint64_t)(((u128)m * p) >> 64);
  uint64_t r = hi - mp;
  return (hi < mp) ? r + p : r;
}

/* Overflow-safe (a + b) mod p for a, b in [0, p) (iterate.cl:49-54). */
uint64_t ora_addmod(uint64_t a, uint64_t b, uint64_t p) {
  uint64_t c = (a >= p - b) ? p : 0;
  return a + b - c;
}

/* Plain modular product via 128-bit remainder (independent of Montgomery form). */
static inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t p) {
  return (uint64_t)(((u128)a * b) % p);
}

/* Per-modulus Montgomery constants as produced by getsegprimes
 * (getsegprimes.cl:214-226): q, one = R mod p, two, nmo = p - one, r2 = R^2 mod p. */
typedef struct { uint64_t p, q, one, two, nmo, r2; } ora_mont;

void ora_mont_init(uint64_t p, ora_mont* m) {
  m->p = p;
  m->q = ora_invert(p);
  m->one = (0 - p) % p;
  m->nmo = p - m->one;
  m->two = ora_addmod(m->one, m->one, p);
  uint64_t r2 = ora_addmod(m->two, m->two, p);
  for (int i = 0; i < 5; ++i) r2 = ora_mmul(r2, r2, p, m->q); /* 4^(2^5) = 2^64 */
  m->r2 = r2;
}

/* x -> x * 2^64 mod p */
uint64_t ora_to_mont(uint64_t x, const ora_mont* m) { return ora_mmul(x % m->p, m->r2, m->p, m->q); }

/* ---------------------------------------------------------- candidate set */

/* Odd primes 7..113 whose multiples the reference's bit tables remove
 * (getsegprimes.cl:115-167); 2, 3, 5 are removed by the mod-30 wheel
 * (getsegprimes.cl:112-113, cl_sieve.cpp:360-419). */
static const uint32_t SMALL_SIEVE[27] = {7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61,
                                         67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113};

/* 1 if N is coprime to every prime <= 113. */
int ora_sieve_survivor(uint64_t N) {
  if ((N & 1) == 0 || N % 3 == 0 || N % 5 == 0) return 0;
  for (int i = 0; i < 27; ++i)
    if (N % SMALL_SIEVE[i] == 0) return 0;
  return 1;
}

/* Base-2 strong probable-prime test (getsegprimes.cl:79-108), written here with
 * plain 128-bit remainders: N-1 = d*2^t, pass if 2^d = 1 or 2^(d*2^s) = -1 for
 * some 0 <= s < t.  (For d = 1 the reference's shift by 64 is masked to a shift
 * by 0; for every N = 2^t+1 that survives the 113-sieve both forms agree, which
 * tests/test_oracle.py checks with big integers.) */
int ora_spsp2(uint64_t N) {
  uint64_t d = N - 1;
  int t = 0;
  while ((d & 1) == 0) { d >>= 1; ++t; }
  uint64_t a = 1, b = 2 % N, e = d;
  while (e) {
    if (e & 1) a = mulmod(a, b, N);
    b = mulmod(b, b, N);
    e >>= 1;
  }
  if (a == 1 || a == N - 1) return 1;
  for (int s = 1; s < t; ++s) {
    a = mulmod(a, a, N);
    if (a == N - 1) return 1;
  }
  return 0;
}

/* Candidate predicate: p in the sieve set iff it survives the 113-sieve and is
 * a base-2 strong probable prime (getsegprimes.cl:184-229). */
int ora_is_candidate(uint64_t p) { return p > 113 && ora_sieve_survivor(p) && ora_spsp2(p); }

/* Ascending candidates in [pmin, pmax).  Returns the count; writes at most cap. */
uint64_t ora_candidates(uint64_t pmin, uint64_t pmax, uint64_t* out, uint64_t cap) {
  uint64_t n = 0;
  if (pmin >= pmax) return 0;
  for (uint64_t p = pmin | 1; p < pmax && p >= pmin; p += 2) {
    if (ora_is_candidate(p)) {
      if (out && n < cap) out[n] = p;
      ++n;
    }
    if (p >= UINT64_MAX - 1) break;
  }
  return n;
}

/* Mod-30 wheel start >= start and its index into the gap table {4,2,4,2,4,6,2,6}
 * as computed by findWheelOffset (cl_sieve.cpp:360-419): the first N >= start
 * coprime to 30, and idx such that N + gap[idx] is the next number coprime to 30
 * (gap[i] follows the residue res[i] below). */
void ora_find_wheel_offset(uint64_t* start, int32_t* index) {
  static const int32_t res[8] = {7, 11, 13, 17, 19, 23, 29, 1};
  uint64_t N = *start;
  while (N % 2 == 0 || N % 3 == 0 || N % 5 == 0) ++N;
  int32_t r = (int32_t)(N % 30), idx = -1;
  for (int i = 0; i < 8; ++i)
    if (res[i] == r) idx = i;
  *start = N;
  *index = idx;
}

/* ------------------------------------------------------------- power table */

/* Primes <= limit by a plain sieve of Eratosthenes (replaces the reference's
 * primesieve_generate_primes(2, nmin) call, cl_sieve.cpp:797). Caller frees. */
uint32_t* ora_small_primes(uint32_t limit, uint64_t* count) {
  uint8_t* comp = (uint8_t*)calloc((size_t)limit + 1, 1);
  uint64_t n = 0, cap = limit < 100 ? 64 : (uint64_t)(1.3 * limit / (__builtin_log((double)limit) - 1.1)) + 64;
  uint32_t* pr = (uint32_t*)malloc(cap * sizeof(uint32_t));
  for (uint64_t i = 2; i <= limit; ++i) {
    if (comp[i]) continue;
    if (n == cap) { cap *= 2; pr = (uint32_t*)realloc(pr, cap * sizeof(uint32_t)); }
    pr[n++] = (uint32_t)i;
    for (uint64_t j = i * i; j <= limit; j += i) comp[j] = 1;
  }
  free(comp);
  *count = n;
  return pr;
}

/* Legendre exponent of prime q in nmin! and the bit just below its MSB
 * (getPower, cl_sieve.cpp:773-788). */
void ora_get_power(uint32_t q, uint32_t nmin, uint32_t* exp, uint32_t* curbit) {
  uint32_t total = 0;
  uint64_t qk = q;
  while (qk <= nmin) { total += (uint32_t)(nmin / qk); qk *= q; }
  uint32_t cb = 0x80000000u;
  if (total > 1) cb >>= (__builtin_clz(total) + 1);
  *exp = total;
  *curbit = cb;
}

/* Power table with runs of equal exponent merged into products below 2^64-1
 * (setupPowerTable, cl_sieve.cpp:804-838).  Entry 0 is the prime 2.  Returns
 * the number of terms; bases/exps/curbits are malloc'd (caller frees). */
uint64_t ora_power_table(uint32_t nmin, uint64_t** bases, uint32_t** exps, uint32_t** curbits) {
  uint64_t np;
  uint32_t* pr = ora_small_primes(nmin, &np);
  uint64_t* B = (uint64_t*)malloc((np + 1) * sizeof(uint64_t));
  uint32_t* E = (uint32_t*)malloc((np + 1) * sizeof(uint32_t));
  uint32_t* C = (uint32_t*)malloc((np + 1) * sizeof(uint32_t));
  uint64_t m = 0;
  if (np > 0) {
    ora_get_power(pr[0], nmin, &E[0], &C[0]);
    B[0] = pr[0];
    m = 1;
    uint64_t i = 1;
    while (i < np) {
      uint32_t e, c;
      ora_get_power(pr[i], nmin, &e, &c);
      B[m] = pr[i]; E[m] = e; C[m] = c;
      for (++i; i < np; ++i) {
        uint32_t e2, c2;
        ora_get_power(pr[i], nmin, &e2, &c2);
        if (e2 != E[m]) break;
        u128 pp = (u128)B[m] * pr[i];
        if (pp < (u128)UINT64_MAX) B[m] = (uint64_t)pp;
        else break;
      }
      ++m;
    }
  }
  free(pr);
  *bases = B; *exps = E; *curbits = C;
  return m;
}

/* mont(nmin! mod p) from the power table (setup.cl:49-103): entry 0 (base 2)
 * by square-and-double, the others by left-to-right square-and-multiply. */
uint64_t ora_setup_residue(const ora_mont* M, const uint64_t* B, const uint32_t* E,
                           const uint32_t* C, uint64_t m) {
  uint64_t p = M->p, q = M->q, s6 = M->one;
  if (m == 0) return s6;
  {
    uint32_t e = E[0], cb = C[0];
    uint64_t a = M->two;
    while (cb) {
      a = ora_mmul(a, a, p, q);
      if (e & cb) a = ora_addmod(a, a, p);
      cb >>= 1;
    }
    s6 = a;
  }
  for (uint64_t i = 1; i < m; ++i) {
    uint32_t e = E[i], cb = C[i];
    uint64_t base = ora_mmul(B[i], M->r2, p, q);
    uint64_t pw = base;
    if (e != 1) {
      uint64_t a = base;
      while (cb) {
        a = ora_mmul(a, a, p, q);
        if (e & cb) a = ora_mmul(a, base, p, q);
        cb >>= 1;
      }
      pw = a;
    }
    s6 = ora_mmul(s6, pw, p, q);
  }
  return s6;
}

/* n! mod p by the direct product (independent check of the power table). */
uint64_t ora_factorial_mod(uint32_t n, uint64_t p) {
  uint64_t r = 1 % p;
  for (uint32_t i = 2; i <= n; ++i) r = mulmod(r, i, p);
  return r;
}

/* ----------------------------------------------------------- verification */

/* Strong probable-prime test to one base, Montgomery form (verifyprime.c:53-92). */
static int strong_prp_base(uint32_t base, const ora_mont* M, int t, uint64_t d) {
  uint64_t p = M->p, q = M->q;
  uint64_t mb = ora_mmul(base, M->r2, p, q), a = M->one, b = mb, e = d;
  while (e) {
    if (e & 1) a = ora_mmul(a, b, p, q);
    b = ora_mmul(b, b, p, q);
    e >>= 1;
  }
  if (a == M->one || a == M->nmo) return 1;
  for (int s = 1; s < t; ++s) {
    a = ora_mmul(a, a, p, q);
    if (a == M->nmo) return 1;
  }
  return 0;
}

/* Deterministic primality for 64-bit odd p: strong tests to the 7 bases
 * {2, 325, 9375, 28178, 450775, 9780504, 1795265022} with bases reduced mod p
 * and zero bases skipped (isPrime, verifyprime.c:96-132). */
int ora_is_prime(uint64_t p) {
  static const uint32_t bases[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
  if (p % 2 == 0) return 0;
  ora_mont M;
  ora_mont_init(p, &M);
  uint64_t d = p - 1;
  int t = 0;
  while ((d & 1) == 0) { d >>= 1; ++t; }
  for (int i = 0; i < 7; ++i) {
    uint32_t b = bases[i];
    if (b >= p) { b = (uint32_t)(b % p); if (b == 0) continue; }
    if (!strong_prp_base(b, &M, t, d)) return 0;
  }
  return 1;
}

/* O(n) CPU re-verification that p | n! - c... i.e. n! = +1 (c = -1) or
 * n! = -1 (c = +1) mod p, starting from 34! (verify, verifyprime.c:137-164). */
int ora_verify(uint64_t p, uint32_t n, int32_t c) {
  const u128 f34 = ((u128)0xde1bc4d19efcac82ULL << 64) | 0x445da75b00000000ULL;
  uint64_t r = (uint64_t)(f34 % p);
  for (uint32_t i = 35; i <= n; ++i) r = mulmod(r, i, p);
  if (r == 1 && c == -1) return 1;
  if (r == p - 1 && c == 1) return 1;
  return 0;
}

/* ------------------------------------------------------------------ sieve */

/* Per-candidate residue iteration (iterate.cl:54-80): for n in [nmin, nmax):
 * s7 <- s7 + one; if s6 == one -> factor (p, -n); if s6 == nmo -> (p, +n);
 * s6 <- s6 * s7.  On exit s6 = mont(nmax!), s7 = mont(nmax). */
uint32_t ora_iterate(const ora_mont* M, uint32_t nmin, uint32_t nmax, uint64_t* s6io, uint64_t* s7io,
                     ora_factor* fac, uint32_t cap) {
  uint64_t p = M->p, q = M->q, one = M->one, nmo = M->nmo, s6 = *s6io, s7 = *s7io;
  uint32_t nf = 0;
  for (uint32_t n = nmin; n < nmax; ++n) {
    s7 = ora_addmod(s7, one, p);
    if (s6 == one || s6 == nmo) {
      if (nf < cap) { fac[nf].p = p; fac[nf].nc = (s6 == one) ? -(int32_t)n : (int32_t)n; fac[nf].pad = 0; }
      ++nf;
    }
    s6 = ora_mmul(s6, s7, p, q);
  }
  *s6io = s6;
  *s7io = s7;
  return nf;
}

/* Order used by the host after draining factors: by p, then by |n|
 * (factorcompare, cl_sieve.cpp:422-436). */
static int factor_cmp(const void* a, const void* b) {
  const ora_factor* x = (const ora_factor*)a;
  const ora_factor* y = (const ora_factor*)b;
  if (x->p != y->p) return x->p < y->p ? -1 : 1;
  int32_t nx = x->nc < 0 ? -x->nc : x->nc, ny = y->nc < 0 ? -y->nc : y->nc;
  return nx < ny ? -1 : (nx > ny ? 1 : 0);
}

void ora_sort_factors(ora_factor* f, uint64_t n) { qsort(f, n, sizeof(ora_factor), factor_cmp); }

/*
 * Whole sieve over p in [pmin, pmax), n in [nmin, nmax): the device-side result
 * of the reference (candidate count, checksum fold of s6 + s7 per candidate,
 * unfiltered factor list sorted by (p, |n|)).  Host filtering is ora_finalize.
 *   check.cl:64 (s6 + s7 per candidate), check.cl:100-103 (candidate count).
 * Returns 0, or -1 if raw_cap was too small (n_raw still holds the true count).
 * Optional per-candidate outputs (ascending p): cand_p, cand_s6, cand_s7.
 */
int ora_sieve_device(uint64_t pmin, uint64_t pmax, uint32_t nmin, uint32_t nmax, int nthreads,
                     ora_factor* raw, uint64_t raw_cap, uint64_t* n_raw, uint64_t* dev_checksum,
                     uint64_t* primecount, uint64_t* cand_p, uint64_t* cand_s6, uint64_t* cand_s7,
                     uint64_t cand_cap) {
  uint64_t ncand = ora_candidates(pmin, pmax, NULL, 0);
  uint64_t* cands = (uint64_t*)malloc((ncand ? ncand : 1) * sizeof(uint64_t));
  ora_candidates(pmin, pmax, cands, ncand);
  uint64_t *B = NULL;
  uint32_t *E = NULL, *C = NULL;
  uint64_t m = ora_power_table(nmin, &B, &E, &C);
  uint64_t sum = 0, total_f = 0;
  int rc = 0;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
  {
    ora_factor local[64];
    uint64_t lsum = 0;
#pragma omp for schedule(dynamic, 16)
    for (uint64_t i = 0; i < ncand; ++i) {
      ora_mont M;
      ora_mont_init(cands[i], &M);
      uint64_t s6 = ora_setup_residue(&M, B, E, C, m);
      uint64_t s7 = ora_mmul(nmin, M.r2, M.p, M.q);
      uint32_t nf = ora_iterate(&M, nmin, nmax, &s6, &s7, local, 64);
      if (nf > 64) { /* rare: recompute with a buffer that fits */
        ora_factor* big = (ora_factor*)malloc((size_t)nf * sizeof(ora_factor));
        uint64_t a6 = ora_setup_residue(&M, B, E, C, m), a7 = ora_mmul(nmin, M.r2, M.p, M.q);
        ora_iterate(&M, nmin, nmax, &a6, &a7, big, nf);
#pragma omp critical
        {
          for (uint32_t k = 0; k < nf; ++k) {
            if (total_f < raw_cap) raw[total_f] = big[k];
            ++total_f;
          }
        }
        free(big);
      } else if (nf > 0) {
#pragma omp critical
        {
          for (uint32_t k = 0; k < nf; ++k) {
            if (total_f < raw_cap) raw[total_f] = local[k];
            ++total_f;
          }
        }
      }
      lsum += s6 + s7;
      if (i < cand_cap) {
        if (cand_p) cand_p[i] = M.p;
        if (cand_s6) cand_s6[i] = s6;
        if (cand_s7) cand_s7[i] = s7;
      }
    }
#pragma omp atomic
    sum += lsum;
  }
  if (total_f > raw_cap) rc = -1;
  else ora_sort_factors(raw, total_f);
  *n_raw = total_f;
  *dev_checksum = sum;
  *primecount = ncand;
  free(cands); free(B); free(E); free(C);
  return rc;
}

/*
 * Host-side result filtering (getResults, cl_sieve.cpp:497-550): every factor is
 * re-verified (a failure returns -1 with *bad set to its index); factors whose p
 * is not prime are discarded (2-PRP composites); kept factors add n + c to the
 * checksum.  `f` must be sorted.  Kept records are compacted to the front of f.
 * Returns the number kept.
 */
int64_t ora_finalize(ora_factor* f, uint64_t n, uint64_t* checksum, int64_t* bad) {
  uint64_t kept = 0, lastgood = 0;
  int64_t first_bad = -1;
#pragma omp parallel for schedule(dynamic, 8)
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t fn = (uint32_t)(f[i].nc < 0 ? -f[i].nc : f[i].nc);
    int32_t fc = f[i].nc < 0 ? -1 : 1;
    if (!ora_verify(f[i].p, fn, fc)) {
#pragma omp critical
      if (first_bad < 0 || (int64_t)i < first_bad) first_bad = (int64_t)i;
    }
  }
  *bad = first_bad;
  if (first_bad >= 0) return -1;
  for (uint64_t i = 0; i < n; ++i) {
    uint64_t fp = f[i].p;
    uint32_t fn = (uint32_t)(f[i].nc < 0 ? -f[i].nc : f[i].nc);
    int32_t fc = f[i].nc < 0 ? -1 : 1;
    if (fp == lastgood || ora_is_prime(fp)) {
      lastgood = fp;
      *checksum += (uint32_t)(fn + (uint32_t)fc);
      f[kept++] = f[i];
    }
  }
  return (int64_t)kept;
}

/* ------------------------------------------------------------- CLI numbers */

/* Number parsing of the -p/-P/-n/-N/-v options (parse_uint64, putil.c:33-115):
 * strtoull base 0, then an optional suffix: K/M/G/T/P (x1000^k), e<k> (x10^k),
 * k/m/g/t/p (x1024^k), b<k> (x2^k).  0 = ok, -1 = cannot parse, -2 = out of range. */
int ora_parse_uint64(uint64_t* result, const char* str, uint64_t lo, uint64_t hi) {
  uint64_t num;
  uint32_t ex = 0;
  char* tail;
  errno = 0;
  num = strtoull(str, &tail, 0);
  if (errno != 0 || num > hi) return -2;
  switch (*tail) {
    case 'P': ex += 3; /* fallthrough */
    case 'T': ex += 3; /* fallthrough */
    case 'G': ex += 3; /* fallthrough */
    case 'M': ex += 3; /* fallthrough */
    case 'K': ex += 3;
      if (tail[1] != '\0') return -1;
      for (; ex > 0; ex -= 3) {
        if (num > hi / 1000) return -2;
        num *= 1000;
      }
      break;
    case 'e': case 'E':
      ex = (uint32_t)strtoul(tail + 1, &tail, 0);
      if (errno != 0) return -2;
      if (*tail != '\0') return -1;
      while (ex-- > 0) {
        if (num > hi / 10) return -2;
        num *= 10;
      }
      break;
    case 'p': ex += 10; /* fallthrough */
    case 't': ex += 10; /* fallthrough */
    case 'g': ex += 10; /* fallthrough */
    case 'm': ex += 10; /* fallthrough */
    case 'k': ex += 10;
      if (tail[1] != '\0') return -1;
      if (num > (hi >> ex)) return -2;
      num <<= ex;
      break;
    case 'b': case 'B':
      ex = (uint32_t)strtoul(tail + 1, &tail, 0);
      if (errno != 0) return -2;
      if (*tail != '\0') return -1;
      while (ex-- > 0) {
        if (num > (hi >> 1)) return -2;
        num <<= 1;
      }
      break;
    case '\0':
      break;
    default:
      return -1;
  }
  if (num < lo) return -2;
  *result = num;
  return 0;
}
