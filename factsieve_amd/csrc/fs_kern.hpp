// fs_kern.hpp -- per-candidate device routines shared by the single-job kernels
// (fs_device.hip) and the batched / serving kernels (fs_batch.hip).
//
// Each routine works on one candidate modulus p held in registers:
//   setup:   residue <- residue * (setup program terms [start, end))
//   iterate: for n in [n0, n1): report n! = +-1, then residue <- residue * (n+1)
//   fold:    mont(n! mod p) + mont(nmax mod p) for the checksum, plus validation
// Residues enter and leave in canonical form [0, p).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "fs_arith.hpp"
#include "../../include/factsieve.h"

// Bounded factor append; `tag` goes to the record's pad word (job id in the
// batched form, 0 otherwise).
__device__ __forceinline__ void fs_append_factor(uint32_t* nf, uint32_t* ovf, fs_factor* fac, uint32_t cap,
                                                 uint64_t p, int32_t nc, int32_t tag) {
  const uint32_t idx = atomicAdd(nf, 1u);
  if (idx < cap) {
    fac[idx].p = p;
    fac[idx].nc = nc;
    fac[idx].pad = tag;
  } else {
    atomicOr(ovf, 1u);
  }
}

// Wave64 sum of a 64-bit value (result valid in every lane).
__device__ __forceinline__ uint64_t fs_wave_sum(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// ------------------------------------------------------------- candidates

#define FS_BLOCK 256

// The small-prime masks (1591 entries of 64 bits: entry off(j) + r is the mask
// of prime kSmallQ[j] for B = r mod q), built at compile time into the code
// object's constant data.
struct fs_sieve_tab_t {
  uint64_t m[FS_NMASK];
};
constexpr fs_sieve_tab_t fs_make_sieve_tab() {
  fs_sieve_tab_t t{};
  for (int j = 0; j < FS_NSMALL; ++j)
    for (uint32_t r = 0; r < kSmallQ[j]; ++r) t.m[fs_small_off(j) + r] = fs_small_mask(kSmallQ[j], r);
  return t;
}
static __constant__ const fs_sieve_tab_t c_sieve_tab = fs_make_sieve_tab();

// LDS copy of the small-prime masks (coalesced loads of the constant table).
__device__ __forceinline__ void fs_sieve_table_init(uint64_t* s_tab) {
  for (uint32_t k = threadIdx.x; k < FS_NMASK; k += FS_BLOCK) s_tab[k] = c_sieve_tab.m[k];
  __syncthreads();
}

// One 256-thread block of candidate generation over [base, high), base odd:
// thread t of block `blk` owns the 64 odd numbers base + 128*(blk*256 + t) + 2i
// as one bit mask, built by OR-ing the masks of the 29 small primes indexed by
// B mod q (the reference's mod-30 wheel plus bit tables, getsegprimes.cl:112-197).
// Each wave then deals its survivors out round-robin (exclusive prefix of
// popcounts + select-bit) so all 64 lanes run the base-2 strong probable-prime
// test (getsegprimes.cl:199-229) on consecutive survivors, keeping one pass bit
// per round.  The block then reserves room for all its passers with a single
// reserve(total) (one atomic per block, not per candidate), and each lane
// stores its passers as store(index, N, q = N^-1 mod 2^64, r2 = 2^128 mod N).
// With split > 1, `split` blocks share one range tile: each builds the same
// masks and tests only the rounds r with r % split == sub, so a short range
// still spreads its tests over many waves.  All threads must call it.
__device__ __forceinline__ uint64_t fs_primegen_survivor(uint64_t base, uint64_t tw, uint32_t wbase, uint32_t k,
                                                         const uint64_t* s_mask, const uint32_t* s_excl) {
  uint32_t lo = 0, hi = 63;  // owner lane: last L with excl[L] <= k
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (s_excl[wbase + mid] <= k) lo = mid;
    else hi = mid - 1;
  }
  uint64_t m = s_mask[wbase + lo];
  for (uint32_t j = k - s_excl[wbase + lo]; j > 0; --j) m &= m - 1;
  const uint32_t bit = __builtin_ctzll(m);
  return base + (tw + lo) * 128 + 2 * (uint64_t)bit;
}

template <class Reserve, class Store>
__device__ __forceinline__ void fs_primegen_block(uint64_t base, uint64_t high, uint64_t blk, uint32_t sub,
                                                  uint32_t split, const uint64_t* s_tab, uint64_t* s_mask,
                                                  uint32_t* s_excl, uint32_t* s_wsum, Reserve reserve, Store store) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t wbase = threadIdx.x & ~63u;
  const uint64_t t = blk * FS_BLOCK + threadIdx.x;
  const uint64_t span = high > base ? high - base : 0;
  uint64_t surv = 0;
  if (t * 128 < span) {
    const uint64_t B = base + t * 128;
    uint64_t comp = 0;
#pragma unroll
    for (int j = 0; j < FS_NSMALL; ++j) comp |= s_tab[fs_small_off(j) + (uint32_t)(B % kSmallQ[j])];
    surv = ~comp;
    const uint64_t left = (high - B + 1) >> 1;  // odd numbers B + 2i < high
    if (left < 64) surv &= (((uint64_t)1 << left) - 1);
    if (B <= 113) {  // the small primes themselves are never candidates
      for (uint32_t i = 0; i < 64; ++i)
        if (B + 2 * i <= 113) surv &= ~((uint64_t)1 << i);
    }
  }
  s_mask[threadIdx.x] = surv;
  const uint32_t cnt = __popcll(surv);
  uint32_t incl = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(incl, o, 64);
    if (lane >= (uint32_t)o) incl += v;
  }
  s_excl[threadIdx.x] = incl - cnt;
  const uint32_t total = __shfl(incl, 63, 64);
  __syncthreads();
  // 1. test: round r handles survivor k = lane + 64 r of this wave (r < 64);
  // this block takes the rounds sub, sub + split, ...
  const uint64_t tw = blk * FS_BLOCK + wbase;
  uint64_t passed = 0;
  for (uint32_t k = lane + 64 * sub, r = sub; k < total; k += 64 * split, r += split) {
    const uint64_t N = fs_primegen_survivor(base, tw, wbase, k, s_mask, s_excl);
    if (fs_candidate_probable_prime(N)) passed |= (uint64_t)1 << r;
  }
  // 2. one reservation per block; lane offsets from a block-wide prefix
  const uint32_t np = __popcll(passed);
  uint32_t pincl = np;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t v = __shfl_up(pincl, o, 64);
    if (lane >= (uint32_t)o) pincl += v;
  }
  if (lane == 63) s_wsum[threadIdx.x >> 6] = pincl;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t acc = 0;
    for (uint32_t w = 0; w < FS_BLOCK / 64; ++w) {
      const uint32_t v = s_wsum[w];
      s_wsum[w] = acc;
      acc += v;
    }
    s_wsum[FS_BLOCK / 64] = acc ? reserve(acc) : 0;
  }
  __syncthreads();
  uint32_t idx = s_wsum[FS_BLOCK / 64] + s_wsum[threadIdx.x >> 6] + (pincl - np);
  // 3. store the passers (N recomputed from its round)
  while (passed) {
    const uint32_t r = __builtin_ctzll(passed);
    passed &= passed - 1;
    const uint64_t N = fs_primegen_survivor(base, tw, wbase, lane + 64 * r, s_mask, s_excl);
    const uint64_t q = fs_invert(N), one = fs_r64_mod(N);
    store(idx++, N, q, fs_r2_from_one(one, N, q));
  }
}

// ------------------------------------------------------------------ setup

// FP64 and hybrid programs come in 32-byte blocks of four terms (multipliers);
// FS_PROG_SQUARE on a block's first term squares the running product first.
// A block is one scalar load, issued one block ahead; start and end are
// multiples of 4 (block boundaries) and the program words are 32-byte aligned.
struct alignas(32) fs_prog_block {
  uint64_t w[4];
  // the square flag (bit 63), extracted by a scalar shift of the high half:
  // the compiler otherwise tests it with a 64-bit vector compare per block
  __device__ __forceinline__ bool square() const {
    uint32_t f;
    asm("s_lshr_b32 %0, %1, 31" : "=s"(f) : "s"((uint32_t)(w[0] >> 32)));
    return f != 0;
  }
};

// FP64 program (p < 2^53): multipliers < 2^50 as double bits.
// A square is fs_fp_mulmod with the residue itself as the multiplier when every
// modulus of the wave is below FS_FP_SQUARE_PMAX (|s| < 0.8 p < 2^50, see
// fs_fp_square_ok); above it, two Montgomery products on the canonical value.
__device__ __forceinline__ uint64_t fs_cand_setup_fp64(uint64_t p, uint64_t q, uint64_t r2, uint64_t s_in,
                                                       const double* __restrict__ progd, uint32_t start,
                                                       uint32_t end) {
  const fs_prog_block* __restrict__ prog = (const fs_prog_block*)progd;
  const double pd = (double)p, pinv = 1.0 / pd;
  double s = (start == 0) ? 1.0 : fs_canon_to_fp(s_in, p);
  const bool fpsq = __all(fs_fp_square_ok(p));
  auto square = [&](const fs_prog_block& blk) {  // first use of a block: waits for its load
    if (blk.square()) {
      if (fpsq) {
        s = fs_fp_mulmod(s, s, pd, pinv);
      } else {
        const uint64_t a = fs_fp_to_canon(s, p);
        s = fs_canon_to_fp(fs_mul_mod(a, a, p, q, r2), p);
      }
    }
  };
  auto terms4 = [&](const fs_prog_block& blk) {
    s = fs_fp_mulmod(s, __longlong_as_double(blk.w[0] & ~FS_PROG_SQUARE), pd, pinv);
    s = fs_fp_mulmod(s, __longlong_as_double(blk.w[1]), pd, pinv);
    s = fs_fp_mulmod(s, __longlong_as_double(blk.w[2]), pd, pinv);
    s = fs_fp_mulmod(s, __longlong_as_double(blk.w[3]), pd, pinv);
  };
  // Scalar loads complete out of order, so any wait is for all of them: a
  // block's load is issued right after the previous block's first use (its
  // wait) and overlaps that block's four terms.  The empty asm (a memory
  // clobber that also claims to update s) pins each load there: the load
  // cannot sink below it and the terms cannot rise above it.
  const uint32_t b0 = start >> 2, b1 = end >> 2;
  if (b0 < b1) {
    fs_prog_block x = prog[b0];
    uint32_t b = b0;
    for (; b + 2 < b1; b += 2) {
      square(x);
      const fs_prog_block y = prog[b + 1];
      asm volatile("" : "+v"(s)::"memory");
      terms4(x);
      square(y);
      x = prog[b + 2];
      asm volatile("" : "+v"(s)::"memory");
      terms4(y);
    }
    square(x);
    terms4(x);
    if (b + 1 < b1) {
      const fs_prog_block y = prog[b + 1];
      square(y);
      terms4(y);
    }
  }
  return fs_fp_to_canon(s, p);
}

// Hybrid program: multipliers < 2^60 (about three primes of the power table
// per word).  Every product is a Montgomery product, so the running value is
// x * 2^-E mod p; the E of the program is removed after the last block by one
// product with Montgomery(2^E).  Below 2^64/3 a multiply is fs_redc62_mul
// (factor 2^-62) and a square fs_redc64_mul (2^-64), both lazy (value below
// 2p); between 2^64/3 and 2^63 both are fs_redc64_mul; from 2^63 on both are
// the canonical subtractive form (fs_mont_mul).  Word [terms] holds the E of
// the first range, word [terms + 1] the E of the other two.  The blocks are
// read one ahead as in the FP64 program.
template <class Mul, class Sq>
__device__ __forceinline__ uint64_t fs_setup_blocks(uint64_t s, const fs_prog_block* __restrict__ prog, uint32_t b0,
                                                    uint32_t b1, Mul mul, Sq sq) {
  auto square = [&](const fs_prog_block& blk) {
    if (blk.square()) s = sq(s);
  };
  auto terms4 = [&](const fs_prog_block& blk) {
    s = mul(s, blk.w[0] & ~FS_PROG_SQUARE);
    s = mul(s, blk.w[1]);
    s = mul(s, blk.w[2]);
    s = mul(s, blk.w[3]);
  };
  if (b0 < b1) {
    fs_prog_block x = prog[b0];
    uint32_t b = b0;
    for (; b + 2 < b1; b += 2) {
      square(x);
      const fs_prog_block y = prog[b + 1];
      asm volatile("" : "+v"(s)::"memory");
      terms4(x);
      square(y);
      x = prog[b + 2];
      asm volatile("" : "+v"(s)::"memory");
      terms4(y);
    }
    square(x);
    terms4(x);
    if (b + 1 < b1) {
      const fs_prog_block y = prog[b + 1];
      square(y);
      terms4(y);
    }
  }
  return s;
}

__device__ __forceinline__ uint64_t fs_cand_setup_hybrid(uint64_t p, uint64_t q, uint64_t r2, uint64_t s_in,
                                                         const uint64_t* __restrict__ progw, uint32_t terms,
                                                         uint32_t start, uint32_t end) {
  const fs_prog_block* __restrict__ prog = (const fs_prog_block*)progw;
  const uint32_t b0 = start >> 2, b1 = end >> 2;
  uint64_t s = (start == 0) ? 1 : s_in;
  uint64_t E;
  if (p < ((uint64_t)1 << 63)) {
    const uint64_t qn = 0 - q;
    auto sq = [&](uint64_t v) {
      const uint64_t a = v >= p ? v - p : v;  // a < p: a*a < p*2^64
      return fs_redc64_mul(a, a, p, qn);
    };
    if (p < FS_REDC62_PMAX) {
      const uint32_t pn = (uint32_t)qn;
      s = fs_setup_blocks(s, prog, b0, b1, [&](uint64_t v, uint64_t m) { return fs_redc62_mul(v, m, p, pn); }, sq);
      E = progw[terms];
    } else {
      s = fs_setup_blocks(s, prog, b0, b1, [&](uint64_t v, uint64_t m) { return fs_redc64_mul(v, m, p, qn); }, sq);
      E = progw[terms + 1];
    }
    s = s >= p ? s - p : s;
  } else {
    s = fs_setup_blocks(
        s, prog, b0, b1, [&](uint64_t v, uint64_t m) { return fs_mont_mul(v, m, p, q); },
        [&](uint64_t v) { return fs_mont_mul(v, v, p, q); });
    E = progw[terms + 1];
  }
  if (end == terms && E != 0) {
    const uint64_t one = fs_mont_mul(1, r2, p, q);
    s = fs_mont_mul(s, fs_mont_pow2(E, p, q, one), p, q);
  }
  return s;
}

// Reference power-table evaluation (setup.cl:49-103) on the Montgomery path:
// prog = [bases[0..cnt) | (exp | curbit << 32)[0..cnt)], entry 0 = the prime 2.
__device__ __forceinline__ uint64_t fs_cand_setup_mont(uint64_t p, uint64_t q, uint64_t r2, uint64_t s_in,
                                                       const uint64_t* __restrict__ prog, uint32_t cnt,
                                                       uint32_t start, uint32_t end) {
  const uint64_t one = fs_mont_mul(1, r2, p, q);
  uint64_t s6 = (start == 0) ? one : fs_mont_mul(s_in, r2, p, q);
  uint32_t k = start;
  if (k == 0 && end > 0) {  // base 2: square-and-double
    const uint64_t w = prog[cnt];
    const uint32_t e = (uint32_t)w;
    uint32_t cb = (uint32_t)(w >> 32);
    uint64_t a = fs_add_mod(one, one, p);
    while (cb) {
      a = fs_mont_mul(a, a, p, q);
      if (e & cb) a = fs_add_mod(a, a, p);
      cb >>= 1;
    }
    s6 = a;
    k = 1;
  }
  for (; k < end; ++k) {
    const uint64_t w = prog[cnt + k];
    const uint32_t e = (uint32_t)w;
    uint32_t cb = (uint32_t)(w >> 32);
    const uint64_t b = fs_mont_mul(prog[k], r2, p, q);
    uint64_t a = b;
    if (e != 1) {
      while (cb) {
        a = fs_mont_mul(a, a, p, q);
        if (e & cb) a = fs_mont_mul(a, b, p, q);
        cb >>= 1;
      }
    }
    s6 = fs_mont_mul(s6, a, p, q);
  }
  return fs_mont_mul(s6, 1, p, q);
}

// ---------------------------------------------------------------- iterate

// Unroll depths of the iterate loops: U steps share one +-1 branch and one
// scalar load of U multipliers.
#define FS_UNROLL_DEFAULT 8

// The +-1 test of a block covers two residues per compare: if v[k] = +-1 then
// the next residue v[k+1] = +-m[k] exactly (up to a multiple of p it is +-m[k];
// the symmetric step leaves it at most m[k] in magnitude), so |v[k+1]| <= m[k]
// holds whenever v[k] or v[k+1] is +-1.  It also holds for other residues of
// magnitude <= m[k] (probability about 2 m / p per test).  The tests' wave masks
// (ballots: each compare writes a scalar register pair) are or-ed on the scalar
// unit, and a block whose mask is not zero takes a uniform branch in which
// every lane tests every residue of the block exactly.

// FP64 path.  nt[j] = (double)(n0 + 1 + j), read with scalar loads, with at
// least U entries past n1 - n0; six FP64 operations per step (four when
// p * n1 <= 2^53 and the product s*m is exact: EXACT) and one compare per two
// steps; U steps (2, 4 or 8) share one branch.
template <int U, bool EXACT, class Rec>
__device__ __forceinline__ uint64_t fs_cand_iter_fp64_loop(uint64_t p, uint64_t s_in, uint32_t n0, uint32_t n1,
                                                           const double* __restrict__ nt, Rec rec) {
  static_assert(U == 2 || U == 4 || U == 8, "unroll depth");
  const double pd = (double)p, pinv = 1.0 / pd;
  double s = fs_canon_to_fp(s_in, p);
  uint32_t n = n0;
  const uint32_t nu = n0 + ((n1 - n0) & ~(uint32_t)(U - 1));
  double m[U];
#pragma unroll
  for (int k = 0; k < U; ++k) m[k] = nt[k];
  for (; n < nu; n += U) {
    // next block, loaded one iteration ahead (left to the compiler's placement:
    // pinning it as in the hybrid loop measured slower here)
    const double* mn = nt + (n - n0) + U;
    double x[U], v[U + 1];
#pragma unroll
    for (int k = 0; k < U; ++k) x[k] = mn[k];
    v[0] = s;
#pragma unroll
    for (int k = 0; k < U; ++k)
      v[k + 1] = EXACT ? fs_fp_mulmod_exact(v[k], m[k], pd, pinv) : fs_fp_mulmod(v[k], m[k], pd, pinv);
    s = v[U];
    uint64_t hit = 0;  // wave masks of the tests, or-ed on the scalar unit
#pragma unroll
    for (int k = 0; k < U; k += 2) hit |= __ballot(__builtin_fabs(v[k + 1]) <= m[k]);
#pragma unroll
    for (int k = 0; k < U; ++k) m[k] = x[k];
    if (hit) {
#pragma unroll
      for (int k = 0; k < U; ++k)
        if (__builtin_fabs(v[k]) == 1.0) rec(v[k] > 0.0 ? -(int32_t)(n + k) : (int32_t)(n + k));
    }
  }
  for (; n < n1; ++n) {
    if (__builtin_fabs(s) == 1.0) rec(s > 0.0 ? -(int32_t)n : (int32_t)n);
    s = EXACT ? fs_fp_mulmod_exact(s, nt[n - n0], pd, pinv) : fs_fp_mulmod(s, nt[n - n0], pd, pinv);
  }
  return fs_fp_to_canon(s, p);
}

// FP64 path, paired steps (fs_fp_pair_ok(p, n1)): one product by
// M = a*(a+1) per two multipliers.  pt[2j] = (n0+1+2j)*(n0+2+2j) and
// pt[2j+1] = n0+2+2j, read with scalar loads, with at least U entries (U/2
// pairs) past index n1 - n0.  With v[i] the residue of nn = nb + 2i and a+1 =
// nn+2 the pair's second multiplier, v[i+1] = v[i] * M = (residue of nn+1) *
// (a+1); a+1 being below 0.4844 p and every residue of magnitude <= 0.503 p,
// v[i+1] == +-(a+1) exactly when the residue of nn+1 is +-1, and |v[i+1]| == 1
// exactly when the residue of nn+2 is.  So a block of H pairs from position
// nb owns positions nb+1 .. nb+U, and one compare per pair, |v[i+1]| <= nb+U
// (the block's last a+1, a scalar), fires for every +-1 among them (and, with
// probability about 2U n/p, for other small residues).  The taken branch
// recomputes the odd positions with single steps and records every +-1 below
// n1 in order.  Position n0 is tested once at the start; the positions after
// the last whole block belong to the single-step tail.  Per two steps: one
// six-operation product (four when EXACT) and one compare.
template <int U, bool EXACT, class Rec>
__device__ __forceinline__ uint64_t fs_cand_iter_fp64_pair_loop(uint64_t p, uint64_t s_in, uint32_t n0, uint32_t n1,
                                                                const double* __restrict__ nt,
                                                                const double* __restrict__ pt, Rec rec) {
  static_assert(U == 2 || U == 4 || U == 8 || U == 16, "unroll depth");
  constexpr int H = U / 2;
  const double pd = (double)p, pinv = 1.0 / pd;
  double s = fs_canon_to_fp(s_in, p);
  if (n0 < n1 && __builtin_fabs(s) == 1.0) rec(s > 0.0 ? -(int32_t)n0 : (int32_t)n0);
  uint32_t n = n0;
  const uint32_t nu = n0 + ((n1 - n0) & ~(uint32_t)(U - 1));
  // One block of H paired steps from position nb with the pairs in w, and its
  // +-1 records.
  auto block = [&](const double (&w)[U], uint32_t nb) {
    double v[H + 1];
    v[0] = s;
#pragma unroll
    for (int i = 0; i < H; ++i)
      v[i + 1] = EXACT ? fs_fp_mulmod_exact(v[i], w[2 * i], pd, pinv) : fs_fp_mulmod(v[i], w[2 * i], pd, pinv);
    s = v[H];
    // wave masks of the compares (each one v_cmp into a scalar register
    // pair), or-ed on the scalar unit; the exact tests below run for the whole
    // wave when any lane's compare fired
    uint64_t hit = 0;
#pragma unroll
    for (int i = 0; i < H; ++i) hit |= __ballot(__builtin_fabs(v[i + 1]) <= w[U - 1]);
    if (hit) {
#pragma unroll
      for (int i = 0; i < H; ++i) {
        const uint32_t nn = nb + 2 * i;
        const double u = fs_fp_mulmod(v[i], nt[nn - n0], pd, pinv);  // residue of nn + 1
        if (nn + 1 < n1 && __builtin_fabs(u) == 1.0) rec(u > 0.0 ? -(int32_t)(nn + 1) : (int32_t)(nn + 1));
        if (nn + 2 < n1 && __builtin_fabs(v[i + 1]) == 1.0)
          rec(v[i + 1] > 0.0 ? -(int32_t)(nn + 2) : (int32_t)(nn + 2));
      }
    }
  };
  // Two register sets for the pairs, alternating by block.  Scalar loads
  // complete out of order (a wait is for all of them), so each block first
  // takes its own set (the empty asm that reads it waits for its load), then
  // issues the next block's load into the other set, pinned above this block's
  // steps by the asm that claims to update s; no set is copied.
  double wa[U], wb[U];
#pragma unroll
  for (int k = 0; k < U; ++k) wa[k] = pt[k];
  while (n < nu) {
#pragma unroll
    for (int k = 0; k < U; ++k) asm volatile("" ::"s"(wa[k]));
#pragma unroll
    for (int k = 0; k < U; ++k) wb[k] = pt[(n - n0) + U + k];
    asm volatile("" : "+v"(s)::"memory");
    block(wa, n);
    n += U;
    if (n >= nu) break;
#pragma unroll
    for (int k = 0; k < U; ++k) asm volatile("" ::"s"(wb[k]));
#pragma unroll
    for (int k = 0; k < U; ++k) wa[k] = pt[(n - n0) + U + k];
    asm volatile("" : "+v"(s)::"memory");
    block(wb, n);
    n += U;
  }
  for (; n < n1; ++n) {  // positions nu + 1 .. n1 - 1
    s = EXACT ? fs_fp_mulmod_exact(s, nt[n - n0], pd, pinv) : fs_fp_mulmod(s, nt[n - n0], pd, pinv);
    if (n + 1 < n1 && __builtin_fabs(s) == 1.0) rec(s > 0.0 ? -(int32_t)(n + 1) : (int32_t)(n + 1));
  }
  return fs_fp_to_canon(s, p);
}

// Per wave: paired steps when the pair table covers the launch (pairs) and
// every active lane's modulus allows them, else single steps; within each, the
// exact-product step when every active lane's modulus allows it.  pt is always
// a valid table pointer (never selected against a null pointer), which keeps
// its loads scalar.
template <int U = FS_UNROLL_DEFAULT, int UP = U, class Rec>
__device__ __forceinline__ uint64_t fs_cand_iter_fp64(uint64_t p, uint64_t s_in, uint32_t n0, uint32_t n1,
                                                      const double* __restrict__ nt, const double* __restrict__ pt,
                                                      bool pairs, Rec rec) {
  if (pairs && __all(fs_fp_pair_ok(p, n1))) {
    if (__all(fs_fp_pair_exact_ok(p, n1))) return fs_cand_iter_fp64_pair_loop<UP, true>(p, s_in, n0, n1, nt, pt, rec);
    return fs_cand_iter_fp64_pair_loop<UP, false>(p, s_in, n0, n1, nt, pt, rec);
  }
  if (__all(fs_fp_exact_ok(p, n1))) return fs_cand_iter_fp64_loop<U, true>(p, s_in, n0, n1, nt, rec);
  return fs_cand_iter_fp64_loop<U, false>(p, s_in, n0, n1, nt, rec);
}

// Hybrid path (p <= FS_HYBRID_PMAX): signed residue |s| <= 0.51 p (0.76 p
// with the high-word quotient), exact 64x32 products mod 2^64 with an FP64
// quotient; nt as for the FP64 path.
// n! = +-1 (mod p) iff s == +-1.  U steps share one test on the quotient
// estimates y[k] ~ v[k]/p that the steps compute anyway, taken at the odd
// positions (two residues each, as above): |y| <= 1.5 * m[U-2] / p holds for
// every |v| <= m[U-2] (the estimate's error is below 2^-50 |v|/p + 2^-21/p),
// and the branch then tests every residue exactly.  HI: the quotient from the
// high word of s (fs_sh_mulmod_hi, when fs_sh_hi_ok(p, n1)), whose estimates pass
// |y| <= 1.5 * 2^32/p for every |v| < 2^33, so for every |v| <= m < 2^31.
template <int U, bool HI, class Rec>
__device__ __forceinline__ uint64_t fs_cand_iter_hybrid_loop(uint64_t p, uint64_t s_in, uint32_t n0, uint32_t n1,
                                                             const double* __restrict__ nt, Rec rec) {
  static_assert(U == 2 || U == 4 || U == 8, "unroll depth");
  const double pinv = 1.0 / (double)p, pinv32 = pinv * 4294967296.0, thr_hi = 1.5 * pinv32, thr15 = 1.5 * pinv;
  const uint64_t pn = 0 - p, c31 = p << 31;
  int64_t s = fs_canon_to_signed(s_in, p);
  uint32_t n = n0;
  const uint32_t nu = n0 + ((n1 - n0) & ~(uint32_t)(U - 1));
  double m[U];
#pragma unroll
  for (int k = 0; k < U; ++k) m[k] = nt[k];
  for (; n < nu; n += U) {
    // Scalar loads complete out of order (a wait is for all of them): take
    // this block's multipliers first, then issue the next block's load, pinned
    // above this block's steps by the asm that claims to update s.
#pragma unroll
    for (int k = 0; k < U; ++k) asm volatile("" ::"s"(m[k]));
    const double* mn = nt + (n - n0) + U;
    double x[U], y[U];
    int64_t v[U + 1];
#pragma unroll
    for (int k = 0; k < U; ++k) x[k] = mn[k];
    asm volatile("" : "+v"(s)::"memory");
    uint32_t mm = n + 1;
    asm volatile("" : "+s"(mm));  // keep the multiplier a 32-bit scalar (no 64-bit induction variable)
    v[0] = s;
#pragma unroll
    for (int k = 0; k < U; ++k)
      v[k + 1] = HI ? fs_sh_mulmod_hi(v[k], mm + k, m[k], c31, pn, pinv32, &y[k])
                    : fs_sh_mulmod(v[k], mm + k, m[k], c31, pn, pinv, pinv32, &y[k]);
    s = v[U];
    double ymin = __builtin_fabs(y[1]);
#pragma unroll
    for (int k = 3; k < U; k += 2) ymin = __builtin_fmin(ymin, __builtin_fabs(y[k]));
    const double thr = HI ? thr_hi : thr15 * m[U - 2];  // m[U-2]: the block's largest even-position multiplier
#pragma unroll
    for (int k = 0; k < U; ++k) m[k] = x[k];
    if (__ballot(ymin <= thr)) {  // whole wave: exact tests where any lane's filter passed
#pragma unroll
      for (int k = 0; k < U; ++k)
        if ((v[k] == 1) | (v[k] == -1)) rec(v[k] == 1 ? -(int32_t)(n + k) : (int32_t)(n + k));
    }
  }
  for (; n < n1; ++n) {
    if ((s == 1) | (s == -1)) rec(s == 1 ? -(int32_t)n : (int32_t)n);
    uint32_t mm = n + 1;
    asm volatile("" : "+s"(mm));
    double y;
    s = HI ? fs_sh_mulmod_hi(s, mm, nt[n - n0], c31, pn, pinv32, &y)
           : fs_sh_mulmod(s, mm, nt[n - n0], c31, pn, pinv, pinv32, &y);
  }
  return fs_signed_to_canon(s, p);
}

// Hybrid path, paired steps (fs_sh_pair_ok(p, n1)): one product by
// M = a*(a+1) per two multipliers (fs_sh_mulmod_pair), the words of M formed
// on the scalar unit from the uniform a, its double read from the FP64 pair
// table pt (as for the FP64 path, pt[2j] = (n0+1+2j)*(n0+2+2j)), with the
// same double-buffered scalar loads.  Each step leaves the estimate y of its
// result over p, which the next step needs for its quotient.  As on the FP64
// path a block of H pairs from nb owns positions nb+1 .. nb+U: its results
// are +-1 or +-(a+1) whenever one of those residues is +-1, and its filter
// |y| <= 1.5 (n1 + 2)/p over the H result estimates then holds (by the
// estimate's error bound).  The taken branch re-tests the block's positions
// below n1 exactly with single steps; position n0 is tested once at the start
// and the positions after the last whole block by the single-step tail.  Per
// two steps: thirteen operations and one of the filter, plus one per block.
template <int U, class Rec>
__device__ __forceinline__ uint64_t fs_cand_iter_hybrid_pair_loop(uint64_t p, uint64_t s_in, uint32_t n0,
                                                                  uint32_t n1, const double* __restrict__ nt,
                                                                  const double* __restrict__ pt, Rec rec) {
  static_assert(U == 2 || U == 4 || U == 8 || U == 16, "unroll depth");
  constexpr int H = U / 2;
  const double pinv = 1.0 / (double)p, pinv32 = pinv * 4294967296.0;
  const double thr = 1.5 * (double)(n1 + 2) * pinv;
  const uint64_t pn = 0 - p, c31 = p << 31, cpair = fs_sh_pair_c(pn);
  int64_t s = fs_canon_to_signed(s_in, p);
  if (n0 < n1 && ((s == 1) | (s == -1))) rec(s == 1 ? -(int32_t)n0 : (int32_t)n0);
  double ys = fs_sh_estimate(s, pinv, pinv32);
  uint32_t n = n0;
  const uint32_t nu = n0 + ((n1 - n0) & ~(uint32_t)(U - 1));
  // The block keeps only its input and the running minimum of |y| while it
  // steps; the rare taken branch recomputes the block's residues from its
  // input with single steps (fs_sh_mulmod: other representatives of the same
  // residues, within 0.51 p, so the same +-1 tests), which keeps the loop's
  // registers low.
  auto block = [&](const double (&w)[U], uint32_t nb) {
    const int64_t sb = s;
    uint32_t a = nb + 1;
    asm volatile("" : "+s"(a));  // a uniform 32-bit scalar: the words of M on the scalar unit
    double ymin = 1.0;
#pragma unroll
    for (int i = 0; i < H; ++i) {
      const uint32_t ai = a + 2 * i;
      const uint64_t M = (uint64_t)ai * (ai + 1);
      s = fs_sh_mulmod_pair(s, ys, (uint32_t)M, (uint32_t)(M >> 32), w[2 * i], pn, cpair, pinv, pinv32, &ys);
      ymin = i == 0 ? __builtin_fabs(ys) : __builtin_fmin(ymin, __builtin_fabs(ys));
    }
    if (__ballot(ymin <= thr)) {  // whole wave: exact tests where any lane's filter passed
      int64_t v = sb;
#pragma unroll 1
      for (uint32_t nn = nb; nn < nb + U; ++nn) {
        double yu;
        v = fs_sh_mulmod(v, nn + 1, nt[nn - n0], c31, pn, pinv, pinv32, &yu);  // residue of nn + 1
        if (nn + 1 < n1 && ((v == 1) | (v == -1))) rec(v == 1 ? -(int32_t)(nn + 1) : (int32_t)(nn + 1));
      }
    }
  };
  double wa[U], wb[U];
#pragma unroll
  for (int k = 0; k < U; ++k) wa[k] = pt[k];
  while (n < nu) {
#pragma unroll
    for (int k = 0; k < U; ++k) asm volatile("" ::"s"(wa[k]));
#pragma unroll
    for (int k = 0; k < U; ++k) wb[k] = pt[(n - n0) + U + k];
    asm volatile("" : "+v"(s)::"memory");
    block(wa, n);
    n += U;
    if (n >= nu) break;
#pragma unroll
    for (int k = 0; k < U; ++k) asm volatile("" ::"s"(wb[k]));
#pragma unroll
    for (int k = 0; k < U; ++k) wa[k] = pt[(n - n0) + U + k];
    asm volatile("" : "+v"(s)::"memory");
    block(wb, n);
    n += U;
  }
  for (; n < n1; ++n) {  // positions nu + 1 .. n1 - 1
    double y;
    s = fs_sh_mulmod(s, n + 1, nt[n - n0], c31, pn, pinv, pinv32, &y);
    if (n + 1 < n1 && ((s == 1) | (s == -1))) rec(s == 1 ? -(int32_t)(n + 1) : (int32_t)(n + 1));
  }
  return fs_signed_to_canon(s, p);
}

// Per wave: paired steps when the pair table covers the launch (pairs) and
// every active lane's modulus allows them, else the high-word quotient when
// every active lane's modulus allows it, else the full quotient.  pt is always a valid table
// pointer (never selected against a null pointer), which keeps its loads
// scalar.
template <int U = FS_UNROLL_DEFAULT, int UP = U, class Rec>
__device__ __forceinline__ uint64_t fs_cand_iter_hybrid(uint64_t p, uint64_t s_in, uint32_t n0, uint32_t n1,
                                                        const double* __restrict__ nt, const double* __restrict__ pt,
                                                        bool pairs, Rec rec) {
  if (pairs && __all(fs_sh_pair_ok(p, n1))) return fs_cand_iter_hybrid_pair_loop<UP>(p, s_in, n0, n1, nt, pt, rec);
  if (__all(fs_sh_hi_ok(p, n1))) return fs_cand_iter_hybrid_loop<U, true>(p, s_in, n0, n1, nt, rec);
  return fs_cand_iter_hybrid_loop<U, false>(p, s_in, n0, n1, nt, rec);
}

// Reference formulation (iterate.cl:54-80): Montgomery residue s6 and
// Montgomery counter s7 advanced by one each step.  Any odd p < 2^64.
template <class Rec>
__device__ __forceinline__ uint64_t fs_cand_iter_mont(uint64_t p, uint64_t q, uint64_t r2, uint64_t s_in,
                                                      uint32_t n0, uint32_t n1, Rec rec) {
  const uint64_t one = fs_mont_mul(1, r2, p, q), nmo = p - one;
  uint64_t s6 = fs_mont_mul(s_in, r2, p, q);
  uint64_t s7 = fs_mont_mul(n0 < p ? (uint64_t)n0 : (uint64_t)n0 % p, r2, p, q);
  for (uint32_t n = n0; n < n1; ++n) {
    s7 = fs_add_mod(s7, one, p);
    if (s6 == one || s6 == nmo) rec((s6 == one) ? -(int32_t)n : (int32_t)n);
    s6 = fs_mont_mul(s6, s7, p, q);
  }
  return fs_mont_mul(s6, 1, p, q);
}

// ------------------------------------------------------------------- fold

// mont(s) + mont(nmax mod p) for the checksum (check.cl:60-64); ok is false if
// the residue is not canonical or the Montgomery round trip fails
// (check.cl:66-78 validates the same way with REDC(s7) == nmax).
__device__ __forceinline__ uint64_t fs_cand_fold(uint64_t p, uint64_t q, uint64_t r2, uint64_t s, uint32_t nmax,
                                                 bool& ok) {
  const uint64_t s6 = fs_mont_mul(s, r2, p, q);
  const uint64_t nm = (p > nmax) ? (uint64_t)nmax : (uint64_t)nmax % p;
  const uint64_t s7 = fs_mont_mul(nm, r2, p, q);
  ok = (s < p) && (fs_mont_mul(s6, 1, p, q) == s) && (fs_mont_mul(s7, 1, p, q) == nm);
  return s6 + s7;
}
