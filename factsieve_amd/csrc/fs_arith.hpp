// fs_arith.hpp -- exact modular arithmetic for the sieve kernels (gfx950).
//
// Three formulations of "r <- r * m mod p" are provided; all return exact
// residues, so every kernel built on them produces the same factors and
// checksum (the reference's results do not depend on how a residue is
// computed, only on its value: cl_sieve.cpp / check.cl fold canonical values).
//
//  * Montgomery REDC, R = 2^64, q = p^-1 mod 2^64, subtractive form with a
//    canonical result -- the reference formulation (iterate.cl:19-47).  Valid for
//    every odd p < 2^64.  64x64->128 products are built from 32x32->64
//    v_mad_u64_u32 partial products.
//
//  * FP64 error-free product with a rounded quotient (p < 2^53): the residue is
//    a double s with |s| < 0.8 p; for an integer multiplier m < 2^50,
//      h = fl(s*m), l = fma(s, m, -h)        (h + l == s*m exactly, |l| <= 2^49)
//      k = rint(fl(h * (1/p)))               (three roundings: |fl(..) - s*m/p|
//                                             < 3*2^-53 * 0.8*2^50 = 0.3, so
//                                             |k - s*m/p| < 0.8)
//      s' = fma(-k, p, h) + l                (exact: h - k*p is an integer of
//                                             magnitude < 0.8 p + 2^49 < 2^53)
//    gives s' == s*m - k*p exactly with |s'| < 0.8 p (iterate's m < 2^31 gives
//    |s'| <= 0.5 p + 2^-21 p).  Every integer below 2^53 (p, residues, n) is an
//    exact double.  Because |s| < p - 1 (p > 5),
//    s == +-1 (mod p) iff |s| == 1.  Six full-rate FP64 VALU operations (four
//    when s*m is exact, fs_fp_mulmod_exact); the iterate loops multiply by
//    (n+1)(n+2) at once when n1 <= 2^22 and 64 n1 <= 31 p (fs_fp_pair_ok).
//
//  * Hybrid (p <= 2^64 - 2^46, used above 2^53): signed residue |s| <= 0.51p (and
//    below 2^63: the step keeps |s| <= (1/2 + 2^-19) p), multiplier
//    m < 2^31; the quotient is an FP64 estimate rounded to nearest, and
//    s' = s*m - k*p is formed exactly mod 2^64 from 32x32->64 multiply-adds.
//    The setup program multiplies by 60-bit words with Montgomery reductions
//    (two REDC-by-2^31 stages below 2^64/3, one REDC by 2^64 above).
//
// The error bounds are proved in the design notes shipped with the library and
// are exercised by the parity tests at the edges of each range.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define FS_HD __host__ __device__ __forceinline__

// Setup program word flag (FP64 and hybrid paths): square the running product
// before multiplying by this term; set only on the first term of a 4-term block.
#define FS_PROG_SQUARE 0x8000000000000000ull

// ------------------------------------------------------------------ Montgomery

// p^-1 mod 2^64 by Newton iteration (each step doubles the correct low bits).
FS_HD uint64_t fs_invert(uint64_t p) {
  uint64_t x = p;                 // p*p == 1 mod 8 for odd p: 3 correct bits
  for (int i = 0; i < 5; ++i) x *= 2 - p * x;  // 3 -> 6 -> 12 -> 24 -> 48 -> 96
  return x;
}

// 64x64 -> 128 product from four 32x32->64 partial products.
FS_HD void fs_mul_wide(uint64_t a, uint64_t b, uint64_t& lo, uint64_t& hi) {
  const uint32_t a0 = (uint32_t)a, a1 = (uint32_t)(a >> 32);
  const uint32_t b0 = (uint32_t)b, b1 = (uint32_t)(b >> 32);
  const uint64_t p00 = (uint64_t)a0 * b0;
  const uint64_t p01 = (uint64_t)a0 * b1 + (p00 >> 32);
  const uint64_t p10 = (uint64_t)a1 * b0 + (uint32_t)p01;
  hi = (uint64_t)a1 * b1 + (p01 >> 32) + (p10 >> 32);
  lo = (p10 << 32) | (uint32_t)p00;
}

// High 64 bits of a 64x64 product.
FS_HD uint64_t fs_mul_hi(uint64_t a, uint64_t b) {
  uint64_t lo, hi;
  fs_mul_wide(a, b, lo, hi);
  return hi;
}

// a*b*2^-64 mod p, canonical in [0, p) when a*b < p*2^64.
FS_HD uint64_t fs_mont_mul(uint64_t a, uint64_t b, uint64_t p, uint64_t q) {
  uint64_t lo, hi;
  fs_mul_wide(a, b, lo, hi);
  const uint64_t m = lo * q;
  const uint64_t mp = fs_mul_hi(m, p);
  const uint64_t r = hi - mp;
  return (hi < mp) ? r + p : r;
}

// (a + b) mod p for a, b in [0, p), overflow-safe.
FS_HD uint64_t fs_add_mod(uint64_t a, uint64_t b, uint64_t p) {
  const uint64_t c = (a >= p - b) ? p : 0;
  return a + b - c;
}

// Plain-domain product a*b mod p (a, b canonical) via two Montgomery products.
FS_HD uint64_t fs_mul_mod(uint64_t a, uint64_t b, uint64_t p, uint64_t q, uint64_t r2) {
  return fs_mont_mul(fs_mont_mul(a, b, p, q), r2, p, q);
}

// 2^128 mod p from one = 2^64 mod p: (4*one)^(2^5) in Montgomery form.
FS_HD uint64_t fs_r2_from_one(uint64_t one, uint64_t p, uint64_t q) {
  const uint64_t two = fs_add_mod(one, one, p);
  uint64_t r2 = fs_add_mod(two, two, p);
  for (int i = 0; i < 5; ++i) r2 = fs_mont_mul(r2, r2, p, q);
  return r2;
}

// ------------------------------------------------------------------ FP64 path

// s*m mod p for |s| < 0.8p, integer |m| < 2^50, odd 5 < p < 2^53 (see header).
FS_HD double fs_fp_mulmod(double s, double m, double p, double pinv) {
#pragma clang fp contract(off)
  const double h = s * m;
  const double l = __builtin_fma(s, m, -h);
  const double k = __builtin_rint(h * pinv);
  const double r = __builtin_fma(-k, p, h);
  return r + l;
}

// The FP64 step when the product is exact: |s| <= 0.51p (the iterate loops'
// invariant: they start from the symmetric residue, |s| <= p/2, and every step
// leaves |s'| <= (1/2 + 2^-21) p), integer m, and p * m <= 2^53
// (fs_fp_exact_ok), so |s*m| <= 0.51 * 2^53 and h = s*m needs no error term.
// k = rint(h * fl(1/p)) is within 1/2 + 2^-51 X of X = s*m/p, and
// fma(-k, p, h) = s*m - k*p exactly (an integer of magnitude < 0.51 p).  Three
// FP64 operations and the round instead of six.
FS_HD bool fs_fp_exact_ok(uint64_t p, uint32_t mmax) {
  return (unsigned __int128)p * mmax <= ((unsigned __int128)1 << 53);
}
FS_HD double fs_fp_mulmod_exact(double s, double m, double p, double pinv) {
#pragma clang fp contract(off)
  const double h = s * m;
  const double k = __builtin_rint(h * pinv);
  return __builtin_fma(-k, p, h);
}

// Paired FP64 step: one product by M = a*(a+1) advances the residue over two
// consecutive multipliers a, a+1 <= n1.  fs_fp_pair_ok(p, n1): n1 <= 2^22 and
// 64 n1 <= 31 p.  Then M < n1^2 <= 2^44, so k = rint(fl(h * fl(1/p))) is
// within 1/2 + 3 * 2^-53 * 0.51 * 2^44 < 0.503 of X = s*M/p and every step keeps
// |s'| <= 0.503 p (fs_fp_mulmod, whose bounds need only M < 2^50 and |s| < 0.8p);
// M itself may exceed p.  And a+1 <= n1 <= (31/64) p < 0.4844 p, so of the
// values congruent to +-(a+1) only +-(a+1) itself has magnitude <= 0.503 p
// (|a+1 - p| > 0.5156 p): a paired step's result congruent to +-(a+1), which it
// is whenever the skipped residue (the one of a, times a+1) is +-1, is exactly
// +-(a+1).  The residues the loop keeps are tested for +-1 directly.
// fs_fp_pair_exact_ok adds p * n1^2 <= 2^53: |s*M| < 0.51 * 2^53 and
// fs_fp_mulmod_exact applies.
FS_HD bool fs_fp_pair_ok(uint64_t p, uint32_t n1) {
  return n1 <= (1u << 22) && 64 * (uint64_t)n1 <= 31 * p;
}
FS_HD bool fs_fp_pair_exact_ok(uint64_t p, uint32_t n1) {
  return fs_fp_pair_ok(p, n1) && (unsigned __int128)p * ((uint64_t)n1 * n1) <= ((unsigned __int128)1 << 53);
}

// FP64 square: fs_fp_mulmod(s, s) for |s| < 0.8 p needs the multiplier s below
// 2^50, which p < 2^50 gives.  The quotient estimate then misses X = s^2/p
// (< 0.64 p < 0.64 * 2^50) by less than 3 * 2^-53 * 0.64 * 2^50 = 0.24, so the
// result keeps |s'| < 0.74 p, and h - k p stays below 0.74 p + 2^47 < 2^53 (exact).
#define FS_FP_SQUARE_PMAX ((uint64_t)1 << 50)
FS_HD bool fs_fp_square_ok(uint64_t p) { return p < FS_FP_SQUARE_PMAX; }

// Canonical [0, p) integer from a symmetric FP64 residue.
FS_HD uint64_t fs_fp_to_canon(double s, uint64_t p) {
  const int64_t v = (int64_t)s;
  return v < 0 ? (uint64_t)(v + (int64_t)p) : (uint64_t)v;
}

// Symmetric FP64 residue (|s| <= p/2) from a canonical value.
FS_HD double fs_canon_to_fp(uint64_t r, uint64_t p) {
  return (r > (p >> 1)) ? -(double)(p - r) : (double)r;
}

// ----------------------------------------------------------------- hybrid path

// Largest modulus of the hybrid path (exclusive bound on pmax is this + 1).
#define FS_HYBRID_PMAX (0xFFFFFFFFFFFFFFFFull - ((uint64_t)1 << 46) + 1)

// Signed hybrid step: s*m mod p for an int64 residue |s| <= 0.51 p, 0 < m < 2^31
// and m < 2p (iterate multiplies by n + 1 <= p whenever s != 0, since p | n! for
// n >= p), 3 <= p <= FS_HYBRID_PMAX; c31 = 2^31 p mod 2^64, pn = 2^64 - p,
// pinv = fl(1/p), pinv32 = 2^32 * pinv, md = (double)m.
// y = s/p + err with |err| <= 2^-51.9 |s/p| + 2^-21/p (the second term from
// rounding s0 * pinv, s0 < 2^32), so y*m misses X = s*m/p by at most
// m*2^-51.9 + (m/p)*2^-21 + 2^-22 (final rounding) < 2^-19 for m < 2p, and
// k = trunc(y*m + 2^31 + 1/2) is round(X) + 2^31: r = s*m - (k - 2^31) p has
// |r| <= (1/2 + 2^-19) p < 2^63 and is formed exactly mod 2^64 as
// s*m + c31 + k*pn from four 32x32->64 multiply-adds.  Because |r| < p - 1,
// r == +-1 (mod p) iff r == +-1.
// *yq receives y, which the iterate loops use as a filter for s == +-1: by the
// error bound above, |y| <= 1.5 pinv holds for s in {-1, 0, 1} and fails for
// every |s| >= 2.
FS_HD int64_t fs_sh_mulmod(int64_t s, uint32_t m, double md, uint64_t c31, uint64_t pn, double pinv,
                           double pinv32, double* yq) {
#pragma clang fp contract(off)
  const uint32_t s0 = (uint32_t)s;
  const int32_t s1 = (int32_t)(s >> 32);
  const double y = __builtin_fma((double)s1, pinv32, (double)s0 * pinv);
  *yq = y;
  const uint32_t k = (uint32_t)__builtin_fma(y, md, 2147483648.5);
  const uint64_t a = (uint64_t)s0 * m + c31;
  const uint64_t b = (uint64_t)k * (uint32_t)pn + a;
  const uint32_t hi = (uint32_t)((uint64_t)(uint32_t)s1 * m + (b >> 32)) + k * (uint32_t)(pn >> 32);
  return (int64_t)(((uint64_t)hi << 32) | (uint32_t)b);
}

FS_HD int64_t fs_sh_mulmod(int64_t s, uint32_t m, double md, uint64_t c31, uint64_t pn, double pinv,
                           double pinv32) {
  double y;
  return fs_sh_mulmod(s, m, md, c31, pn, pinv, pinv32, &y);
}

// The signed step with the quotient from the high word of s only, for moduli
// far above the multiplier (fs_sh_hi_ok).  y = s1 * 2^32/p misses s/p by
// s0/p in [0, 2^32/p), so y*m falls short of X = s*m/p by at most
// 2^32 m / p, and the rounding terms of fs_sh_mulmod add less than 2^-21
// (m < 2^29): r = s*m - (k - 2^31) p has |r| <= (1/2 + 2^32 m/p + 2^-21) p.
// Two regimes keep r an exact int64 with |r| < p - 1, whatever the input
// |s| <= 0.76 p:
//   p < 2^63 and p >= m * 2^34:  |r| <= (3/4 + 2^-21) p < 2^63;
//   p < FS_SH_HI_PMAX and p >= m * 2^39:  |r| <= (1/2 + 2^-7 + 2^-21) p.
// Two FP64 operations fewer than fs_sh_mulmod (no conversion of s0, no FMA).
// *yq = y is a filter for s == +-1: s in {-1, 0, 1} has s1 in {-1, 0}, so
// |y| <= 1.5 * pinv32 holds for it (and also for other |s| < 2^33, which the
// exact test then rejects).  A residue +-1 times m <= p/2^34 gets k = 2^31
// (|y*m| <= 1/4), so the step leaves exactly +-m, as the paired test needs.
#define FS_SH_HI_SHIFT 39
#define FS_SH_HI_SHIFT_LOW 34
#define FS_SH_HI_PMAX (0xFFFFFFFFFFFFFFFFull - ((uint64_t)1 << 58) + 1)
// Largest multiplier the high-word step takes for p (0 when it takes none).
FS_HD uint64_t fs_sh_hi_mmax(uint64_t p) {
  if (p >= FS_SH_HI_PMAX) return 0;
  return p < ((uint64_t)1 << 63) ? p >> FS_SH_HI_SHIFT_LOW : p >> FS_SH_HI_SHIFT;
}
// Bound on |r| / p of the high-word step, and so on the residues of its loop.
#define FS_SH_HI_RBOUND 0.7500005
FS_HD bool fs_sh_hi_ok(uint64_t p, uint32_t mmax) { return fs_sh_hi_mmax(p) >= (uint64_t)mmax; }

FS_HD int64_t fs_sh_mulmod_hi(int64_t s, uint32_t m, double md, uint64_t c31, uint64_t pn, double pinv32,
                              double* yq) {
#pragma clang fp contract(off)
  const uint32_t s0 = (uint32_t)s;
  const int32_t s1 = (int32_t)(s >> 32);
  const double y = (double)s1 * pinv32;
  *yq = y;
  const uint32_t k = (uint32_t)__builtin_fma(y, md, 2147483648.5);
  const uint64_t a = (uint64_t)s0 * m + c31;
  const uint64_t b = (uint64_t)k * (uint32_t)pn + a;
  const uint32_t hi = (uint32_t)((uint64_t)(uint32_t)s1 * m + (b >> 32)) + k * (uint32_t)(pn >> 32);
  return (int64_t)(((uint64_t)hi << 32) | (uint32_t)b);
}

// Paired signed step: s*M mod p for M = a*(a+1), the product of two
// consecutive multipliers a, a+1 <= n1, for |s| <= 0.51 p (fs_sh_pair_ok:
// n1 <= 2^22, so M < 2^44, and 2^45 <= p < 2^64 - 2^58).  y is the
// estimate of s/p that the previous step left (fs_sh_estimate): within
// 2^-52 |s/p| + 2^-21/p of it (two roundings of the sum and one of
// fl(1/p), plus the rounding of s0 * fl(1/p) with s0 < 2^32), so y*M misses
// X = s*M/p by less than 2^44 * 2^-52 * 0.51 + 2^44 * 2^-21 / 2^45 < 0.0022.
// kd = fma(y, M, 1.5 * 2^52) is
// 1.5 * 2^52 + k with k = round(y*M), |k| < 2^51, exactly (spacing 1 there),
// so the bits of kd are 0x4338000000000000 + k: k mod 2^32 is their low word
// k0 and floor(k / 2^32) their high word kb1 minus 0x43380000.  r = s*M - k*p
// then has |r| <= (1/2 + 0.0022) p < 0.989 * 2^63 and is formed exactly mod 2^64 as
// s*M + k*pn (pn = 2^64 - p) from two 32x32->64 multiply-adds and four 32-bit
// products: k*pn = k0*pn + kb1*pn0*2^32 - 0x43380000*pn0*2^32, the last term
// a per-modulus constant folded into the first multiply-add's addend
// (fs_sh_pair_c).  MS0, MS1 are the words of M (scalar).  *yr = the estimate
// of r/p for the next step: nine operations for the step and four for the
// estimate.  Because a+1 <= 2^22 < 0.49 p, a result congruent to +-(a+1), as
// it is when the skipped residue (that of s times a) is +-1, is exactly
// +-(a+1).
#define FS_SH_PAIR_PLIM (0xFFFFFFFFFFFFFFFFull - ((uint64_t)1 << 58) + 1)  // 2^64 - 2^58, exclusive
FS_HD bool fs_sh_pair_ok(uint64_t p, uint32_t n1) {
  return n1 <= (1u << 22) && p >= ((uint64_t)1 << 45) && p < FS_SH_PAIR_PLIM;
}
FS_HD uint64_t fs_sh_pair_c(uint64_t pn) { return (uint64_t)(0u - 0x43380000u * (uint32_t)pn) << 32; }
FS_HD double fs_sh_estimate(int64_t s, double pinv, double pinv32) {
#pragma clang fp contract(off)
  return __builtin_fma((double)(int32_t)((uint64_t)s >> 32), pinv32, (double)(uint32_t)s * pinv);
}
FS_HD int64_t fs_sh_mulmod_pair(int64_t s, double y, uint32_t MS0, uint32_t MS1, double Md, uint64_t pn,
                                uint64_t cpair, double pinv, double pinv32, double* yr) {
#pragma clang fp contract(off)
  const double kd = __builtin_fma(y, Md, 6755399441055744.0);  // 1.5 * 2^52
  const uint64_t kb = __builtin_bit_cast(uint64_t, kd);
  const uint32_t k0 = (uint32_t)kb, kb1 = (uint32_t)(kb >> 32);
  const uint32_t s0 = (uint32_t)s, s1 = (uint32_t)((uint64_t)s >> 32);
  const uint64_t u = (uint64_t)k0 * (uint32_t)pn + ((uint64_t)s0 * MS0 + cpair);
  const uint32_t hi = (uint32_t)(u >> 32) + s1 * MS0 + s0 * MS1 + k0 * (uint32_t)(pn >> 32) + kb1 * (uint32_t)pn;
  const int64_t r = (int64_t)(((uint64_t)hi << 32) | (uint32_t)u);
  *yr = fs_sh_estimate(r, pinv, pinv32);
  return r;
}

// Symmetric signed residue from a canonical one, and back.
FS_HD int64_t fs_canon_to_signed(uint64_t r, uint64_t p) { return r > (p >> 1) ? (int64_t)(r - p) : (int64_t)r; }
FS_HD uint64_t fs_signed_to_canon(int64_t s, uint64_t p) { return s < 0 ? (uint64_t)s + p : (uint64_t)s; }

// --------------------------------------------- lazy Montgomery (REDC by 2^64)

// s*m*2^-64 mod p for s < 2p, m < 2^63, odd p < 2^63, qn = -p^-1 mod 2^64;
// result < 2p (s*m < p*2^64, so (s*m + u*p) / 2^64 < 2p < 2^64).  The setup
// program's square on the hybrid path below 2^63, and its multiply between
// 2^64/3 and 2^63.  t = s*m = th:t1:t0 and
// u = (t mod 2^64) * qn mod 2^64 = u1:u0 from 32x32 partial products; the
// low 64 bits of t + u*p are zero, so the carry out of bit 63 is (t1 != 0)
// once bits 0..31 have been folded into z00 (t1 + bits 32..63 of
// t0 + u*p is 0 or 2^32, and 2^32 exactly when t1 != 0).
FS_HD uint64_t fs_redc64_mul(uint64_t s, uint64_t m, uint64_t p, uint64_t qn) {
  const uint32_t s0 = (uint32_t)s, s1 = (uint32_t)(s >> 32);
  const uint32_t m0 = (uint32_t)m, m1 = (uint32_t)(m >> 32);
  const uint64_t x00 = (uint64_t)s0 * m0;
  const uint64_t x01 = (uint64_t)s0 * m1 + (x00 >> 32);
  const uint64_t x10 = (uint64_t)s1 * m0 + (uint32_t)x01;
  const uint64_t th = (uint64_t)s1 * m1 + (x01 >> 32) + (x10 >> 32);
  const uint32_t t0 = (uint32_t)x00, t1 = (uint32_t)x10;
  const uint64_t y = (uint64_t)t0 * (uint32_t)qn;
  const uint32_t u0 = (uint32_t)y;
  const uint32_t u1 = (uint32_t)(y >> 32) + t0 * (uint32_t)(qn >> 32) + t1 * (uint32_t)qn;
  const uint64_t z00 = (uint64_t)u0 * (uint32_t)p + t0;  // bits 0..31 are zero
  const uint64_t z01 = (uint64_t)u0 * (uint32_t)(p >> 32) + (z00 >> 32);
  const uint64_t z10 = (uint64_t)u1 * (uint32_t)p + (uint32_t)z01;
  const uint64_t w = (uint64_t)u1 * (uint32_t)(p >> 32) + (z01 >> 32);
  return th + w + (z10 >> 32) + (t1 != 0);
}

// Largest modulus (exclusive) of fs_redc62_mul: its first stage leaves a value
// below 3p, which must fit in 64 bits.
#define FS_REDC62_PMAX 0x5555555555555555ull

// s*m*2^-62 mod p for s < 2p, m < 2^60, odd p < 2^64/3, pn = -p^-1 mod 2^32;
// result < 2p.  Two REDC-by-2^31 stages over m = m1*2^31 + m0 (m0 < 2^31,
// m1 < 2^29): y = (s*m0 + u*p) / 2^31 < 3p, then (s*m1 + y + u'*p) / 2^31
// < p/2 + p + 3p*2^-31 < 2p, with u, u' < 2^31 chosen to clear the low 31
// bits.  Reducing by 2^31 rather than 2^32 keeps every 32x32+64 partial sum
// below 2^64 (u*p0 < 2^63 next to s0*m0 < 2^63), and the division by 2^31 is
// one 64-bit shift and one shift-add: (L + X*2^32) / 2^31 = 2X + L/2^31.
// Eight 32x32->64 multiply-adds and two 32-bit multiplies per 60-bit word,
// about three primes of a factorial's power table: the setup program's
// multiply on the hybrid path below 2^64/3.
FS_HD uint64_t fs_redc62_mul(uint64_t s, uint64_t m, uint64_t p, uint32_t pn) {
  const uint32_t m0 = (uint32_t)m & 0x7FFFFFFFu, m1 = (uint32_t)(m >> 31);
  const uint32_t p0 = (uint32_t)p, p1 = (uint32_t)(p >> 32);
  const uint32_t s0 = (uint32_t)s, s1 = (uint32_t)(s >> 32);
  const uint64_t a = (uint64_t)s0 * m0;  // < 2^63
  const uint32_t u = ((uint32_t)a * pn) & 0x7FFFFFFFu;
  const uint64_t l = (uint64_t)u * p0 + a;  // < 2^64, low 31 bits zero
  const uint64_t x = (uint64_t)u * p1 + (uint64_t)s1 * m0;
  const uint64_t y = (x << 1) + (l >> 31);  // < 3p
  const uint64_t a2 = (uint64_t)s0 * m1 + (uint32_t)y;  // < 2^62
  const uint32_t u2 = ((uint32_t)a2 * pn) & 0x7FFFFFFFu;
  const uint64_t l2 = (uint64_t)u2 * p0 + a2;
  const uint64_t x2 = (uint64_t)u2 * p1 + ((uint64_t)s1 * m1 + (y >> 32));
  return (x2 << 1) + (l2 >> 31);
}

// 2^e * 2^64 mod p (Montgomery form of 2^e) for e >= 1: left-to-right
// square-and-double starting from 2, one = 2^64 mod p.
FS_HD uint64_t fs_mont_pow2(uint64_t e, uint64_t p, uint64_t q, uint64_t one) {
  uint64_t a = fs_add_mod(one, one, p);
  const int top = 63 - __builtin_clzll(e);
  for (int b = top - 1; b >= 0; --b) {
    a = fs_mont_mul(a, a, p, q);
    if ((e >> b) & 1) a = fs_add_mod(a, a, p);
  }
  return a;
}

// ------------------------------------------------------------ candidate tests

// Base-2 strong probable-prime test for odd N > 2 (any N < 2^64), Montgomery
// arithmetic with q = N^-1 mod 2^64 and one = 2^64 mod N (getsegprimes.cl:79-108,
// left-to-right square-and-double over the odd part d of N-1).
FS_HD bool fs_spsp2(uint64_t N, uint64_t q, uint64_t one) {
  const uint64_t nmo = N - one;
  uint64_t d = N - 1;
  const int t = __builtin_ctzll(d);
  d >>= t;
  uint64_t a = fs_add_mod(one, one, N);  // 2 in Montgomery form
  if (d > 1) {
    const int top = 63 - __builtin_clzll(d);
    for (int b = top - 1; b >= 0; --b) {
      a = fs_mont_mul(a, a, N, q);
      if ((d >> b) & 1) a = fs_add_mod(a, a, N);
    }
  }
  if (a == one || a == nmo) return true;
  for (int s = 1; s < t; ++s) {
    a = fs_mont_mul(a, a, N, q);
    if (a == nmo) return true;
  }
  return false;
}

// 2^64 mod N for odd N >= 3: 2^64 - N above 2^63; below, 2^63 mod N from one
// exact signed step 2^33 * 2^30 (valid as 2^33 <= 0.51 N for N >= 2^34), then
// doubled; a plain 64-bit remainder below 2^34.
FS_HD uint64_t fs_r64_mod(uint64_t N) {
  if (N > ((uint64_t)1 << 63)) return 0 - N;
  if (N < ((uint64_t)1 << 34)) return (0 - N) % N;
  const double pinv = 1.0 / (double)N;
  const int64_t h = fs_sh_mulmod((int64_t)1 << 33, 1u << 30, 1073741824.0, N << 31, 0 - N, pinv, pinv * 4294967296.0);
  const uint64_t r = fs_signed_to_canon(h, N);
  return fs_add_mod(r, r, N);
}

// Largest N (exclusive) of the FP64 probable-prime test.
#define FS_SPRP_FP_MAX ((uint64_t)1 << 50)

// Base-2 strong probable-prime test of odd 5 <= N < 2^50 in FP64 arithmetic
// (same predicate as fs_spsp2).  Residues are integer doubles |a| < 0.8 N:
// a square is fs_fp_mulmod with the multiplier a itself (|a| < 2^50; a^2/N
// < 0.64 N keeps the quotient error below 3*2^-53*0.64*2^50 = 0.24, so
// |a'| < 0.74 N, and |h - k N| < 0.74 N + 2^46 < 2^53); a doubling is reduced
// back to |a| <= N/2.  a == +-1 (mod N) iff |a| == 1.
FS_HD bool fs_spsp2_fp(uint64_t N) {
  const double nd = (double)N, ninv = 1.0 / nd, half = 0.5 * nd;
  uint64_t d = N - 1;
  const int t = __builtin_ctzll(d);
  d >>= t;
  double a = 2.0;
  if (d > 1) {
    const int top = 63 - __builtin_clzll(d);
    for (int b = top - 1; b >= 0; --b) {
      a = fs_fp_mulmod(a, a, nd, ninv);
      if ((d >> b) & 1) {
        a = a + a;
        a = a > half ? a - nd : (a < -half ? a + nd : a);
      }
    }
  }
  if (__builtin_fabs(a) == 1.0) return true;
  for (int s = 1; s < t; ++s) {
    a = fs_fp_mulmod(a, a, nd, ninv);
    if (a == -1.0) return true;
  }
  return false;
}

// The candidate test of a sieve survivor N (odd, > 113): base-2 strong
// probable prime, in FP64 below FS_SPRP_FP_MAX and in Montgomery arithmetic
// above.
FS_HD bool fs_candidate_probable_prime(uint64_t N) {
  if (N < FS_SPRP_FP_MAX) return fs_spsp2_fp(N);
  const uint64_t q = fs_invert(N);
  return fs_spsp2(N, q, fs_r64_mod(N));
}

// Odd primes 3..113 removed before the probable-prime test: 3 and 5 (the
// reference's mod-30 wheel) and 7..113 (its bit tables, getsegprimes.cl:115-167).
#define FS_NSMALL 29
#define FS_NMASK 1591  // sum of the 29 primes: one 64-bit mask per residue class
static constexpr uint32_t kSmallQ[FS_NSMALL] = {3,  5,  7,  11, 13, 17, 19, 23, 29, 31,
                                                37, 41, 43, 47, 53, 59, 61, 67, 71, 73,
                                                79, 83, 89, 97, 101, 103, 107, 109, 113};
FS_HD constexpr uint32_t fs_small_off(int j) {
  uint32_t o = 0;
  for (int k = 0; k < j; ++k) o += kSmallQ[k];
  return o;
}
static_assert(fs_small_off(FS_NSMALL) == FS_NMASK, "mask table size");

// Mask of residue class r of prime q: bit i set iff q divides B + 2i for odd
// B = r (mod q), i.e. i = i0 (mod q) with i0 = (q - r) * 2^-1 mod q.
FS_HD constexpr uint64_t fs_small_mask(uint32_t q, uint32_t r) {
  const uint32_t i0 = ((q - r) % q) * ((q + 1) >> 1) % q;
  uint64_t m = 0;
  for (uint32_t i = i0; i < 64; i += q) m |= (uint64_t)1 << i;
  return m;
}
