// fs_hostemu.cpp -- the per-candidate loops of the sieve kernels executed on the
// host with the same inline arithmetic (fs_arith.hpp), one candidate at a time.
//
// These entry points exist so that the CPU test suite can check each arithmetic
// formulation (FP64 / hybrid / Montgomery) against the oracle at sizes and
// moduli where a GPU run would be needed otherwise.  The sieve driver never
// calls them: fs_sieve, fs_batch and the server run only the gfx950 kernels.
#include <stdint.h>
#include <string.h>
#include <numeric>
#include <vector>
#include "fs_arith.hpp"
#include "fs_internal.hpp"
#include "fs_kern.hpp"
#include "../../include/factsieve.h"

extern "C" int fs_hostemu_mulmod(int path, uint64_t p, int64_t s, uint64_t m, int64_t* raw, uint64_t* res) {
  if ((p & 1) == 0 || p < 7 || !raw || !res) return FS_E_ARG;
  const uint64_t as = s < 0 ? (uint64_t)0 - (uint64_t)s : (uint64_t)s;
  if (path == FS_PATH_FP64) {
    if (p >= ((uint64_t)1 << 53) || m >= ((uint64_t)1 << 50) || (double)as >= 0.8 * (double)p) return FS_E_ARG;
    const double pd = (double)p;
    const double r = fs_fp_mulmod((double)s, (double)m, pd, 1.0 / pd);
    *raw = (int64_t)r;
    *res = fs_fp_to_canon(r, p);
    return FS_OK;
  }
  if (path == FS_PATH_HYBRID) {
    if (p > FS_HYBRID_PMAX || m >= ((uint64_t)1 << 31) || (p < ((uint64_t)1 << 31) && m >= 2 * p) ||
        (double)as > 0.51 * (double)p ||
        as >= ((uint64_t)1 << 63))
      return FS_E_ARG;
    const double pinv = 1.0 / (double)p;
    const int64_t r = fs_sh_mulmod(s, (uint32_t)m, (double)m, p << 31, 0 - p, pinv, pinv * 4294967296.0);
    *raw = r;
    *res = fs_signed_to_canon(r, p);
    return FS_OK;
  }
  return FS_E_ARG;
}

extern "C" int fs_hostemu_fp_exact(uint64_t p, int64_t s, uint32_t m, int64_t* raw, uint64_t* res) {
  if ((p & 1) == 0 || p < 7 || m == 0 || !fs_fp_exact_ok(p, m) || !raw || !res) return FS_E_ARG;
  const uint64_t as = s < 0 ? (uint64_t)0 - (uint64_t)s : (uint64_t)s;
  if ((double)as > 0.51 * (double)p) return FS_E_ARG;
  const double pd = (double)p;
  const double r = fs_fp_mulmod_exact((double)s, (double)m, pd, 1.0 / pd);
  *raw = (int64_t)r;
  *res = fs_fp_to_canon(r, p);
  return FS_OK;
}

extern "C" int fs_hostemu_mulmod_hi(uint64_t p, int64_t s, uint32_t m, int64_t* raw, uint64_t* res) {
  if ((p & 1) == 0 || p > FS_HYBRID_PMAX || m == 0 || !fs_sh_hi_ok(p, m) || !raw || !res) return FS_E_ARG;
  const uint64_t as = s < 0 ? (uint64_t)0 - (uint64_t)s : (uint64_t)s;
  if ((double)as > FS_SH_HI_RBOUND * (double)p) return FS_E_ARG;
  const double pinv32 = (1.0 / (double)p) * 4294967296.0;
  double y;
  const int64_t r = fs_sh_mulmod_hi(s, m, (double)m, p << 31, 0 - p, pinv32, &y);
  *raw = r;
  *res = fs_signed_to_canon(r, p);
  return FS_OK;
}

extern "C" int fs_hostemu_fp_pair(uint64_t p, int64_t s, uint32_t a, int64_t* raw, uint64_t* res, int* exact) {
  if ((p & 1) == 0 || p < 7 || p >= ((uint64_t)1 << 53) || a == 0 || a >= (1u << 22) || !fs_fp_pair_ok(p, a + 1) ||
      !raw || !res || !exact)
    return FS_E_ARG;
  const uint64_t as = s < 0 ? (uint64_t)0 - (uint64_t)s : (uint64_t)s;
  if ((double)as > 0.51 * (double)p) return FS_E_ARG;
  const double pd = (double)p, M = (double)((uint64_t)a * (a + 1));
  *exact = fs_fp_pair_exact_ok(p, a + 1);
  const double r = *exact ? fs_fp_mulmod_exact((double)s, M, pd, 1.0 / pd) : fs_fp_mulmod((double)s, M, pd, 1.0 / pd);
  *raw = (int64_t)r;
  *res = fs_fp_to_canon(r, p);
  return FS_OK;
}

extern "C" int fs_hostemu_sh_pair(uint64_t p, int64_t s, uint32_t a, int64_t* raw, uint64_t* res) {
  if ((p & 1) == 0 || a == 0 || a >= (1u << 22) || !fs_sh_pair_ok(p, a + 1) || !raw || !res) return FS_E_ARG;
  const uint64_t as = s < 0 ? (uint64_t)0 - (uint64_t)s : (uint64_t)s;
  if ((double)as > 0.51 * (double)p) return FS_E_ARG;
  const double pinv = 1.0 / (double)p, pinv32 = pinv * 4294967296.0;
  const uint64_t pn = 0 - p, M = (uint64_t)a * (a + 1);
  double y;
  const int64_t r = fs_sh_mulmod_pair(s, fs_sh_estimate(s, pinv, pinv32), (uint32_t)M, (uint32_t)(M >> 32),
                                      (double)M, pn, fs_sh_pair_c(pn), pinv, pinv32, &y);
  *raw = r;
  *res = fs_signed_to_canon(r, p);
  return FS_OK;
}

extern "C" int fs_hostemu_setup_mul(uint64_t p, uint64_t s, uint64_t m, int square, uint64_t* raw, uint64_t* res,
                                    uint32_t* shift) {
  if ((p & 1) == 0 || p < 3 || p > FS_HYBRID_PMAX || m >= ((uint64_t)1 << 60) || !raw || !res || !shift)
    return FS_E_ARG;
  const uint64_t q = fs_invert(p);
  if (p < ((uint64_t)1 << 63)) {
    if (s >= 2 * p) return FS_E_ARG;
    if (square) {
      const uint64_t a = s >= p ? s - p : s;  // as the setup loop squares
      *raw = fs_redc64_mul(a, a, p, 0 - q);
      *shift = 64;
    } else if (p < FS_REDC62_PMAX) {
      *raw = fs_redc62_mul(s, m, p, (uint32_t)(0 - q));
      *shift = 62;
    } else {
      *raw = fs_redc64_mul(s, m, p, 0 - q);
      *shift = 64;
    }
    *res = *raw >= p ? *raw - p : *raw;
  } else {
    if (s >= p) return FS_E_ARG;
    *raw = *res = square ? fs_mont_mul(s, s, p, q) : fs_mont_mul(s, m, p, q);
    *shift = 64;
  }
  return FS_OK;
}

extern "C" int fs_hostemu_setup(int path, uint64_t p, uint32_t nmin, uint64_t* res) {
  if ((p & 1) == 0 || p < 3) return FS_E_ARG;
  const uint64_t q = fs_invert(p), one = fs_r64_mod(p), r2 = fs_r2_from_one(one, p, q);
  const int64_t T = fs_setup_program(nmin, path, nullptr);
  if (T < 0) return (int)T;
  std::vector<uint64_t> w((size_t)(path == FS_PATH_MONT ? 2 * T : T) + 2);
  fs_setup_program(nmin, path, w.data());
  if (path == FS_PATH_FP64) {
    if (p >= ((uint64_t)1 << 53)) return FS_E_ARG;
    const double pd = (double)p, pinv = 1.0 / pd;
    double s = 1.0;
    for (int64_t k = 0; k < T; ++k) {
      if (w[k] & FS_PROG_SQUARE) {  // as fs_cand_setup_fp64 squares (one modulus: its own wave)
        if (fs_fp_square_ok(p)) {
          s = fs_fp_mulmod(s, s, pd, pinv);
        } else {
          const uint64_t a = fs_fp_to_canon(s, p);
          s = fs_canon_to_fp(fs_mul_mod(a, a, p, q, r2), p);
        }
      }
      double m;
      const uint64_t bits = w[k] & ~FS_PROG_SQUARE;
      memcpy(&m, &bits, 8);
      s = fs_fp_mulmod(s, m, pd, pinv);
    }
    *res = fs_fp_to_canon(s, p);
  } else if (path == FS_PATH_HYBRID) {
    if (p > FS_HYBRID_PMAX) return FS_E_ARG;
    // the device routine's products (fs_cand_setup_hybrid), one term at a time
    uint64_t s = 1;
    const uint64_t qn = 0 - q;
    for (int64_t k = 0; k < T; ++k) {
      const uint64_t m = w[k] & ~FS_PROG_SQUARE;
      if (p < ((uint64_t)1 << 63)) {
        if (w[k] & FS_PROG_SQUARE) {
          const uint64_t a = s >= p ? s - p : s;
          s = fs_redc64_mul(a, a, p, qn);
        }
        s = (p < FS_REDC62_PMAX) ? fs_redc62_mul(s, m, p, (uint32_t)qn) : fs_redc64_mul(s, m, p, qn);
      } else {
        if (w[k] & FS_PROG_SQUARE) s = fs_mont_mul(s, s, p, q);
        s = fs_mont_mul(s, m, p, q);
      }
    }
    s = s >= p ? s - p : s;
    const uint64_t E = (p < FS_REDC62_PMAX) ? w[T] : w[T + 1];
    if (E != 0) s = fs_mont_mul(s, fs_mont_pow2(E, p, q, one), p, q);
    *res = s;
  } else if (path == FS_PATH_MONT) {
    uint64_t s6 = one;
    for (int64_t k = 0; k < T; ++k) {
      const uint32_t e = (uint32_t)w[T + k];
      uint32_t cb = (uint32_t)(w[T + k] >> 32);
      if (k == 0) {
        uint64_t a = fs_add_mod(one, one, p);
        while (cb) {
          a = fs_mont_mul(a, a, p, q);
          if (e & cb) a = fs_add_mod(a, a, p);
          cb >>= 1;
        }
        s6 = a;
        continue;
      }
      const uint64_t b = fs_mont_mul(w[k], r2, p, q);
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
    *res = fs_mont_mul(s6, 1, p, q);
  } else {
    return FS_E_ARG;
  }
  return FS_OK;
}

extern "C" int fs_hostemu_iterate(int path, uint64_t p, uint32_t n0, uint32_t n1, uint64_t* res,
                                  fs_factor* fac, uint32_t cap, uint32_t* nf) {
  if ((p & 1) == 0 || p < 3 || n0 > n1 || *res >= p) return FS_E_ARG;
  // fs_sh_mulmod's multiplier bound m < 2p (n1 < 2^32 <= 2p for p >= 2^31).
  if (path == FS_PATH_HYBRID && *res != 0 && p < ((uint64_t)1 << 31) && (uint64_t)n1 >= 2 * p) return FS_E_ARG;
  const uint64_t q = fs_invert(p), one = fs_r64_mod(p), r2 = fs_r2_from_one(one, p, q);
  uint32_t cnt = 0;
  auto rec = [&](int32_t nc) {
    if (cnt < cap) {
      fac[cnt].p = p;
      fac[cnt].nc = nc;
      fac[cnt].pad = 0;
    }
    ++cnt;
  };
  // The device loops' blocks of U = 8 steps from n0 with their +-1 filter (one
  // test per two residues, see fs_kern.hpp), then single exact steps; a +-1
  // residue inside a block whose filter does not fire is an error.
  constexpr uint32_t U = FS_UNROLL_DEFAULT;
  const uint32_t nu = n0 + ((n1 - n0) & ~(U - 1));
  if (path == FS_PATH_FP64) {
    if (p >= ((uint64_t)1 << 53)) return FS_E_ARG;
    const bool exact = fs_fp_exact_ok(p, n1);  // the device loops' choice (per wave there)
    const double pd = (double)p, pinv = 1.0 / pd;
    auto step = [&](double s, uint32_t m) {
      return exact ? fs_fp_mulmod_exact(s, (double)m, pd, pinv) : fs_fp_mulmod(s, (double)m, pd, pinv);
    };
    double s = fs_canon_to_fp(*res, p);
    uint32_t n = n0;
    if (fs_fp_pair_ok(p, n1)) {
      // The paired loop: one product by M = a*(a+1) per two positions; a
      // block from nb owns positions nb+1 .. nb+U and its filter is one
      // compare per pair, |v| <= nb+U on the product.  The odd positions'
      // residues are formed here by single steps to check each pair against
      // two single steps and that the filter fires for every block holding a
      // +-1; position n0 is tested at the start and the positions after the
      // last whole block by single steps.
      const bool pexact = fs_fp_pair_exact_ok(p, n1);
      auto pstep = [&](double x, uint64_t M) {
        return pexact ? fs_fp_mulmod_exact(x, (double)M, pd, pinv) : fs_fp_mulmod(x, (double)M, pd, pinv);
      };
      if (n0 < n1 && __builtin_fabs(s) == 1.0) rec(s > 0.0 ? -(int32_t)n0 : (int32_t)n0);
      for (; n < nu; n += U) {
        double v[U / 2 + 1], u[U / 2];
        v[0] = s;
        bool fire = false, any = false;
        for (uint32_t i = 0; i < U / 2; ++i) {
          const uint64_t a = n + 1 + 2 * i;
          v[i + 1] = pstep(v[i], a * (a + 1));
          if (__builtin_fabs(v[i + 1]) > 0.51 * pd) return FS_E_VALIDATE;  // the loop's invariant
          fire |= __builtin_fabs(v[i + 1]) <= (double)(n + U);
          u[i] = fs_fp_mulmod(v[i], (double)a, pd, pinv);
          if (fs_fp_to_canon(pstep(u[i], a + 1), p) != fs_fp_to_canon(v[i + 1], p)) return FS_E_VALIDATE;
        }
        for (uint32_t i = 0; i < U / 2; ++i) {
          const double r[2] = {u[i], v[i + 1]};  // residues of n + 2i + 1 and n + 2i + 2
          for (uint32_t k = 0; k < 2; ++k) {
            const uint32_t nn = n + 2 * i + 1 + k;
            if (nn >= n1 || __builtin_fabs(r[k]) != 1.0) continue;
            any = true;
            rec(r[k] > 0.0 ? -(int32_t)nn : (int32_t)nn);
          }
        }
        if (any && !fire) return FS_E_VALIDATE;  // the filter would have missed it
        s = v[U / 2];
      }
      for (; n < n1; ++n) {
        s = pstep(s, n + 1);
        if (n + 1 < n1 && __builtin_fabs(s) == 1.0) rec(s > 0.0 ? -(int32_t)(n + 1) : (int32_t)(n + 1));
      }
      *res = fs_fp_to_canon(s, p);
      *nf = cnt;
      return FS_OK;
    }
    for (; n < nu; n += U) {
      double v[U + 1];
      v[0] = s;
      for (uint32_t k = 0; k < U; ++k) v[k + 1] = step(v[k], n + 1 + k);
      bool hit = false;
      for (uint32_t k = 0; k < U; k += 2) hit |= __builtin_fabs(v[k + 1]) <= (double)(n + 1 + k);
      for (uint32_t k = 0; k < U; ++k) {
        if (__builtin_fabs(v[k]) != 1.0) continue;
        if (!hit) return FS_E_VALIDATE;  // the filter would have missed it
        rec(v[k] > 0.0 ? -(int32_t)(n + k) : (int32_t)(n + k));
      }
      s = v[U];
    }
    for (; n < n1; ++n) {
      if (__builtin_fabs(s) == 1.0) rec(s > 0.0 ? -(int32_t)n : (int32_t)n);
      s = step(s, n + 1);
    }
    *res = fs_fp_to_canon(s, p);
  } else if (path == FS_PATH_HYBRID && fs_sh_pair_ok(p, n1)) {
    // The paired loop: one product by M = a*(a+1) per two positions, each
    // leaving the estimate y of its result over p; a block from nb owns
    // positions nb+1 .. nb+U and its filter is |y| <= 1.5 (n1 + 2)/p over its
    // U/2 results.  Here each pair is also checked against two single steps
    // and the loop's residue bound, and a +-1 residue in a block whose filter
    // does not fire is an error.
    const double pinv = 1.0 / (double)p, pinv32 = pinv * 4294967296.0;
    const double thr = 1.5 * (double)(n1 + 2) * pinv;
    const uint64_t pn = 0 - p, cpair = fs_sh_pair_c(pn), c31 = p << 31;
    int64_t s = fs_canon_to_signed(*res, p);
    if (n0 < n1 && ((s == 1) | (s == -1))) rec(s == 1 ? -(int32_t)n0 : (int32_t)n0);
    double ys = fs_sh_estimate(s, pinv, pinv32);
    uint32_t n = n0;
    for (; n < nu; n += U) {
      int64_t v[U / 2 + 1];
      v[0] = s;
      bool fire = false, any = false;
      for (uint32_t i = 0; i < U / 2; ++i) {
        const uint64_t a = n + 1 + 2 * i, M = a * (a + 1);
        v[i + 1] = fs_sh_mulmod_pair(v[i], ys, (uint32_t)M, (uint32_t)(M >> 32), (double)M, pn, cpair, pinv,
                                     pinv32, &ys);
        if (__builtin_fabs((double)v[i + 1]) > 0.51 * (double)p) return FS_E_VALIDATE;  // the loop's bound
        fire |= __builtin_fabs(ys) <= thr;
        double yu;
        const int64_t u = fs_sh_mulmod(v[i], (uint32_t)a, (double)a, c31, pn, pinv, pinv32, &yu);
        const int64_t two = fs_sh_mulmod(u, (uint32_t)(a + 1), (double)(a + 1), c31, pn, pinv, pinv32, &yu);
        if (fs_signed_to_canon(two, p) != fs_signed_to_canon(v[i + 1], p)) return FS_E_VALIDATE;
      }
      int64_t w = v[0];
      for (uint32_t nn = n; nn < n + U; ++nn) {  // positions n + 1 .. n + U
        double yu;
        w = fs_sh_mulmod(w, nn + 1, (double)(nn + 1), c31, pn, pinv, pinv32, &yu);
        if (nn + 1 >= n1 || ((w != 1) & (w != -1))) continue;
        any = true;
        rec(w == 1 ? -(int32_t)(nn + 1) : (int32_t)(nn + 1));
      }
      if (any && !fire) return FS_E_VALIDATE;  // the filter would have missed it
      s = v[U / 2];
    }
    for (; n < n1; ++n) {
      double y;
      s = fs_sh_mulmod(s, n + 1, (double)(n + 1), c31, pn, pinv, pinv32, &y);
      if (n + 1 < n1 && ((s == 1) | (s == -1))) rec(s == 1 ? -(int32_t)(n + 1) : (int32_t)(n + 1));
    }
    *res = fs_signed_to_canon(s, p);
  } else if (path == FS_PATH_HYBRID) {
    if (p > FS_HYBRID_PMAX) return FS_E_ARG;
    const bool hi = fs_sh_hi_ok(p, n1);  // the device loops' choice (per wave there)
    const double pinv = 1.0 / (double)p, pinv32 = pinv * 4294967296.0;
    const double lim = (hi ? FS_SH_HI_RBOUND : 0.51) * (double)p;  // the loop's residue bound
    bool out = false;
    auto step = [&](int64_t s, uint32_t m, double* y) {
      const int64_t r = hi ? fs_sh_mulmod_hi(s, m, (double)m, p << 31, 0 - p, pinv32, y)
                           : fs_sh_mulmod(s, m, (double)m, p << 31, 0 - p, pinv, pinv32, y);
      // (the bound assumes m < 2p, which the sieve's launches meet: its residue is 0 for n >= p)
      out |= (uint64_t)m < 2 * p && __builtin_fabs((double)r) > lim;
      return r;
    };
    int64_t s = fs_canon_to_signed(*res, p);
    uint32_t n = n0;
    for (; n < nu; n += U) {
      int64_t v[U + 1];
      double y[U];
      v[0] = s;
      for (uint32_t k = 0; k < U; ++k) v[k + 1] = step(v[k], n + 1 + k, &y[k]);
      double ymin = __builtin_fabs(y[1]);
      for (uint32_t k = 3; k < U; k += 2) ymin = __builtin_fmin(ymin, __builtin_fabs(y[k]));
      const double thr = hi ? 1.5 * pinv32 : (1.5 * pinv) * (double)(n + U - 1);
      for (uint32_t k = 0; k < U; ++k) {
        if ((v[k] != 1) & (v[k] != -1)) continue;
        if (!(ymin <= thr)) return FS_E_VALIDATE;  // the filter would have missed it
        rec(v[k] == 1 ? -(int32_t)(n + k) : (int32_t)(n + k));
      }
      s = v[U];
    }
    for (; n < n1; ++n) {
      if ((s == 1) | (s == -1)) rec(s == 1 ? -(int32_t)n : (int32_t)n);
      double y;
      s = step(s, n + 1, &y);
    }
    if (out) return FS_E_VALIDATE;  // a residue left the bound the step's next input needs
    *res = fs_signed_to_canon(s, p);
  } else if (path == FS_PATH_MONT) {
    const uint64_t nmo = p - one;
    uint64_t s6 = fs_mont_mul(*res, r2, p, q);
    uint64_t s7 = fs_mont_mul(n0 < p ? (uint64_t)n0 : (uint64_t)n0 % p, r2, p, q);
    for (uint32_t n = n0; n < n1; ++n) {
      s7 = fs_add_mod(s7, one, p);
      if (s6 == one || s6 == nmo) rec((s6 == one) ? -(int32_t)n : (int32_t)n);
      s6 = fs_mont_mul(s6, s7, p, q);
    }
    *res = fs_mont_mul(s6, 1, p, q);
  } else {
    return FS_E_ARG;
  }
  *nf = cnt;
  return FS_OK;
}

// The candidate set of [low, high) as the getsegprimes kernel forms it: the
// 64-odd-number masks of the 29 small primes, then the base-2 SPRP test.
extern "C" int64_t fs_hostemu_candidates(uint64_t low, uint64_t high, uint64_t* out, uint64_t cap) {
  if (low >= high) return 0;
  const uint64_t base = low | 1;
  if (base >= high) return 0;
  int64_t n = 0;
  for (uint64_t B = base; B < high; B += 128) {
    uint64_t comp = 0;
    for (int j = 0; j < FS_NSMALL; ++j) comp |= fs_small_mask(kSmallQ[j], (uint32_t)(B % kSmallQ[j]));
    uint64_t surv = ~comp;
    const uint64_t left = (high - B + 1) >> 1;
    if (left < 64) surv &= (((uint64_t)1 << left) - 1);
    while (surv) {
      const int bit = __builtin_ctzll(surv);
      surv &= surv - 1;
      const uint64_t N = B + 2 * (uint64_t)bit;
      if (N <= 113) continue;
      if (fs_candidate_probable_prime(N)) {
        if (out && (uint64_t)n < cap) out[n] = N;
        ++n;
      }
    }
    if (B > UINT64_MAX - 128) break;
  }
  return n;
}
