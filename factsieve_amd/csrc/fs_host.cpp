// fs_host.cpp -- host side of the FactSieve sieve on the C ABI: setup programs,
// CPU factor verifier, number parsing, the segment driver (cl_sieve), result
// drain, factors.txt, checkpoint/resume and the self-test.
//
// Reference counterparts: cl_sieve.cpp (driver, power table, drain, checkpoint,
// self-test), verifyprime.c (CPU verifier), putil.c (number parsing).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <errno.h>
#include <math.h>
#include <time.h>
#include <sys/resource.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <functional>
#include <future>
#include <map>
#include <memory>
#include <new>
#include <string>
#include <mutex>
#include <thread>
#include <vector>
#include "fs_arith.hpp"
#include "fs_internal.hpp"
#include "../../include/factsieve.h"

typedef unsigned __int128 u128;

// ============================================================ small primes

// Primes <= limit (odd-only bit sieve of Eratosthenes).  Replaces the
// reference's primesieve_generate_primes(2, nmin) call (cl_sieve.cpp:797).
static std::vector<uint32_t> primes_upto_uncached(uint32_t limit) {
  std::vector<uint32_t> out;
  if (limit < 2) return out;
  out.push_back(2);
  if (limit < 3) return out;
  // base primes up to sqrt(limit), then segments of 2^18 odd numbers (a
  // 32 KiB bit set, cache resident) marked by each base prime in turn
  uint32_t r = 1;
  while ((uint64_t)(r + 1) * (r + 1) <= limit) ++r;
  std::vector<uint32_t> base;
  {
    std::vector<char> c(r + 1, 0);
    for (uint32_t x = 3; x <= r; x += 2) {
      if (c[x]) continue;
      base.push_back(x);
      for (uint64_t y = (uint64_t)x * x; y <= r; y += 2 * x) c[y] = 1;
    }
  }
  const uint64_t nodd = ((uint64_t)limit - 1) / 2;  // odd numbers 3..limit: index (x-3)/2
  constexpr uint64_t kSeg = 1u << 18;
  std::vector<uint64_t> comp(kSeg / 64);
  std::vector<uint64_t> next(base.size());  // next odd multiple index of each base prime
  for (size_t k = 0; k < base.size(); ++k) next[k] = ((uint64_t)base[k] * base[k] - 3) / 2;
  for (uint64_t s0 = 0; s0 < nodd; s0 += kSeg) {
    const uint64_t s1 = std::min(nodd, s0 + kSeg);
    std::fill(comp.begin(), comp.end(), 0);
    for (size_t k = 0; k < base.size(); ++k) {
      uint64_t j = next[k];
      for (; j < s1; j += base[k]) comp[(j - s0) >> 6] |= (uint64_t)1 << ((j - s0) & 63);
      next[k] = j;
    }
    for (uint64_t w = 0; w < (s1 - s0 + 63) / 64; ++w) {
      uint64_t m = ~comp[w];
      while (m) {
        const uint64_t i = s0 + w * 64 + __builtin_ctzll(m);
        m &= m - 1;
        if (i >= s1) break;
        out.push_back((uint32_t)(2 * i + 3));
      }
    }
  }
  return out;
}

std::vector<uint32_t> fs_primes_upto(uint32_t limit) {
  static std::mutex mu;
  static uint32_t last_limit = 0;
  static std::vector<uint32_t> last;
  std::lock_guard<std::mutex> lk(mu);
  if (last.empty() || last_limit != limit) {
    last = primes_upto_uncached(limit);
    last_limit = limit;
  }
  return last;
}

// Legendre exponent of prime q in n! (getPower, cl_sieve.cpp:773-782).
static uint32_t legendre(uint32_t q, uint32_t n) {
  uint32_t e = 0;
  for (uint64_t qk = q; qk <= n; qk *= q) e += (uint32_t)(n / qk);
  return e;
}

// ============================================================ setup programs

extern "C" int64_t fs_power_table(uint32_t nmin, uint64_t* bases, uint32_t* exps, uint32_t* curbits) {
  const std::vector<uint32_t> pr = fs_primes_upto(nmin);
  int64_t m = 0;
  size_t i = 0;
  while (i < pr.size()) {
    const uint32_t e = legendre(pr[i], nmin);
    uint64_t b = pr[i];
    ++i;
    // entry 0 (the prime 2) stays alone; later runs of equal exponent are
    // merged while the product stays below 2^64 - 1 (cl_sieve.cpp:818-835)
    if (m > 0) {
      while (i < pr.size() && legendre(pr[i], nmin) == e) {
        const u128 pp = (u128)b * pr[i];
        if (pp >= (u128)UINT64_MAX) break;
        b = (uint64_t)pp;
        ++i;
      }
    }
    if (bases) bases[m] = b;
    if (exps) exps[m] = e;
    if (curbits) {
      uint32_t cb = 0x80000000u;
      if (e > 1) cb >>= (__builtin_clz(e) + 1);
      curbits[m] = cb;
    }
    ++m;
  }
  return m;
}

// Horner program over the binary digits of the Legendre exponents:
// nmin! = prod_j C_j^(2^j), C_j = product of primes q whose exponent has bit j
// set; evaluated as (((C_J)^2 C_{J-1})^2 ...) C_0 with each C_j split into
// multipliers below `bound`.  Returns the program as multipliers, 0 = square.
std::vector<uint64_t> fs_horner_program(uint32_t nmin, uint64_t bound) {
  std::vector<uint64_t> prog;
  const std::vector<uint32_t> pr = fs_primes_upto(nmin);
  if (pr.empty()) return prog;
  std::vector<uint32_t> ex(pr.size());
  uint32_t emax = 0;
  for (size_t i = 0; i < pr.size(); ++i) {
    ex[i] = legendre(pr[i], nmin);
    emax = std::max(emax, ex[i]);
  }
  const int J = 31 - __builtin_clz(emax);
  // primes of each level j (bit j of the exponent set), ascending, in one pass
  std::vector<std::vector<uint32_t>> lvl(J + 1);
  for (size_t i = 0; i < pr.size(); ++i)
    for (uint32_t e = ex[i]; e; e &= e - 1) lvl[__builtin_ctz(e)].push_back(pr[i]);
  // Each level's primes are packed first-fit-decreasing: a word starts with the
  // largest unused prime and takes the largest unused prime that still fits,
  // until none does.  "Largest unused index <= x" is a union-find over the
  // ascending list (slot k stands for g[k-1]; slot 0 means none).
  std::vector<uint32_t> g;
  std::vector<uint32_t> up;
  auto find = [&](uint32_t x) {
    while (up[x] != x) {
      up[x] = up[up[x]];
      x = up[x];
    }
    return x;
  };
  for (int j = J; j >= 0; --j) {
    if (j != J) prog.push_back(0);
    g.swap(lvl[j]);
    const uint32_t n = (uint32_t)g.size();
    up.resize(n + 1);
    for (uint32_t k = 0; k <= n; ++k) up[k] = k;
    uint32_t lo = 0;  // smallest index that may be unused
    for (;;) {
      uint32_t k = find(n);
      if (k == 0) break;
      up[k] = k - 1;
      uint64_t cur = g[k - 1];
      for (;;) {
        const uint64_t lim = (bound - 1) / cur;  // cur * q < bound
        while (lo < n && up[lo + 1] != lo + 1) ++lo;
        if (lo == n || lim < g[lo]) break;  // not even the smallest unused prime fits
        const uint32_t pos =
            lim >= g[n - 1] ? n : (uint32_t)(std::upper_bound(g.begin(), g.begin() + k, lim) - g.begin());
        k = find(pos);
        if (k == 0) break;
        up[k] = k - 1;
        cur *= g[k - 1];
      }
      prog.push_back(cur);
    }
  }
  return prog;
}

extern "C" int fs_select_path(uint64_t pmax) {
  if (pmax <= ((uint64_t)1 << 53)) return FS_PATH_FP64;
  if (pmax <= FS_HYBRID_PMAX + 1) return FS_PATH_HYBRID;  // every p < pmax is covered
  return FS_PATH_MONT;
}

// The setup program of (nmin, path) in one build: its term count, and its
// words (the layout fs_setup_program writes) when `out` is given.
static int64_t build_program(uint32_t nmin, int path, std::vector<uint64_t>* out);

extern "C" int64_t fs_setup_program(uint32_t nmin, int path, uint64_t* words) {
  if (!words) return build_program(nmin, path, nullptr);
  std::vector<uint64_t> w;
  const int64_t T = build_program(nmin, path, &w);
  if (T < 0) return T;
  // exactly the documented words: T (FP64), T + 2 (HYBRID), 2T (MONT)
  const size_t len = path == FS_PATH_MONT ? 2 * (size_t)T : path == FS_PATH_HYBRID ? (size_t)T + 2 : (size_t)T;
  std::copy(w.begin(), w.begin() + len, words);
  return T;
}

static int64_t build_program(uint32_t nmin, int path, std::vector<uint64_t>* out) {
  uint64_t* words = nullptr;
  if (path == FS_PATH_MONT) {
    const int64_t m = fs_power_table(nmin, nullptr, nullptr, nullptr);
    if (out) {
      out->assign((size_t)(2 * m) + 2, 0);
      words = out->data();
      std::vector<uint64_t> b(m);
      std::vector<uint32_t> e(m), c(m);
      fs_power_table(nmin, b.data(), e.data(), c.data());
      for (int64_t i = 0; i < m; ++i) {
        words[i] = b[i];
        words[m + i] = (uint64_t)e[i] | ((uint64_t)c[i] << 32);
      }
    }
    return m;
  }
  if (path != FS_PATH_FP64 && path != FS_PATH_HYBRID) {
    fs_set_error_msg("fs_setup_program: unknown path");
    return FS_E_ARG;
  }
  // FP64: multipliers below 2^50 (the FP64 step's bound); HYBRID: below 2^60, the
  // bound of the two-stage REDC by 2^31 (fs_redc62_mul).
  const uint64_t bound = (path == FS_PATH_FP64) ? ((uint64_t)1 << 50) : ((uint64_t)1 << 60);
  const std::vector<uint64_t> horner = fs_horner_program(nmin, bound);
  // Blocks of 4 terms: a square is a flag (bit 63) on the first term of a
  // block, so each group C_j is padded with multipliers 1 to a whole block.
  std::vector<uint64_t> prog;
  bool pending_square = false;
  for (size_t i = 0; i <= horner.size(); ++i) {
    if (i == horner.size() || horner[i] == 0) {
      while (prog.size() % 4) prog.push_back(1);
      if (i == horner.size()) break;
      if (pending_square) prog.insert(prog.end(), {FS_PROG_SQUARE | 1, 1, 1, 1});  // empty group
      pending_square = true;
      continue;
    }
    prog.push_back(horner[i] | (pending_square ? FS_PROG_SQUARE : 0));
    pending_square = false;
  }
  if (pending_square) prog.insert(prog.end(), {FS_PROG_SQUARE | 1, 1, 1, 1});
  if (out) {
    out->assign(prog.size() + 2, 0);
    words = out->data();
    // HYBRID: powers of 2^-1 the Montgomery products accumulate; E62 with
    // multiplies by 2^-62 (below 2^64/3), E64 with multiplies by 2^-64 (above);
    // squares are by 2^-64 in both
    uint64_t E62 = 0, E64 = 0;
    for (size_t i = 0; i < prog.size(); ++i) {
      const uint64_t m = prog[i] & ~FS_PROG_SQUARE;
      if (prog[i] & FS_PROG_SQUARE) {
        E62 = 2 * E62 + 64;
        E64 = 2 * E64 + 64;
      }
      E62 += 62;
      E64 += 64;
      if (path == FS_PATH_FP64) {
        const double d = (double)m;  // exact: < 2^50
        uint64_t bits;
        memcpy(&bits, &d, 8);
        words[i] = bits | (prog[i] & FS_PROG_SQUARE);
      } else {
        words[i] = prog[i];
      }
    }
    if (path == FS_PATH_HYBRID) {
      words[prog.size()] = E62;
      words[prog.size() + 1] = E64;
    }
  }
  return (int64_t)prog.size();
}

namespace {
std::mutex g_prog_mu;
std::map<std::pair<uint32_t, int>, std::shared_ptr<const fs_program>> g_prog_cache;
size_t g_prog_words = 0;
}  // namespace

std::shared_ptr<const fs_program> fs_program_cached(uint32_t nmin, int path) {
  std::lock_guard<std::mutex> lk(g_prog_mu);
  auto it = g_prog_cache.find(std::make_pair(nmin, path));
  return it != g_prog_cache.end() ? it->second : nullptr;
}

std::shared_ptr<const fs_program> fs_program_get(uint32_t nmin, int path) {
  std::mutex& mu = g_prog_mu;
  auto& cache = g_prog_cache;
  size_t& cached_words = g_prog_words;
  const auto key = std::make_pair(nmin, path);
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  try {
    auto pr = std::make_shared<fs_program>();
    const int64_t T = build_program(nmin, path, &pr->words);
    if (T < 0) return nullptr;
    pr->terms = (uint32_t)T;
    std::lock_guard<std::mutex> lk(mu);
    // bounded: 4096 programs or 2^25 words (256 MB); holders keep their own reference
    if (cache.size() >= 4096 || cached_words + pr->words.size() > ((size_t)1 << 25)) {
      cache.clear();
      cached_words = 0;
    }
    if (cache.emplace(key, pr).second) cached_words += pr->words.size();
    return pr;
  } catch (const std::bad_alloc&) {
    fs_set_error_msg("setup program: host allocation failed");
    return nullptr;
  }
}

// ============================================================ CPU verifier

// Deterministic primality for odd 64-bit p: strong probable-prime tests to the
// seven bases {2, 325, 9375, 28178, 450775, 9780504, 1795265022} (Jim Sinclair's
// set, as used by isPrime, verifyprime.c:96-132), bases reduced mod p; below
// 4,759,123,141 the three bases {2, 7, 61} suffice (Jaeschke 1993: the least
// strong pseudoprime to all three is 4,759,123,141).
extern "C" int fs_is_prime(uint64_t p) {
  static const uint32_t sinclair[7] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
  static const uint32_t small[3] = {2, 7, 61};
  const bool below = p < 4759123141ull;
  const uint32_t* bases = below ? small : sinclair;
  const int nb = below ? 3 : 7;
  if (p < 2) return 0;
  if (p < 4) return 1;
  if ((p & 1) == 0) return 0;
  const uint64_t q = fs_invert(p);
  const uint64_t one = (0 - p) % p;
  const uint64_t r2 = fs_r2_from_one(one, p, q);
  const uint64_t nmo = p - one;
  uint64_t d = p - 1;
  const int t = __builtin_ctzll(d);
  d >>= t;
  for (int i = 0; i < nb; ++i) {
    uint64_t b = bases[i];
    if (b >= p) {
      b %= p;
      if (b == 0) continue;
    }
    const uint64_t bm = fs_mont_mul(b, r2, p, q);
    uint64_t a = one, x = bm;
    for (uint64_t e = d; e; e >>= 1) {
      if (e & 1) a = fs_mont_mul(a, x, p, q);
      x = fs_mont_mul(x, x, p, q);
    }
    if (a == one || a == nmo) continue;
    bool ok = false;
    for (int s = 1; s < t && !ok; ++s) {
      a = fs_mont_mul(a, a, p, q);
      ok = (a == nmo);
    }
    if (!ok) return 0;
  }
  return 1;
}

// O(n) recomputation of n! mod p; true iff n! = +1 (c = -1) or -1 (c = +1)
// (verify, verifyprime.c:137-164).
// a*b*2^-64 mod p for a, b < p, odd p (q = p^-1 mod 2^64), with the host's
// 64x64->128 multiply: m*p matches a*b in the low word, so the high words'
// difference is the exact quotient of a*b - m*p by 2^64, in (-p, p).
static inline uint64_t mont_host(uint64_t a, uint64_t b, uint64_t p, uint64_t q) {
  const u128 t = (u128)a * b;
  const uint64_t m = (uint64_t)t * q;
  const uint64_t mp = (uint64_t)(((u128)m * p) >> 64), hi = (uint64_t)(t >> 64);
  return hi >= mp ? hi - mp : hi - mp + p;
}

// lo * (lo+1) * ... * hi mod p for odd p >= 3 and 2 <= lo <= hi < p (the
// verify loop of verifyprime.c:137-164): four interleaved chains of
// Montgomery products, chain j taking the multipliers lo + j + 4k, each kept in
// Montgomery form and advanced by 4 with a modular add.
static uint64_t range_product_mod(uint64_t p, uint64_t lo, uint64_t hi) {
  const uint64_t q = fs_invert(p), one = (0 - p) % p, r2 = fs_r2_from_one(one, p, q);
  const uint64_t four = fs_add_mod(fs_add_mod(one, one, p), fs_add_mod(one, one, p), p);
  uint64_t r[4] = {one, one, one, one}, im[4];
  for (int j = 0; j < 4; ++j) im[j] = mont_host((lo + j) % p, r2, p, q);
  const uint64_t count = hi - lo + 1;  // hi < p, so no wrap
  for (uint64_t k = count / 4; k; --k) {
    for (int j = 0; j < 4; ++j) {
      r[j] = mont_host(r[j], im[j], p, q);
      im[j] = fs_add_mod(im[j], four, p);
    }
  }
  for (uint64_t j = 0; j < count % 4; ++j) r[0] = mont_host(r[0], im[j], p, q);  // the last 0..3 multipliers
  const uint64_t prod = mont_host(mont_host(r[0], r[1], p, q), mont_host(r[2], r[3], p, q), p, q);
  return mont_host(prod, 1, p, q);
}

// n! mod p (0 when p <= n); p odd, >= 3.
static uint64_t factorial_mod(uint64_t p, uint32_t n) {
  if (n >= p) return 0;
  return n < 2 ? 1 : range_product_mod(p, 2, n);
}

static bool plus_minus_one(uint64_t r, uint64_t p, int32_t c) { return c == -1 ? r == 1 : r == p - 1; }

// An even p never divides n! +- 1 (odd for n >= 2; 2 for n < 2 and c = +1,
// which p > 2 cannot divide).
extern "C" int fs_verify_factor(uint64_t p, uint32_t n, int32_t c) {
  if (p < 3 || (p & 1) == 0 || (c != 1 && c != -1)) return 0;
  return plus_minus_one(factorial_mod(p, n), p, c);
}

// Number parsing of the -p/-P/-n/-N/-v options with the suffixes of
// parse_uint64 (putil.c:33-115): K/M/G/T/P = x1000^k, e<k> = x10^k,
// k/m/g/t/p = x1024^k, b<k> = x2^k.  0 ok, -1 cannot parse, -2 out of range.
extern "C" int fs_parse_uint64(uint64_t* result, const char* str, uint64_t lo, uint64_t hi) {
  char* tail = nullptr;
  errno = 0;
  uint64_t v = strtoull(str, &tail, 0);
  if (errno != 0 || v > hi) return -2;
  const char suf = *tail;
  auto scale = [&](uint64_t f, uint32_t times) -> int {
    for (uint32_t k = 0; k < times; ++k) {
      if (v > hi / f) return -2;
      v *= f;
    }
    return 0;
  };
  switch (suf) {
    case '\0':
      break;
    case 'K': case 'M': case 'G': case 'T': case 'P': {
      if (tail[1] != '\0') return -1;
      const char* order = "KMGTP";
      const uint32_t k = (uint32_t)(strchr(order, suf) - order) + 1;
      if (scale(1000, k)) return -2;
      break;
    }
    case 'k': case 'm': case 'g': case 't': case 'p': {
      if (tail[1] != '\0') return -1;
      const char* order = "kmgtp";
      const uint32_t sh = 10 * ((uint32_t)(strchr(order, suf) - order) + 1);
      if (v > (hi >> sh)) return -2;
      v <<= sh;
      break;
    }
    case 'e': case 'E': case 'b': case 'B': {
      char* t2 = nullptr;
      const unsigned long k = strtoul(tail + 1, &t2, 0);
      if (errno != 0) return -2;
      if (*t2 != '\0') return -1;
      if (suf == 'e' || suf == 'E') {
        if (scale(10, (uint32_t)k)) return -2;
      } else {
        for (unsigned long j = 0; j < k; ++j) {
          if (v > (hi >> 1)) return -2;
          v <<= 1;
        }
      }
      break;
    }
    default:
      return -1;
  }
  if (v < lo) return -2;
  *result = v;
  return 0;
}

// ============================================================ driver

extern "C" void fs_default_options(fs_options* o) {
  memset(o, 0, sizeof *o);
  o->path = FS_PATH_AUTO;
  o->verify_threads = 2;
  o->write_files = 0;
  o->resume = 0;
  o->quiet = 1;
  o->checkpoint_secs = 60;
}

// Checkpoint record, byte-identical to workStatus (cl_sieve.h:9-12): 8 x u64 then
// 2 x u32, 72 bytes, state_sum = wrapping sum of the other fields.
typedef fs_work_status WorkStatus;
static_assert(sizeof(WorkStatus) == 72, "workStatus layout");

static uint64_t state_sum(const WorkStatus& s) {
  return s.pmin + s.pmax + s.p + s.checksum + s.primecount + s.factorcount + s.last_trickle + s.nmin +
         s.nmax;
}

static std::string in_dir(const char* dir, const char* name) {
  if (!dir || !*dir) return name;
  std::string s = dir;
  if (s.back() != '/') s += '/';
  return s + name;
}

// write_state (cl_sieve.cpp:100-129): stateA.ckp / stateB.ckp.
static bool write_state_file(WorkStatus& st, bool file_a, const char* dir) {
  st.state_sum = state_sum(st);
  const std::string fn = in_dir(dir, file_a ? "stateA.ckp" : "stateB.ckp");
  FILE* f = fopen(fn.c_str(), "wb");
  if (!f) {
    fprintf(stderr, "Cannot open %s !!!\n", fn.c_str());
    return false;
  }
  if (fwrite(&st, sizeof st, 1, f) != 1) {
    fprintf(stderr, "Cannot write checkpoint to file. Continuing...\n");
    fclose(f);
    return false;
  }
  return fclose(f) == 0;
}

// Alternates the two files, as the reference's driver does.
static void write_state(WorkStatus& st, bool& write_a_next, const char* dir) {
  if (write_state_file(st, write_a_next, dir)) write_a_next = !write_a_next;
}

extern "C" int fs_write_checkpoint(const char* dir, fs_work_status* st, int which) {
  if (!st || (which != 0 && which != 1)) return FS_E_ARG;
  if (!write_state_file(*st, which == 0, dir)) {
    fs_set_error_msg("cannot write checkpoint file");
    return FS_E_IO;
  }
  return FS_OK;
}

// read_state (cl_sieve.cpp:132-222): a file is good when it parses, matches the
// ranges and its state_sum checks; with two good files the larger p wins.
static int read_state(WorkStatus& st, bool& write_a_next, const char* dir, bool verbose) {
  WorkStatus s[2];
  bool good[2] = {false, false};
  const char* names[2] = {"stateA.ckp", "stateB.ckp"};
  for (int k = 0; k < 2; ++k) {
    const std::string fn = in_dir(dir, names[k]);
    FILE* f = fopen(fn.c_str(), "rb");
    if (!f) continue;
    if (fread(&s[k], sizeof(WorkStatus), 1, f) != 1) {
      fprintf(stderr, "Cannot parse %s !!!\n", names[k]);
      if (verbose) printf("Cannot parse %s !!!\n", names[k]);
    } else if (s[k].pmin != st.pmin || s[k].pmax != st.pmax || s[k].nmin != st.nmin ||
               s[k].nmax != st.nmax) {
      fprintf(stderr, "Invalid checkpoint file %s !!!\n", names[k]);
      if (verbose) printf("Invalid checkpoint file %s !!!\n", names[k]);
    } else if (state_sum(s[k]) != s[k].state_sum) {
      fprintf(stderr, "Checksum error in %s !!!\n", names[k]);
      if (verbose) printf("Checksum error in %s !!!\n", names[k]);
    } else {
      good[k] = true;
    }
    fclose(f);
  }
  if (good[0] && good[1]) {
    if (s[0].p > s[1].p) good[1] = false;
    else good[0] = false;
  }
  for (int k = 0; k < 2; ++k) {
    if (good[k]) {
      st = s[k];
      write_a_next = (k == 1);
      if (verbose) printf("Resuming from checkpoint in %s\n", names[k]);
      return 1;
    }
  }
  return 0;
}

extern "C" int fs_read_checkpoint(const char* dir, fs_work_status* st, int* which) {
  if (!st) return FS_E_ARG;
  bool write_a_next = true;
  WorkStatus w = *st;
  if (!read_state(w, write_a_next, dir, false)) return 0;
  *st = w;
  if (which) *which = write_a_next ? 1 : 0;  // reading B makes A the next file written
  return 1;
}

// Sort key of the drained factor list: p, then |n| (factorcompare, cl_sieve.cpp:422-436).
static bool factor_less(const fs_factor& a, const fs_factor& b) {
  if (a.p != b.p) return a.p < b.p;
  const int32_t na = a.nc < 0 ? -a.nc : a.nc, nb = b.nc < 0 ? -b.nc : b.nc;
  return na < nb;
}

void fs_sort_factors(fs_factor* f, size_t n) { std::stable_sort(f, f + n, factor_less); }

// Verify every factor on `threads` CPU threads; returns the index of the first
// failure or -1 (the OpenMP loop of cl_sieve.cpp:497-517).
// The reference verifies the factors in an OpenMP loop, one factor per
// iteration (cl_sieve.cpp:491-515).  Here each factor's product 2..n is cut
// into ranges of 2^15 multipliers, the ranges of all factors share the
// threads, and each factor's partial products are then multiplied together:
// a few large-n factors still use every thread.
// Worker threads kept between calls (the role of the OpenMP thread team the
// reference's verify loop runs on, cl_sieve.cpp:491-515): run(t, fn) executes
// fn on the calling thread and on up to t - 1 pool threads.  fn must take its
// work items from a shared counter until none are left, so any number of
// callers finishes the region: run returns once the caller's own fn has
// returned and every helper that entered fn has left it; a helper that comes
// later finds the region closed and does not enter (it may also enter the
// next region's fn more than once, which then finds no items left).  Entry
// and exit are atomics (an entering helper counts itself active before it
// checks that the region is open; the caller closes the region before it
// reads the active count).  Like an OpenMP team under its default wait
// policy, a helper polls for the next region for a while (kSpin) before it
// sleeps on a condition variable, so drains a few hundred microseconds apart
// find their helpers awake.  One region at a time; the threads are joined at
// exit.
namespace {
class WorkerPool {
 public:
  void run(int t, const std::function<void()>& fn) {
    std::lock_guard<std::mutex> region(region_mu_);
    const int helpers = t - 1;
    {
      std::lock_guard<std::mutex> lk(mu_);
      while ((int)threads_.size() < helpers) {
        const int idx = (int)threads_.size();
        threads_.emplace_back([this, idx] { loop(idx); });
      }
    }
    task_.store(&fn);
    want_.store(helpers);
    open_.store(true);
    gen_.fetch_add(1);
    if (sleepers_.load() > 0) {  // a helper counted as sleeping has not yet seen this region
      std::lock_guard<std::mutex> lk(mu_);
      cv_.notify_all();
    }
    fn();
    open_.store(false);
    while (active_.load() != 0) std::this_thread::yield();
  }
  ~WorkerPool() {
    {
      std::lock_guard<std::mutex> lk(mu_);
      stop_.store(true);
    }
    cv_.notify_all();
    for (auto& th : threads_) th.join();
  }

 private:
  static constexpr std::chrono::microseconds kSpin{1000};
  void loop(int idx) {
    uint64_t seen = 0;
    for (;;) {
      const auto t0 = std::chrono::steady_clock::now();
      while (gen_.load() == seen && !stop_.load() && std::chrono::steady_clock::now() - t0 < kSpin)
        std::this_thread::yield();
      if (gen_.load() == seen && !stop_.load()) {
        std::unique_lock<std::mutex> lk(mu_);
        sleepers_.fetch_add(1);
        cv_.wait(lk, [&] { return stop_.load() || gen_.load() != seen; });
        sleepers_.fetch_sub(1);
      }
      if (stop_.load()) return;
      seen = gen_.load();
      if (idx >= want_.load()) continue;
      active_.fetch_add(1);
      if (open_.load()) (*task_.load())();
      active_.fetch_sub(1);
    }
  }
  std::mutex region_mu_, mu_;
  std::condition_variable cv_;
  std::vector<std::thread> threads_;
  std::atomic<const std::function<void()>*> task_{nullptr};
  std::atomic<uint64_t> gen_{0};
  std::atomic<int> want_{0}, active_{0}, sleepers_{0};
  std::atomic<bool> open_{false}, stop_{false};
};

WorkerPool& worker_pool() {
  static WorkerPool pool;
  return pool;
}
}  // namespace

int64_t fs_verify_all(const fs_factor* f, size_t n, int threads, signed char* prime) {
  if (n == 0) return -1;
  struct Range {
    size_t i;
    uint64_t lo, hi;
  };
  // ranges of at most 2^15 multipliers, smaller when the whole list is short
  // (about eight ranges per thread)
  uint64_t work = 0;
  for (size_t i = 0; i < n; ++i) work += (uint64_t)(f[i].nc < 0 ? -(int64_t)f[i].nc : f[i].nc);
  const uint64_t kRange = std::min<uint64_t>(
      1u << 15, std::max<uint64_t>(1u << 11, work / (8 * (uint64_t)std::max(threads, 1)) + 1));
  std::vector<Range> rg;
  std::vector<size_t> first(n + 1);
  std::vector<char> direct(n, 0);  // decided without a product
  for (size_t i = 0; i < n; ++i) {
    first[i] = rg.size();
    const uint64_t p = f[i].p;
    const uint32_t fn = (uint32_t)(f[i].nc < 0 ? -(int64_t)f[i].nc : f[i].nc);
    if (p < 3 || (p & 1) == 0 || fn >= p || fn < 2) {
      direct[i] = 1;
      continue;
    }
    for (uint64_t lo = 2; lo <= fn; lo += kRange) rg.push_back({i, lo, std::min<uint64_t>(lo + kRange - 1, fn)});
  }
  first[n] = rg.size();
  // the primality of each distinct p (the list is sorted by p) goes into the
  // same parallel pass, for the composite filter
  std::vector<size_t> distinct;
  if (prime)
    for (size_t i = 0; i < n; ++i) {
      prime[i] = -1;
      if (i == 0 || f[i].p != f[i - 1].p) distinct.push_back(i);
    }
  const size_t items = rg.size() + distinct.size();
  std::vector<uint64_t> part(rg.size());
  if (threads < 1) threads = 1;
  if ((size_t)threads > items) threads = (int)std::max<size_t>(items, 1);
  std::atomic<size_t> next(0);
  const std::function<void()> job = [&]() {
    for (;;) {
      const size_t t = next.fetch_add(1);
      if (t >= items) break;
      if (t < rg.size()) {
        part[t] = range_product_mod(f[rg[t].i].p, rg[t].lo, rg[t].hi);
      } else {
        const size_t i = distinct[t - rg.size()];
        prime[i] = (signed char)(fs_is_prime(f[i].p) != 0);
      }
    }
  };
  if (threads > 1) worker_pool().run(std::min(threads, 256), job);
  else job();
  for (size_t i = 0; i < n; ++i) {
    const uint64_t p = f[i].p;
    const int32_t fc = f[i].nc < 0 ? -1 : 1;
    bool ok;
    if (direct[i]) {
      const uint32_t fn = (uint32_t)(f[i].nc < 0 ? -(int64_t)f[i].nc : f[i].nc);
      ok = fs_verify_factor(p, fn, fc) != 0;
    } else {
      uint64_t r = 1;
      for (size_t t = first[i]; t < first[i + 1]; ++t) r = (uint64_t)((u128)r * part[t] % p);
      ok = plus_minus_one(r, p, fc);
    }
    if (!ok) return (int64_t)i;
  }
  return -1;
}

extern "C" int64_t fs_verify_factors(const fs_factor* f, uint64_t n, int threads) {
  if (!f && n) {  // FS_E_ARG (-1) would read as "all pass"
    fs_set_error_msg("fs_verify_factors: null factor list");
    return FS_E_VERIFY;
  }
  try {
    return fs_verify_all(f, (size_t)n, threads, nullptr);
  } catch (const std::exception&) {
    fs_set_error_msg("fs_verify_factors: host allocation or thread creation failed");
    return FS_E_NOMEM;
  }
}

// The verifier a drain uses: the GPU (fs_verify_factors_gpu) when the mode asks
// for it, or in auto mode when the factors need more than 2^24 multiplications
// (the CPU loop then takes longer than the launch and transfers); else the CPU
// threads.  Index of the first failure, -1, or an FS_E_* error below -1.
bool fs_verify_uses_gpu(const fs_factor* f, size_t n, int mode) {
  if (mode != FS_VERIFY_AUTO) return mode == FS_VERIFY_GPU;
  uint64_t work = 0;
  for (size_t i = 0; i < n; ++i) {
    const uint64_t fn = (uint64_t)(f[i].nc < 0 ? -(int64_t)f[i].nc : f[i].nc);
    work += fn < f[i].p ? fn : 0;
  }
  return work > ((uint64_t)1 << 24);
}

int64_t fs_verify_select(const fs_factor* f, size_t n, int threads, int mode, fs_stream s, signed char* prime) {
  if (!fs_verify_uses_gpu(f, n, mode)) return fs_verify_all(f, n, threads, prime);
  if (prime)
    for (size_t i = 0; i < n; ++i) prime[i] = -1;
  return fs_verify_factors_gpu(f, n, s);
}

// Filter composites (2-PRP p) and fold kept factors into the checksum
// (cl_sieve.cpp:531-550).  Compacts kept factors to the front; returns count.
size_t fs_filter_factors(fs_factor* f, size_t n, uint64_t* checksum, bool verbose, const signed char* prime) {
  size_t kept = 0;
  uint64_t lastgood = 0;
  for (size_t i = 0; i < n; ++i) {
    const uint64_t fp = f[i].p;
    const uint32_t fn = (uint32_t)(f[i].nc < 0 ? -f[i].nc : f[i].nc);
    const int32_t fc = f[i].nc < 0 ? -1 : 1;
    if (fp == lastgood || (prime && prime[i] >= 0 ? prime[i] != 0 : fs_is_prime(fp) != 0)) {
      lastgood = fp;
      *checksum += (uint32_t)(fn + (uint32_t)fc);
      f[kept++] = f[i];
    } else {
      fprintf(stderr, "discarded 2-PRP factor %llu\n", (unsigned long long)fp);
      if (verbose) printf("discarded 2-PRP factor %llu\n", (unsigned long long)fp);
    }
  }
  return kept;
}

// Upper estimate of candidates in [lo, lo + len): the probable primes have
// density ~1/ln p, bounded here by 1/(ln lo - 1.5) plus slack, and never more
// than the survivors of the 113-sieve (density 0.1188).
uint64_t fs_cand_capacity(uint64_t lo, uint64_t len) {
  const double l = (double)(lo < 3 ? 3 : lo);
  const double sieve_bound = (double)len * 0.1190 + 4096.0;
  const double dens = (log(l) > 4.0) ? 1.0 / (log(l) - 1.5) : 1.0;
  const double est = (double)len * dens * 1.15 + 65536.0;
  const double c = std::min(sieve_bound, est);
  return (uint64_t)std::min(c, 4.0e9);
}

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device buffers, stream, events and pinned read-back words of the plain
// driver.  They outlive a call: a pool keeps one set per concurrently running
// fs_sieve call (per device), so a call reuses the stream, its multiplier
// table and the buffers an earlier call allocated, grows them only when it
// needs more, and re-uploads the setup program and power table only when nmin
// or the path changed.  The pool is never freed (the process exit reclaims it).
struct DevBufs {
  int device = -1;
  fs_cands c{};
  static constexpr size_t kSumOff = 64, kHdr = 128;  // byte offsets of the sums and the records
  static_assert(FS_CTR_WORDS * 4 <= kSumOff && kSumOff + FS_SUM_WORDS * 8 <= kHdr, "header layout");
  uint8_t* blk = nullptr;   // counters | sums | factor records
  uint32_t* ctr = nullptr;
  uint64_t* sum = nullptr;
  fs_factor* fac = nullptr;
  uint32_t fcap = 0;        // this call's factor capacity (the reference's, main.cpp:193)
  uint32_t fac_alloc = 0;   // records allocated (>= fcap)
  void* prog = nullptr;
  size_t prog_alloc = 0;    // bytes
  uint32_t prog_terms = 0;
  uint32_t prog_nmin = 0;
  int prog_path = -1;       // program in prog: (prog_nmin, prog_path), -1 = none
  uint64_t* pt_bases = nullptr;
  uint32_t* pt_exps = nullptr;
  size_t pt_alloc = 0;      // entries
  uint32_t pt_count = 0;
  int64_t pt_nmin = -1;     // power table in pt_*: nmin, -1 = none
  void* verify = nullptr;
  fs_stream s = nullptr;
  fs_event ev[6] = {};
  uint8_t* h_blk = nullptr;   // pinned read-back, laid out as blk
  uint32_t* h_ctr = nullptr;
  uint64_t* h_sum = nullptr;
  // The first factor records, read back with the counters at the end of each
  // segment (one wait covers the segment and its read-back); `staged` holds
  // while nothing has been enqueued on s since that wait.
  static constexpr uint32_t kStagedFactors = 1024;
  fs_factor* h_fac = nullptr;
  bool staged = false;
  // setup programs of the lanes of a job split over several lanes per prime,
  // one after another in lprog (cached by path and the lanes' first n)
  static constexpr int kLanes = FS_LANES;
  void* lprog = nullptr;
  size_t lprog_alloc = 0;  // bytes
  std::vector<uint32_t> lprog_key;
  uint32_t lpoff[kLanes] = {}, lterms[kLanes] = {};
  // side stream of the verify products, and its end event
  hipStream_t vs = nullptr;
  hipEvent_t vev = nullptr;
  int init_side() {
    if (!vs && hipStreamCreateWithFlags(&vs, hipStreamNonBlocking) != hipSuccess) {
      fs_set_error_msg("side stream creation failed");
      return FS_E_HIP;
    }
    if (!vev && hipEventCreateWithFlags(&vev, hipEventDisableTiming) != hipSuccess) {
      fs_set_error_msg("side event creation failed");
      return FS_E_HIP;
    }
    return FS_OK;
  }
  int init() {
    int rc;
    if (hipGetDevice(&device) != hipSuccess) device = 0;
    if ((rc = fs_stream_create(&s))) return rc;
    for (auto& e : ev)
      if ((rc = fs_event_create(&e))) return rc;
    if ((rc = fs_malloc(&verify, fs_verify_bytes())) || (rc = need_factors(kStagedFactors))) return rc;
    if (hipHostMalloc((void**)&h_blk, kHdr + kStagedFactors * sizeof(fs_factor), hipHostMallocDefault) !=
        hipSuccess) {
      fs_set_error_msg("pinned host allocation failed");
      return FS_E_NOMEM;
    }
    h_ctr = (uint32_t*)h_blk;
    h_sum = (uint64_t*)(h_blk + kSumOff);
    h_fac = (fs_factor*)(h_blk + kHdr);
    return FS_OK;
  }
  // factor records for a capacity of `cap`, behind the counters and sums in
  // one allocation (ctr at 0, sum at kSumOff, records at kHdr), so that one
  // copy reads back all three
  int need_factors(uint32_t cap) {
    fcap = cap;
    if (fac_alloc >= cap) return FS_OK;
    fs_free(blk);
    blk = nullptr;
    ctr = nullptr;
    sum = nullptr;
    fac = nullptr;
    fac_alloc = 0;
    int rc = fs_malloc((void**)&blk, kHdr + (size_t)cap * sizeof(fs_factor));
    if (rc) return rc;
    ctr = (uint32_t*)blk;
    sum = (uint64_t*)(blk + kSumOff);
    fac = (fs_factor*)(blk + kHdr);
    fac_alloc = cap;
    return FS_OK;
  }
  // counters, sums and the first factor records to pinned memory on s (one
  // copy), read after the next wait for s
  int stage() {
    staged = false;
    const size_t nrec = std::min(kStagedFactors, fcap);
    if (hipMemcpyAsync(h_blk, blk, kHdr + nrec * sizeof(fs_factor), hipMemcpyDeviceToHost, (hipStream_t)s) !=
        hipSuccess) {
      fs_set_error_msg("counter read-back failed");
      return FS_E_HIP;
    }
    return FS_OK;
  }
  // counters (and sums) to the host through pinned memory, waiting for them
  int read_ctr(uint32_t* out, uint64_t* sums = nullptr) {
    staged = false;
    if (hipMemcpyAsync(h_blk, blk, kHdr, hipMemcpyDeviceToHost, (hipStream_t)s) != hipSuccess) {
      fs_set_error_msg("counter read-back failed");
      return FS_E_HIP;
    }
    int rc = fs_stream_sync(s);
    if (rc) return rc;
    memcpy(out, h_ctr, FS_CTR_WORDS * 4);
    if (sums) memcpy(sums, h_sum, FS_SUM_WORDS * 8);
    return FS_OK;
  }
};

std::mutex g_pool_mu;
std::vector<DevBufs*>* g_pool = new std::vector<DevBufs*>();

// One call's use of a pooled DevBufs; the set goes back to the pool when the
// call ends (its stream drained first, also after an error).
struct Lease {
  DevBufs* d = nullptr;
  int take() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) dev = 0;
    {
      std::lock_guard<std::mutex> lk(g_pool_mu);
      for (size_t k = g_pool->size(); k-- > 0;) {
        if ((*g_pool)[k]->device == dev) {
          d = (*g_pool)[k];
          g_pool->erase(g_pool->begin() + k);
          return FS_OK;
        }
      }
    }
    d = new DevBufs();
    const int rc = d->init();
    if (rc) {
      give_back();  // keeps what it allocated for a later call
      return rc;
    }
    return FS_OK;
  }
  void give_back() {
    if (!d) return;
    if (d->vs) (void)hipStreamSynchronize(d->vs);
    if (d->s) (void)fs_stream_sync(d->s);
    std::lock_guard<std::mutex> lk(g_pool_mu);
    if (d->s) g_pool->push_back(d);  // a set without a stream is never handed out again
    d = nullptr;
  }
  ~Lease() { give_back(); }
};

struct Drain {
  std::vector<fs_factor> kept;  // all kept factors, in drain order (globally sorted)
};

}  // namespace

// getResults (cl_sieve.cpp:439-554): read counters and sums, check the flags,
// read + sort + verify + filter the factors, append them to factors.txt.
// The last drain of a call leaves the counters as they are (clear = false):
// the next call on these buffers clears them before its first segment.
static int drain(DevBufs& d, WorkStatus& st, const fs_options& o, Drain& out, fs_result* res, bool clear = true) {
  uint32_t ctr[FS_CTR_WORDS];
  uint64_t sum[FS_SUM_WORDS];
  int rc;
  const bool staged = d.staged;  // counters, sums and the first records already on the host
  if (staged) {
    memcpy(ctr, d.h_ctr, sizeof ctr);
    memcpy(sum, d.h_sum, sizeof sum);
  } else if ((rc = d.read_ctr(ctr, sum))) {
    return rc;
  }
  st.primecount += sum[0];
  st.checksum += sum[1];
  const bool verbose = !o.quiet;
  if (ctr[FS_CTR_MAXCOUNT] > d.c.cap || ctr[FS_CTR_PRIMEOVF]) {
    fs_set_error_msg("error: gpu prime array overflow");
    return FS_E_OVERFLOW;
  }
  if (ctr[FS_CTR_VALIDATE]) {
    fs_set_error_msg("error: gpu validation failure");
    return FS_E_VALIDATE;
  }
  if (ctr[FS_CTR_POWERTABLE]) {
    fs_set_error_msg("error: power table verification failed");
    return FS_E_POWERTABLE;
  }
  const uint32_t nf = ctr[FS_CTR_NFACTORS];
  if (nf > d.fcap || ctr[FS_CTR_FACTOROVF]) {
    char buf[128];
    snprintf(buf, sizeof buf, "Error: number of results (%u) overflowed array.", nf);
    fs_set_error_msg(buf);
    return FS_E_OVERFLOW;
  }
  if (nf > 0) {
    std::vector<fs_factor> h(nf);
    const uint32_t have = staged ? std::min(nf, DevBufs::kStagedFactors) : 0;
    memcpy(h.data(), d.h_fac, (size_t)have * sizeof(fs_factor));
    if (have < nf && (rc = fs_memcpy_dtoh(h.data() + have, d.fac + have, (size_t)(nf - have) * sizeof(fs_factor),
                                          d.s, 1)))
      return rc;
    fs_sort_factors(h.data(), nf);
    res->raw_factors += nf;
    const char* where = fs_verify_uses_gpu(h.data(), nf, o.verify_device) ? "GPU" : "CPU";
    if (verbose) printf("processing %u factors on %s\n", nf, where);
    std::vector<signed char> prime(nf);
    const int64_t bad = fs_verify_select(h.data(), nf, o.verify_threads, o.verify_device, d.s, prime.data());
    if (bad < -1) return (int)bad;
    if (bad >= 0) {
      char buf[256];
      const fs_factor& f = h[bad];
      snprintf(buf, sizeof buf, "%s factor verification failed!  %llu is not a factor of %u!%+d", where,
               (unsigned long long)f.p, (unsigned)(f.nc < 0 ? -f.nc : f.nc), f.nc < 0 ? -1 : 1);
      fs_set_error_msg(buf);
      fprintf(stderr, "%s\n", buf);
      return FS_E_VERIFY;
    }
    fprintf(stderr, "Verified %u factors.\n", nf);
    if (verbose) printf("Verified %u factors.\n", nf);
    uint64_t ck = 0;
    const size_t kept = fs_filter_factors(h.data(), nf, &ck, verbose, prime.data());
    st.checksum += ck;
    st.factorcount += kept;
    if (o.write_files) {
      const std::string fn = in_dir(o.workdir, "factors.txt");
      FILE* f = fopen(fn.c_str(), "a");
      if (!f) {
        fs_set_error_msg("Cannot open factors.txt !!!");
        return FS_E_IO;
      }
      for (size_t i = 0; i < kept; ++i) {
        const uint32_t fn_ = (uint32_t)(h[i].nc < 0 ? -h[i].nc : h[i].nc);
        if (fprintf(f, "%llu | %u!%+d\n", (unsigned long long)h[i].p, fn_, h[i].nc < 0 ? -1 : 1) < 0) {
          fclose(f);
          fs_set_error_msg("Cannot write to factors.txt !!!");
          return FS_E_IO;
        }
      }
      fclose(f);
    }
    out.kept.insert(out.kept.end(), h.begin(), h.begin() + kept);
  }
  d.staged = false;
  return clear ? fs_clearresult(d.ctr, d.sum, d.s) : FS_OK;
}

// finalizeResults (cl_sieve.cpp:721-771): line-count check, then the checksum trailer.
static int finalize_file(const WorkStatus& st, const fs_options& o) {
  const std::string fn = in_dir(o.workdir, "factors.txt");
  if (st.factorcount) {
    FILE* f = fopen(fn.c_str(), "r");
    if (!f) {
      fs_set_error_msg("Cannot open factors.txt !!!");
      return FS_E_IO;
    }
    char line[256];
    uint64_t lc = 0;
    while (fgets(line, sizeof line, f)) ++lc;
    fclose(f);
    if (lc < st.factorcount) {
      fs_set_error_msg("ERROR: Missing factors in factors.txt !!!");
      return FS_E_IO;
    }
  }
  FILE* f = fopen(fn.c_str(), "a");
  if (!f) {
    fs_set_error_msg("Cannot open factors.txt !!!");
    return FS_E_IO;
  }
  int w;
  if (st.factorcount) w = fprintf(f, "%016llX\n", (unsigned long long)st.checksum);
  else w = fprintf(f, "no factors\n%016llX\n", (unsigned long long)st.checksum);
  fclose(f);
  if (w < 0) {
    fs_set_error_msg("Cannot write to factors.txt !!!");
    return FS_E_IO;
  }
  return FS_OK;
}

int fs_validate_ranges(uint64_t pmin, uint64_t pmax, uint32_t nmin, uint32_t nmax) {
  // setupSearch (cl_sieve.cpp:557-589)
  if (pmin == 0 || pmax == 0) {
    fs_set_error_msg("-p and -P arguments are required");
    return FS_E_ARG;
  }
  if (nmin == 0 || nmax == 0) {
    fs_set_error_msg("-n and -N arguments are required");
    return FS_E_ARG;
  }
  if (nmin > nmax) {
    fs_set_error_msg("nmin <= nmax is required");
    return FS_E_ARG;
  }
  if (pmin > pmax) {
    fs_set_error_msg("pmin <= pmax is required");
    return FS_E_ARG;
  }
  if (pmin < nmin) {
    fs_set_error_msg("pmin must be >= nmin, there are no factors when p <= nmin");
    return FS_E_ARG;
  }
  if (nmax > 0x7fffffffu) {
    fs_set_error_msg("nmax must be < 2^31");
    return FS_E_ARG;
  }
  return FS_OK;
}

uint32_t fs_auto_segment_candidates() {
  static std::atomic<uint32_t> cached{0};  // the compute units of the device (one device per process here)
  uint32_t cus = cached.load();
  if (!cus) {
    if (fs_device_info(nullptr, 0, &cus, nullptr, nullptr) != FS_OK || cus == 0) cus = 256;
    cached.store(cus);
  }
  // 8 waves of 64 lanes on each of the 4 SIMDs of every CU, times 8 rounds
  return cus * 4u * 8u * 64u * 8u;
}

// Candidates per segment of a job: 8 rounds of the GPU's wave slots
// (fs_auto_segment_candidates) times min(8, 2^22 / n-range), so a segment's
// p*n steps stay within 8 rounds x 2^22 (about two seconds of iterate) and the
// segment loop still drains and may checkpoint every few seconds.  Each launch
// ends with its last waves draining the SIMDs, so fewer, longer launches lose
// less (n in [1e5, 2e5), p in [1e12, 1e12 + 1e9): 9 segments -> 2, iterate
// 353 -> 347 ms).  The job's expected candidates (its p-range over ln p) are
// split into segments of equal size.  The segment loop turns a target into a
// p-length of target x (ln p - 1), which holds about (1 - 1/ln p) of the
// target; the target is raised by that factor and 2 % more, so that no sliver
// is left over.
static uint32_t segment_target(uint64_t pmin, uint64_t pmax, uint32_t nmin, uint32_t nmax) {
  const uint64_t base = fs_auto_segment_candidates();
  const uint64_t nr = std::max<uint64_t>(nmax > nmin ? nmax - nmin : 1, 1);
  const uint64_t cap = base * std::min<uint64_t>(8, std::max<uint64_t>(1, ((uint64_t)1 << 22) / nr));
  const double lnp = std::max(log(std::max(0.5 * ((double)pmin + (double)pmax), 3.0)), 2.0);
  const double expect = ((double)pmax - (double)pmin) / lnp;
  const double nseg = std::max(1.0, ceil(expect / (double)cap));
  const double t = std::min(1.02 * expect / nseg * lnp / (lnp - 1.0), (double)cap);
  return (uint32_t)std::max<double>(t, (double)base);
}

// n values where a cheaper iterate loop stops applying to every modulus of
// [plo, phi) (the kernels choose their loop per wave by the launch's last n):
// a launch that ends at such a point runs the n below it with that loop.
// FP64: paired steps while 64 n1 <= 31 p and n1 <= 2^22 (fs_fp_pair_ok, smallest
// p), paired with the exact product while p n1^2 <= 2^53, single steps with the
// exact product while p n1 <= 2^53 (largest p); hybrid: paired steps while
// n1 <= 2^22 (fs_sh_pair_ok, every p of the segment in [2^45, 2^64 - 2^58)),
// then the high-word quotient while n1 <= fs_sh_hi_mmax(p) (p >> 34 below 2^63,
// p >> 39 above; the smallest over the segment is at plo or, past 2^63, at
// max(plo, 2^63)).
static std::vector<uint32_t> loop_boundaries(int path, uint64_t plo, uint64_t phi, uint32_t nmin, uint32_t nmax) {
  std::vector<uint32_t> b;
  auto add = [&](uint64_t n) {
    if (n > nmin && n < nmax) b.push_back((uint32_t)n);
  };
  auto isqrt = [](uint64_t x) {
    uint64_t r = (uint64_t)sqrt((double)x);
    while (r * r > x) --r;
    while ((r + 1) * (r + 1) <= x) ++r;
    return r;
  };
  const uint64_t ptop = phi - 1;  // largest modulus of the segment
  if (path == FS_PATH_FP64 && ptop > 0) {
    const uint64_t pair = std::min<uint64_t>(31 * plo / 64, (uint64_t)1 << 22);
    add(pair);
    add(std::min<uint64_t>(pair, isqrt(((uint64_t)1 << 53) / ptop)));
    add(((uint64_t)1 << 53) / ptop);
  } else if (path == FS_PATH_HYBRID) {
    const bool pairs = fs_sh_pair_ok(plo, 1) && fs_sh_pair_ok(ptop, 1);
    if (pairs) add((uint64_t)1 << 22);
    if (ptop < FS_SH_HI_PMAX) {
      const uint64_t p63 = (uint64_t)1 << 63;
      uint64_t hi = fs_sh_hi_mmax(plo);
      if (ptop >= p63) hi = std::min(hi, fs_sh_hi_mmax(std::max(plo, p63)));
      if (!pairs || hi > ((uint64_t)1 << 22)) add(hi);
    }
  }
  std::sort(b.begin(), b.end());
  b.erase(std::unique(b.begin(), b.end()), b.end());
  return b;
}

// Time the events of the last segment's kernels into the result counters
// (a segment run as lanes records no setup window, its setup being in the one
// launch, and its iterate window starts at the end of candidate generation,
// which the lanes launch follows directly on the stream).
static void account(DevBufs& d, fs_result* r, bool lanes) {
  float ms = 0;
  if (fs_event_elapsed_ms(d.ev[0], d.ev[1], &ms) == FS_OK) r->kernel_ms_primegen += ms;
  if (!lanes && fs_event_elapsed_ms(d.ev[2], d.ev[3], &ms) == FS_OK) r->kernel_ms_setup += ms;
  if (fs_event_elapsed_ms(d.ev[lanes ? 1 : 4], d.ev[5], &ms) == FS_OK) r->kernel_ms_iterate += ms;
}

// Lanes per prime for a job whose one segment holds fewer candidates than half
// the GPU's wave slots (cus x 4 SIMDs x 8 waves x 64 lanes): the n-range is cut
// into L subranges that run side by side on L streams, each from the setup
// program of its own first n, as long as the costliest extra setup is at most
// half a subrange's iterate work (VALU operations per setup term and per step
// of each path).  1 = no split.
// Setup program terms at n, about 0.047 n (FP64, multipliers < 2^50) or 0.037 n
// (hybrid, 60-bit words), and VALU operations per term and per step.
static void lane_costs(int path, double* terms_per_n, double* term_ops, double* step_ops) {
  *terms_per_n = path == FS_PATH_FP64 ? 0.047 : 0.037;
  *term_ops = path == FS_PATH_FP64 ? 6.3 : 18.0;
  *step_ops = path == FS_PATH_FP64 ? 4.0 : 7.2;
}

static uint32_t split_lanes(uint32_t count, uint32_t nmin, uint32_t nmax, int path) {
  const uint64_t slots = fs_auto_segment_candidates() / 8;
  if (count == 0 || 2 * (uint64_t)count > slots || path == FS_PATH_MONT) return 1;
  uint32_t L = (uint32_t)std::min<uint64_t>(DevBufs::kLanes, slots / count);
  L = std::min<uint32_t>(L, (nmax - nmin) / 2048);
  double terms_per_n, term_ops, step_ops;
  lane_costs(path, &terms_per_n, &term_ops, &step_ops);
  for (; L > 1; --L) {
    const double nlast = nmin + (double)(nmax - nmin) * (L - 1) / L;
    if (2.0 * (nlast * terms_per_n + 64) * term_ops <= (double)(nmax - nmin) / L * step_ops) break;
  }
  return std::max<uint32_t>(L, 1);
}

// Lane boundaries nb[0] = nmin < ... < nb[L] = nmax with equal work per lane,
// the setup of nb[k]! (a n per n) plus the lane's steps (c per n):
// a nb[k] + c (nb[k+1] - nb[k]) = W, so nb[k+1] = r nb[k] + W / c with
// r = 1 - a / c, and W follows from nb[L] = nmax.
static void lane_bounds(uint32_t nmin, uint32_t nmax, uint32_t L, int path, uint32_t* nb) {
  double terms_per_n, term_ops, step_ops;
  lane_costs(path, &terms_per_n, &term_ops, &step_ops);
  const double r = 1.0 - terms_per_n * term_ops / step_ops, rl = pow(r, (double)L);
  const double w = (1.0 - r) * ((double)nmax - rl * nmin) / (1.0 - rl);  // W / c
  nb[0] = nmin;
  double x = nmin;
  for (uint32_t k = 1; k < L; ++k) {
    x = r * x + w;
    const uint32_t v = (uint32_t)std::llround(x);
    nb[k] = std::min(std::max(v, nb[k - 1] + 1), nmax - (L - k));  // strictly increasing, inside the range
  }
  nb[L] = nmax;
}

extern "C" int fs_trickle_up_text(double progress, double cputime, double runtime, char* buf, size_t len) {
  if (!buf) return FS_E_ARG;
  const int n = snprintf(buf, len,
                         "<trickle_up>\n"
                         "   <progress>%lf</progress>\n"
                         "   <cputime>%lf</cputime>\n"
                         "   <runtime>%lf</runtime>\n"
                         "</trickle_up>\n",
                         progress, cputime, runtime);
  if (n < 0 || (size_t)n >= len) return FS_E_ARG;
  return n;
}

extern "C" int fs_read_init_data(const char* dir, int* gpu_device_num, char* gpu_type, size_t gpu_type_len) {
  if (!gpu_device_num) return FS_E_ARG;
  FILE* f = fopen(in_dir(dir, "init_data.xml").c_str(), "rb");
  if (!f) return 0;
  std::string text;
  char buf[4096];
  size_t got;
  while ((got = fread(buf, 1, sizeof buf, f)) > 0 && text.size() < (1u << 20)) text.append(buf, got);
  fclose(f);
  auto field = [&](const char* tag, std::string* out) {
    const std::string open = std::string("<") + tag + ">", close = std::string("</") + tag + ">";
    const size_t a = text.find(open);
    if (a == std::string::npos) return false;
    const size_t b = text.find(close, a + open.size());
    if (b == std::string::npos) return false;
    *out = text.substr(a + open.size(), b - a - open.size());
    return true;
  };
  std::string num, type;
  if (gpu_type && gpu_type_len) {
    gpu_type[0] = 0;
    if (field("gpu_type", &type)) snprintf(gpu_type, gpu_type_len, "%s", type.c_str());
  }
  if (!field("gpu_device_num", &num)) {
    fs_set_error_msg("init_data.xml: no <gpu_device_num>");
    return FS_E_ARG;
  }
  size_t a = num.find_first_not_of(" \t\r\n"), b = num.find_last_not_of(" \t\r\n");
  if (a == std::string::npos) {
    fs_set_error_msg("init_data.xml: empty <gpu_device_num>");
    return FS_E_ARG;
  }
  num = num.substr(a, b - a + 1);
  if (num.size() > 2 || num.find_first_not_of("0123456789") != std::string::npos || std::stoi(num) > 63) {
    fs_set_error_msg("init_data.xml: <gpu_device_num> is not a device index 0..63");
    return FS_E_ARG;
  }
  *gpu_device_num = std::stoi(num);
  return 1;
}

namespace {
// The BOINC hooks of one run (all optional).
struct Boinc {
  const fs_boinc_hooks* h;
  double last_fraction = 0.0;
  explicit Boinc(const fs_boinc_hooks* hooks) : h(hooks) {}
  void fraction(double fd) {
    last_fraction = fd;
    if (h && h->fraction_done) h->fraction_done(fd, h->user);
  }
  void begin() {
    if (h && h->begin_critical_section) h->begin_critical_section(h->user);
  }
  void end() {
    if (h && h->end_critical_section) h->end_critical_section(h->user);
  }
  void checkpoint_done() {
    if (h && h->checkpoint_completed) h->checkpoint_completed(h->user);
  }
  // handle_trickle_up (cl_sieve.cpp:47-69): once a day, unless standalone
  void trickle(WorkStatus& st, double run_s) {
    if (!h || h->standalone || !h->send_trickle_up) return;
    const uint64_t now = (uint64_t)time(nullptr);
    if (now - st.last_trickle <= 86400) return;
    st.last_trickle = now;
    struct rusage ru;
    double cpu = 0.0;
    if (getrusage(RUSAGE_SELF, &ru) == 0)
      cpu = (double)ru.ru_utime.tv_sec + ru.ru_utime.tv_usec * 1e-6 + (double)ru.ru_stime.tv_sec +
            ru.ru_stime.tv_usec * 1e-6;
    char msg[512];
    if (fs_trickle_up_text(last_fraction, cpu, run_s + h->starting_elapsed_time, msg, sizeof msg) > 0)
      h->send_trickle_up("factsieve_progress", msg, h->user);
  }
};
}  // namespace

extern "C" int fs_sieve(uint64_t pmin, uint64_t pmax, uint32_t nmin, uint32_t nmax, const fs_options* opt,
                        fs_result* res, fs_factor* factors, uint64_t cap, uint64_t* nfactors) {
  fs_options o;
  if (opt) o = *opt;
  else fs_default_options(&o);
  fs_result r;
  memset(&r, 0, sizeof r);
  r.lanes = 1;
  if (nfactors) *nfactors = 0;
  int rc = fs_validate_ranges(pmin, pmax, nmin, nmax);
  if (rc) return rc;
  const bool verbose = !o.quiet;
  const int path = (o.path == FS_PATH_AUTO) ? fs_select_path(pmax) : o.path;
  if (path != FS_PATH_FP64 && path != FS_PATH_HYBRID && path != FS_PATH_MONT) {
    fs_set_error_msg("unknown arithmetic path");
    return FS_E_ARG;
  }
  if ((path == FS_PATH_FP64 && pmax > ((uint64_t)1 << 53)) ||
      (path == FS_PATH_HYBRID && pmax > FS_HYBRID_PMAX + 1)) {
    fs_set_error_msg("forced arithmetic path does not cover pmax");
    return FS_E_ARG;
  }
  r.path = path;
  // The setup program, when not cached, is built on another host thread while
  // the checkpoint state, the stream and the buffers are prepared.
  const std::shared_ptr<const fs_program> prog_ready = fs_program_cached(nmin, path);
  std::future<std::shared_ptr<const fs_program>> prog_fut;
  if (!prog_ready) {
    try {
      prog_fut = std::async(std::launch::async, [nmin, path] { return fs_program_get(nmin, path); });
    } catch (const std::system_error&) {
      prog_fut = std::async(std::launch::deferred, [nmin, path] { return fs_program_get(nmin, path); });
    }
  }

  WorkStatus st;
  memset(&st, 0, sizeof st);
  st.pmin = pmin;
  st.pmax = pmax;
  st.nmin = nmin;
  st.nmax = nmax;
  st.p = pmin;
  bool write_a_next = true;
  const double t0 = now_s();
  fprintf(stderr, "Starting sieve at p: %llu n: %u\nStopping sieve at P: %llu N: %u\n",
          (unsigned long long)pmin, nmin, (unsigned long long)pmax, nmax);
  if (verbose)
    printf("Starting sieve at p: %llu n: %u\nStopping sieve at P: %llu N: %u\n", (unsigned long long)pmin,
           nmin, (unsigned long long)pmax, nmax);

  if (o.write_files) {
    bool resumed = false;
    if (!o.test_mode && o.resume && read_state(st, write_a_next, o.workdir, verbose)) {
      resumed = true;
      r.resumed = 1;
      fprintf(stderr, "Resuming from checkpoint, current p: %llu\n", (unsigned long long)st.p);
      if (verbose) printf("Current p: %llu\n", (unsigned long long)st.p);
      if (st.p == st.pmax) {
        if (verbose) printf("Workunit complete.\n");
        fprintf(stderr, "Workunit complete.\n");
        r.factorcount = st.factorcount;
        r.primecount = st.primecount;
        r.checksum = st.checksum;
        if (res) *res = r;
        return FS_OK;
      }
    }
    if (!resumed) {
      FILE* f = fopen(in_dir(o.workdir, "factors.txt").c_str(), "w");
      if (!f) {
        fs_set_error_msg("Cannot open factors.txt !!!");
        return FS_E_IO;
      }
      fclose(f);
      st.last_trickle = (uint64_t)time(nullptr);
    }
  }

  Lease lease;
  if ((rc = lease.take())) return rc;
  DevBufs& d = *lease.d;
  Drain dr;
  // factor buffer: 1e6 records, 3e7 below 2^32 (main.cpp:193, cl_sieve.cpp:593-595)
  if ((rc = d.need_factors((pmin < 0xFFFFFFFFull) ? 30000000u : 1000000u))) return rc;
  d.staged = false;
  bool clear_all = true;  // the first segment's clear also zeroes the flags and sums (clearresult)

  // setup program and the reference power table (verifypow, and setup on MONT),
  // uploaded when they differ from what the buffers hold
  {
    const std::shared_ptr<const fs_program> pr = prog_ready ? prog_ready : prog_fut.get();
    if (!pr) return FS_E_NOMEM;
    if (d.prog_path != path || d.prog_nmin != nmin) {
      const std::vector<uint64_t>& w = pr->words;
      const size_t bytes = std::max<size_t>(w.size(), 1) * 8;
      if (bytes > d.prog_alloc) {
        fs_free(d.prog);
        d.prog = nullptr;
        d.prog_alloc = 0;
        if ((rc = fs_malloc(&d.prog, bytes))) return rc;
        d.prog_alloc = bytes;
      }
      d.prog_path = -1;
      if ((rc = fs_memcpy_htod(d.prog, w.data(), w.size() * 8, d.s, 1))) return rc;
      d.prog_terms = pr->terms;
      d.prog_nmin = nmin;
      d.prog_path = path;
    }
    if (d.pt_nmin != (int64_t)nmin) {
      const int64_t m = fs_power_table(nmin, nullptr, nullptr, nullptr);
      std::vector<uint64_t> b((size_t)m + 1);
      std::vector<uint32_t> e((size_t)m + 1), c((size_t)m + 1);
      fs_power_table(nmin, b.data(), e.data(), c.data());
      if (b.size() > d.pt_alloc) {
        fs_free(d.pt_bases);
        fs_free(d.pt_exps);
        d.pt_bases = nullptr;
        d.pt_exps = nullptr;
        d.pt_alloc = 0;
        if ((rc = fs_malloc((void**)&d.pt_bases, b.size() * 8)) || (rc = fs_malloc((void**)&d.pt_exps, e.size() * 4)))
          return rc;
        d.pt_alloc = b.size();
      }
      d.pt_nmin = -1;
      if ((rc = fs_memcpy_htod(d.pt_bases, b.data(), b.size() * 8, d.s, 0)) ||
          (rc = fs_memcpy_htod(d.pt_exps, e.data(), e.size() * 4, d.s, 1)))
        return rc;
      d.pt_count = (uint32_t)m;
      d.pt_nmin = nmin;
    }
    fprintf(stderr, "Compressed %zu power table terms to %u\n", fs_primes_upto(nmin).size(), d.pt_count);
    if (verbose) printf("Compressed %zu power table terms to %u\n", fs_primes_upto(nmin).size(), d.pt_count);
  }

  const uint32_t target = o.seg_candidates ? o.seg_candidates : segment_target(pmin, pmax, nmin, nmax);
  const uint32_t nchunk_max = o.n_chunk ? o.n_chunk : (1u << 22);
  bool verified = false, verify_pending = false;
  uint64_t cand_floor = 0;  // room for candidates at least this large after an overflow
  int overflows = 0;
  double last_ckpt = now_s(), last_prog = now_s();
  Boinc bx(o.boinc);
  const double irsize = 1.0 / (double)(pmax - pmin ? pmax - pmin : 1);
  uint32_t prog_chunk = 1u << 16;

  while (st.p < st.pmax) {
    // segment [st.p, stop)
    const double lnp = log((double)(st.p < 3 ? 3 : st.p));
    double len_d = (double)target * std::max(lnp - 1.0, 1.0);
    len_d = std::max(len_d, 65536.0);
    uint64_t len = (uint64_t)std::min(len_d, 4.0e12);
    uint64_t stop = st.p + len;
    if (stop > st.pmax || stop < st.p) stop = st.pmax;
    len = stop - st.p;
    // Room for the segment's candidates from an upper estimate (the count is
    // not read back: every kernel reads it on the device); a segment whose
    // count overflows it is redone, with room for that count, below.
    const uint64_t need = std::min<uint64_t>(std::max<uint64_t>(fs_cand_capacity(st.p, len), cand_floor), 0xFFFFFE00ull);
    // a one-segment job that underfills the GPU runs on several lanes per
    // prime: lane k is a copy of the candidates at k * stride (a multiple of
    // the 256-thread block) covering n in [nb[k], nb[k + 1]), all lanes in one
    // setup and one iterate launch; the lane count follows the expected
    // number of candidates, len / ln p
    const bool whole = (st.p == pmin && stop == st.pmax);
    const double expect = (double)len / std::max(log((double)st.p + 0.5 * (double)len), 1.0);
    const uint32_t L = (whole && o.n_chunk == 0) ? split_lanes((uint32_t)std::min(expect, 4.0e9), nmin, nmax, path) : 1;
    const uint32_t stride = (uint32_t)((need + 255u) & ~(uint64_t)255u);
    const uint64_t room = L > 1 ? (uint64_t)L * stride : need;
    if (room > d.c.cap) {
      fs_cands_free(&d.c);
      if ((rc = fs_cands_alloc(&d.c, (uint32_t)std::min<uint64_t>(room + room / 4, 0xFFFFFF00ull)))) return rc;
      if (room > d.c.cap) {
        fs_set_error_msg("error: gpu prime array overflow");
        return FS_E_OVERFLOW;
      }
    }
    // getsegprimes writes at most `need` candidates (lane 0's stride with
    // lanes): the kernels' grids cover that many, and a larger count raises
    // the overflow flag instead
    d.c.grid_hint = (uint32_t)need;
    fs_cands gen = d.c;
    gen.cap = L > 1 ? stride : (uint32_t)need;
    if ((rc = clear_all ? fs_clearall(d.ctr, d.sum, d.s) : fs_clearn(d.ctr, d.s)) ||
        (rc = fs_event_record(d.ev[0], d.s)) ||
        (rc = fs_getsegprimes_lanes(st.p, stop, &gen, d.ctr, L, stride, d.s)) ||
        (rc = fs_event_record(d.ev[1], d.s)))
      return rc;
    clear_all = false;
    fs_lane_set lanes{};
    lanes.lanes = L;
    lanes.stride = stride;
    lane_bounds(nmin, nmax, L, path, lanes.nb);
    fs_cands first = gen;  // lane 0, and the candidates the checks read
    fs_cands top = first;  // the last lane: its residues are nmax! at the end
    top.p += (size_t)(L - 1) * stride;
    top.q += (size_t)(L - 1) * stride;
    top.r2 += (size_t)(L - 1) * stride;
    top.res += (size_t)(L - 1) * stride;
    top.nnext += (size_t)(L - 1) * stride;
    if (L > 1) {
      std::vector<uint32_t> key{(uint32_t)path, L};
      key.insert(key.end(), lanes.nb, lanes.nb + L);
      if (key != d.lprog_key) {
        std::vector<uint64_t> all;
        for (uint32_t k = 0; k < L; ++k) {
          const std::shared_ptr<const fs_program> pr = fs_program_get(lanes.nb[k], path);
          if (!pr) return FS_E_NOMEM;
          d.lpoff[k] = (uint32_t)all.size();
          d.lterms[k] = pr->terms;
          all.insert(all.end(), pr->words.begin(), pr->words.end());
          all.resize((all.size() + 3) & ~(size_t)3, 0);  // the next program starts on a 4-word block
        }
        const size_t bytes = std::max<size_t>(all.size(), 4) * 8;
        if (bytes > d.lprog_alloc) {
          fs_free(d.lprog);
          d.lprog = nullptr;
          d.lprog_alloc = 0;
          if ((rc = fs_malloc(&d.lprog, bytes))) return rc;
          d.lprog_alloc = bytes;
        }
        d.lprog_key.clear();
        if ((rc = fs_memcpy_htod(d.lprog, all.data(), all.size() * 8, d.s, 1))) return rc;
        d.lprog_key = key;
      }
      lanes.progs = d.lprog;
      for (uint32_t k = 0; k < L; ++k) {
        lanes.poff[k] = d.lpoff[k];
        lanes.terms[k] = d.lterms[k];
      }
    }

    // power-table / setup cross-check on the first candidate of the first
    // non-empty segment (verify kernels, cl_sieve.cpp:868-907): the products
    // run on a side stream (they read only p, q and r2, written by
    // getsegprimes, which the side stream waits for), enqueued after the
    // segment's setup; the comparison with the setup residue follows setup
    const bool verify_now = !verified;
    auto verify_products = [&]() -> int {
      if (!verify_now) return FS_OK;
      int e;
      if ((e = d.init_side())) return e;
      const fs_stream vs = (fs_stream)d.vs;
      if (hipStreamWaitEvent(d.vs, (hipEvent_t)d.ev[1], 0) != hipSuccess) return FS_E_HIP;
      if ((e = fs_verifyslow(&first, d.ctr, d.verify, nmin, vs)) ||
          (e = fs_verifypow(&first, d.ctr, d.pt_bases, d.pt_exps, d.pt_count, d.verify, vs)) ||
          (e = fs_verifyreduce(&first, d.ctr, d.verify, vs)))
        return e;
      if (hipEventRecord(d.vev, d.vs) != hipSuccess) return FS_E_HIP;
      return hipStreamWaitEvent((hipStream_t)d.s, d.vev, 0) == hipSuccess ? FS_OK : FS_E_HIP;
    };
    const std::vector<uint32_t> cuts = loop_boundaries(path, st.p, stop, nmin, nmax);
    if (L > 1) {
      // setup and iterate of every lane in one launch, timed from the end of
      // candidate generation (an event record between two kernels of a stream
      // also delays the second on the GPU, by a few microseconds each)
      if ((rc = fs_lanes(&d.c, d.ctr, d.fac, d.fcap, &lanes, cuts.data(), (uint32_t)cuts.size(), d.verify, path,
                         d.s)))
        return rc;
      ++r.iterate_launches;
      if ((rc = fs_event_record(d.ev[5], d.s))) return rc;
      if ((rc = verify_products())) return rc;
      if (verify_now && (rc = fs_verifyresult_res0(&first, d.ctr, d.verify, d.s))) return rc;
    } else {
      // setup: n! mod p, in chunks of the program
      if ((rc = fs_event_record(d.ev[2], d.s))) return rc;
      if (d.prog_terms == 0 && (rc = fs_setup(&first, d.ctr, d.prog, 0, 0, 0, nmin, path, d.s))) return rc;
      for (uint32_t s0 = 0; s0 < d.prog_terms; s0 += prog_chunk) {
        const uint32_t s1 = std::min(d.prog_terms, s0 + prog_chunk);
        if ((rc = fs_setup(&first, d.ctr, d.prog, d.prog_terms, s0, s1, nmin, path, d.s))) return rc;
      }
      if ((rc = fs_event_record(d.ev[3], d.s))) return rc;
      if ((rc = verify_products())) return rc;
      if (verify_now && (rc = fs_verifyresult(&first, d.ctr, d.verify, d.s))) return rc;
      // iterate over the n-range in chunks cut where the loops change
      if ((rc = fs_event_record(d.ev[4], d.s))) return rc;
      size_t ci = 0;
      for (uint32_t n0 = nmin; n0 < nmax;) {
        uint32_t n1 = (nmax - n0 > nchunk_max) ? n0 + nchunk_max : nmax;
        while (ci < cuts.size() && cuts[ci] <= n0) ++ci;
        if (ci < cuts.size() && cuts[ci] < n1) n1 = cuts[ci];
        if ((rc = fs_iterate(&first, d.ctr, d.fac, d.fcap, n0, n1, path, d.s))) return rc;
        ++r.iterate_launches;
        n0 = n1;
      }
      if ((rc = fs_event_record(d.ev[5], d.s))) return rc;
    }
    verify_pending = verify_now;  // its flag is read with the segment's counters, before any result is taken
    // fold the checksum of the last lane
    if ((rc = fs_check(&top, d.ctr, d.sum, nmax, d.s))) return rc;
    if ((rc = d.stage())) return rc;
    bx.begin();  // the wait for the segment's kernels (waitOnEvent, cl_sieve.cpp:237-285)
    rc = fs_stream_sync(d.s);
    bx.end();
    if (rc) return rc;
    d.staged = true;
    const uint32_t count = d.h_ctr[FS_CTR_COUNT];
    if (d.h_ctr[FS_CTR_PRIMEOVF]) {
      // more candidates than room: every kernel of the segment saw none, so
      // nothing of it was taken; redo it with room for the count
      if (++overflows > 3) {
        fs_set_error_msg("error: gpu prime array overflow");
        return FS_E_OVERFLOW;
      }
      cand_floor = std::max<uint64_t>(cand_floor, (uint64_t)count + count / 2 + 65536);
      const uint32_t zero = 0;
      if ((rc = fs_memcpy_htod(d.ctr + FS_CTR_PRIMEOVF, &zero, 4, d.s, 1))) return rc;
      d.staged = false;
      continue;
    }
    for (uint32_t k = 1; k < L; ++k) r.setup_mulmods += (uint64_t)count * lanes.terms[k];
    if (L > 1) r.lanes = L;
    r.steps += (uint64_t)count * (nmax - nmin);
    r.setup_mulmods += (uint64_t)count * d.prog_terms;
    ++r.segments;

    const uint64_t nextp = st.p + len;
    const bool last = (nextp >= st.pmax || nextp < st.p);
    st.p = last ? st.pmax : nextp;
    if (verify_pending && count > 0) {  // the power-table cross-check of the first segment (cl_sieve.cpp:903-907)
      if (d.h_ctr[FS_CTR_POWERTABLE]) {
        fs_set_error_msg("error: power table verification failed");
        return FS_E_POWERTABLE;
      }
      verified = true;
      fprintf(stderr, "Verified power table (%llu bytes) starting sieve...\n", (unsigned long long)d.pt_count * 16);
      if (verbose) printf("Verified power table (%llu bytes) starting sieve...\n", (unsigned long long)d.pt_count * 16);
    }
    verify_pending = false;  // an empty segment checked nothing: the next one checks again
    account(d, &r, L > 1);

    const double tn = now_s();
    if (tn - last_prog > 2.0) {  // fraction done every 2 s (cl_sieve.cpp:1119-1124)
      const double fd = (double)(st.p - pmin) * irsize;
      bx.fraction(fd);
      if (verbose) printf("Sieve Progress: %.1f%%\n", fd * 100.0);
      last_prog = tn;
    }
    const bool stop_now = o.stop_after_segments && r.segments >= o.stop_after_segments && !last;
    if (!last && (tn - last_ckpt > (double)o.checkpoint_secs || stop_now)) {
      bx.begin();  // drain + checkpoint in one critical section (cl_sieve.cpp:1131-1135)
      rc = drain(d, st, o, dr, &r);
      if (rc == FS_OK && o.write_files && !o.test_mode) {
        bx.trickle(st, tn - t0);
        write_state(st, write_a_next, o.workdir);
        if (verbose) printf("Checkpoint, current p: %llu\n", (unsigned long long)st.p);
        bx.checkpoint_done();
      }
      bx.end();
      if (rc) return rc;
      last_ckpt = tn;
    }
    if (stop_now) {
      r.factorcount = st.factorcount;
      r.primecount = st.primecount;
      r.checksum = st.checksum;
      r.seconds = now_s() - t0;
      if (res) *res = r;
      return FS_OK;
    }
  }

  st.p = st.pmax;
  if (clear_all && (rc = fs_clearresult(d.ctr, d.sum, d.s))) return rc;  // no segment ran
  bx.begin();  // the last drain and the results file (cl_sieve.cpp:1249-1256)
  bx.fraction(1.0);
  if (verbose) printf("Sieve Progress: %.1f%%\n", 100.0);
  rc = drain(d, st, o, dr, &r, false);
  if (rc == FS_OK && o.write_files) {
    if (!o.test_mode) {
      bx.trickle(st, now_s() - t0);
      write_state(st, write_a_next, o.workdir);
      bx.checkpoint_done();
    }
    rc = finalize_file(st, o);
  }
  bx.end();
  if (rc) return rc;
  r.seconds = now_s() - t0;
  r.factorcount = st.factorcount;
  r.primecount = st.primecount;
  r.checksum = st.checksum;
  fprintf(stderr, "Sieve complete.\nfactors %llu, prime count %llu\n", (unsigned long long)st.factorcount,
          (unsigned long long)st.primecount);
  if (verbose) {
    printf("Sieve finished in %d sec.\n", (int)r.seconds);
    printf("factors %llu, prime count %llu, checksum %016llX\n", (unsigned long long)st.factorcount,
           (unsigned long long)st.primecount, (unsigned long long)st.checksum);
  }
  if (factors && cap) memcpy(factors, dr.kept.data(), std::min<uint64_t>(cap, dr.kept.size()) * sizeof(fs_factor));
  if (nfactors) *nfactors = dr.kept.size();
  if (res) *res = r;
  return FS_OK;
}

// run_test (cl_sieve.cpp:1273-1391): the five reference ranges and their
// expected factor count, candidate count and checksum.
extern "C" int fs_self_test(const fs_options* opt, int* passed_mask) {
  struct Case {
    uint64_t pmin, pmax;
    uint32_t nmin, nmax;
    uint64_t factors, primes, checksum;
  };
  static const Case cases[5] = {
      {100000000ull, 101000000ull, 1000000u, 2000000u, 1071, 54211, 0x000004F4F744CA97ull},
      {1000000000000ull, 1000010000000ull, 10000u, 2000000u, 3, 361727, 0x05052A2D65F3D735ull},
      {127ull, 100000ull, 127u, 1000000u, 42770, 9566, 0x0000000065C074F0ull},
      {2000000000000ull, 2000010000000ull, 2000000u, 6000000u, 3, 352866, 0x09C9742A908451EBull},
      {1000000000000ull, 1000001000000ull, 100000000u, 110000000u, 3, 36249, 0x0081056DC98DA092ull},
  };
  fs_options o;
  if (opt) o = *opt;
  else fs_default_options(&o);
  o.test_mode = 1;
  o.resume = 0;
  const bool verbose = !o.quiet;
  if (verbose) printf("Beginning self test of 5 ranges.\n");
  const double t0 = now_s();
  int good = 0, mask = 0;
  for (int k = 0; k < 5; ++k) {
    fs_result r;
    const int rc = fs_sieve(cases[k].pmin, cases[k].pmax, cases[k].nmin, cases[k].nmax, &o, &r, nullptr, 0, nullptr);
    const bool ok = rc == FS_OK && r.factorcount == cases[k].factors && r.primecount == cases[k].primes &&
                    r.checksum == cases[k].checksum;
    if (ok) {
      ++good;
      mask |= 1 << k;
    }
    fprintf(stderr, "test case %d %s.\n", k + 1, ok ? "passed" : "failed");
    if (verbose) printf("test case %d %s.\n\n", k + 1, ok ? "passed" : "failed");
  }
  if (verbose) {
    printf(good == 5 ? "All test cases completed successfully!\n" : "Self test FAILED!\n");
    printf("Elapsed time: %d sec.\n", (int)(now_s() - t0));
  }
  fprintf(stderr, good == 5 ? "All test cases completed successfully!\n" : "Self test FAILED!\n");
  if (passed_mask) *passed_mask = mask;
  return good;
}
