// fs_device.hip -- gfx950 kernels of the FactSieve sieve and the C-ABI runtime
// layer that replaces the reference's OpenCL stack (simpleCL.c + raw cl* calls).
//
// Kernels (reference counterparts in kernels/*.cl):
//   k_clearn / k_clearresult     clearn.cl, clearresult.cl
//   k_getsegprimes               getsegprimes.cl  (113-sieve + base-2 SPRP, bounded append)
//   k_setup_{fp64,hybrid,mont}   setup.cl         (nmin! mod p)
//   k_iterate_{fp64,hybrid,mont} iterate.cl       (n! mod p for n in [startN, endN), +-1 capture)
//   k_check                      check.cl         (checksum fold, count, validation)
//   k_verify{slow,pow,reduce,result} verify*.cl   (nmin! cross-check on one candidate)
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
#include <string>
#include <algorithm>
#include <map>
#include <mutex>
#include <vector>
#include "fs_arith.hpp"
#include "fs_kern.hpp"
#include "fs_internal.hpp"
#include "../../include/factsieve.h"

// ===================================================================== errors

static thread_local std::string g_last_error;
static void set_error(const char* what, hipError_t e) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  g_last_error = buf;
}
extern "C" void fs_set_error_msg(const char* msg) { g_last_error = msg ? msg : ""; }
extern "C" void fs_set_hip_error(const char* what, int err) { set_error(what, (hipError_t)err); }
extern "C" const char* fs_last_error(void) { return g_last_error.c_str(); }

#define HIPCHK(call, what)              \
  do {                                  \
    hipError_t e__ = (call);            \
    if (e__ != hipSuccess) {            \
      set_error(what, e__);             \
      return FS_E_HIP;                  \
    }                                   \
  } while (0)

static hipStream_t S(fs_stream s) { return (hipStream_t)s; }

// ================================================================== constants


#define FS_VERIFY_BLOCKS 10000u
#define FS_VERIFY_L2 40u


// ============================================================ device helpers

// Candidates of the segment the kernels work on: the count getsegprimes
// reserved, at most cap; none after a candidate-buffer overflow (the driver
// then redoes the segment with room for the count).
__device__ __forceinline__ uint32_t fs_live(const uint32_t* ctr, uint32_t cap) {
  if (ctr[FS_CTR_PRIMEOVF]) return 0;
  const uint32_t n = ctr[FS_CTR_COUNT];
  return n < cap ? n : cap;
}

// ================================================================== kernels

__global__ void k_clearn(uint32_t* ctr) {
  if (threadIdx.x == 0) ctr[FS_CTR_COUNT] = 0;
}

__global__ void k_clearresult(uint32_t* ctr, uint64_t* sum) {
  const uint32_t t = threadIdx.x;
  if (t < FS_SUM_WORDS) sum[t] = 0;
  if (t >= 1 && t < FS_CTR_WORDS) ctr[t] = 0;
}
// clearn and clearresult in one launch (a job's first segment).
__global__ void k_clearall(uint32_t* ctr, uint64_t* sum) {
  const uint32_t t = threadIdx.x;
  if (t < FS_SUM_WORDS) sum[t] = 0;
  if (t < FS_CTR_WORDS) ctr[t] = 0;
}

__global__ void k_set_flag(uint32_t* ctr, uint32_t idx) {
  if (threadIdx.x == 0) atomicOr(&ctr[idx], 1u);
}

// Segmented candidate generation over [low, high) (getsegprimes.cl:170-239);
// the block routine is fs_primegen_block.  Candidate idx (< c.cap) is stored
// at idx + k * stride for every lane k < lanes (one copy per lane of a job
// run as lanes per prime; lanes = 1 otherwise).
__global__ __launch_bounds__(FS_BLOCK) void k_getsegprimes(uint64_t base, uint64_t high, uint32_t split,
                                                           fs_cands c, uint32_t* ctr, uint32_t lanes,
                                                           uint32_t stride) {
  __shared__ uint64_t s_tab[FS_NMASK];
  __shared__ uint64_t s_mask[FS_BLOCK];
  __shared__ uint32_t s_excl[FS_BLOCK];
  __shared__ uint32_t s_wsum[FS_BLOCK / 64 + 1];
  fs_sieve_table_init(s_tab);
  fs_primegen_block(
      base, high, blockIdx.x / split, blockIdx.x % split, split, s_tab, s_mask, s_excl, s_wsum,
      [&](uint32_t n) { return atomicAdd(&ctr[FS_CTR_COUNT], n); },
      [&](uint32_t idx, uint64_t N, uint64_t q, uint64_t r2) {
        if (idx < c.cap) {
          for (uint32_t k = 0; k < lanes; ++k) {
            const size_t g = (size_t)idx + (size_t)k * stride;
            c.p[g] = N;
            c.q[g] = q;
            c.r2[g] = r2;
            c.res[g] = 0;
            c.nnext[g] = 0;
          }
        } else {
          atomicOr(&ctr[FS_CTR_PRIMEOVF], 1u);
        }
      });
}

// ---- setup: res <- res * prod(program[start..end)), program word 0 = square

__global__ __launch_bounds__(FS_BLOCK) void k_setup_fp64(fs_cands c, const uint32_t* ctr,
                                                         const double* __restrict__ prog, uint32_t start,
                                                         uint32_t end, uint32_t nmin) {
  const uint32_t i = blockIdx.x * FS_BLOCK + threadIdx.x;
  if (i >= fs_live(ctr, c.cap)) return;
  c.res[i] = fs_cand_setup_fp64(c.p[i], c.q[i], c.r2[i], c.res[i], prog, start, end);
  c.nnext[i] = nmin;
}

__global__ __launch_bounds__(FS_BLOCK) void k_setup_hybrid(fs_cands c, const uint32_t* ctr,
                                                           const uint64_t* __restrict__ prog, uint32_t terms,
                                                           uint32_t start, uint32_t end, uint32_t nmin) {
  const uint32_t i = blockIdx.x * FS_BLOCK + threadIdx.x;
  if (i >= fs_live(ctr, c.cap)) return;
  c.res[i] = fs_cand_setup_hybrid(c.p[i], c.q[i], c.r2[i], c.res[i], prog, terms, start, end);
  c.nnext[i] = nmin;
}

__global__ __launch_bounds__(FS_BLOCK) void k_setup_mont(fs_cands c, const uint32_t* ctr,
                                                         const uint64_t* __restrict__ prog, uint32_t cnt,
                                                         uint32_t start, uint32_t end, uint32_t nmin) {
  const uint32_t i = blockIdx.x * FS_BLOCK + threadIdx.x;
  if (i >= fs_live(ctr, c.cap)) return;
  c.res[i] = fs_cand_setup_mont(c.p[i], c.q[i], c.r2[i], c.res[i], prog, cnt, start, end);
  c.nnext[i] = nmin;
}

// ---- iterate: for n in [n0, n1): capture n! = +-1, then multiply by n+1

#define FS_REC(p_)                                                                                   \
  [&](int32_t nc) { fs_append_factor(&ctr[FS_CTR_NFACTORS], &ctr[FS_CTR_FACTOROVF], fac, fcap, p_, nc, 0); }

__global__ __launch_bounds__(FS_BLOCK) void k_iterate_fp64(fs_cands c, uint32_t* ctr, fs_factor* fac,
                                                           uint32_t fcap, uint32_t n0, uint32_t n1,
                                                           const double* __restrict__ nt,
                                                           const double* __restrict__ pt) {
  const uint32_t i = blockIdx.x * FS_BLOCK + threadIdx.x;
  if (i >= fs_live(ctr, c.cap)) return;
  const uint64_t p = c.p[i];
  if (c.nnext[i] != n0) atomicOr(&ctr[FS_CTR_VALIDATE], 1u);
  c.res[i] = fs_cand_iter_fp64<FS_UNROLL_DEFAULT, 16>(p, c.res[i], n0, n1, nt, pt, true, FS_REC(p));
  c.nnext[i] = n1;
}

__global__ __launch_bounds__(FS_BLOCK) void k_iterate_hybrid(fs_cands c, uint32_t* ctr, fs_factor* fac,
                                                             uint32_t fcap, uint32_t n0, uint32_t n1,
                                                             const double* __restrict__ nt,
                                                             const double* __restrict__ pt) {
  const uint32_t i = blockIdx.x * FS_BLOCK + threadIdx.x;
  if (i >= fs_live(ctr, c.cap)) return;
  const uint64_t p = c.p[i];
  if (c.nnext[i] != n0) atomicOr(&ctr[FS_CTR_VALIDATE], 1u);
  c.res[i] = fs_cand_iter_hybrid<FS_UNROLL_DEFAULT, 16>(p, c.res[i], n0, n1, nt, pt, true, FS_REC(p));
  c.nnext[i] = n1;
}

__global__ __launch_bounds__(FS_BLOCK) void k_iterate_mont(fs_cands c, uint32_t* ctr, fs_factor* fac,
                                                           uint32_t fcap, uint32_t n0, uint32_t n1) {
  const uint32_t i = blockIdx.x * FS_BLOCK + threadIdx.x;
  if (i >= fs_live(ctr, c.cap)) return;
  const uint64_t p = c.p[i];
  if (c.nnext[i] != n0) atomicOr(&ctr[FS_CTR_VALIDATE], 1u);
  c.res[i] = fs_cand_iter_mont(p, c.q[i], c.r2[i], c.res[i], n0, n1, FS_REC(p));
  c.nnext[i] = n1;
}

// t[j] = first + j; the pair table holds, for a = first + j, the entry
// (a*(a+1), a+1) at index j >> 1 of half (j & 1): pt[(j & 1) * 2 * half + 2 * (j >> 1)].
__global__ void k_fill_ntab(double* t, double* pt, uint32_t half, uint32_t first, uint32_t count) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < count) {
    const uint64_t a = (uint64_t)first + j;
    t[j] = (double)a;
    double* e = pt + (size_t)(j & 1) * 2 * half + 2 * (size_t)(j >> 1);
    e[0] = (double)(a * (a + 1));  // exact while a < 2^26; only used for a < 2^22
    e[1] = (double)(a + 1);
  }
}

// ---- lanes per prime in one launch: lane k of a segment is the copy of its
// candidates (at most stride of them) at k * stride (stride a multiple of
// FS_BLOCK, so a block lies in one lane and the lane's n-range, program and
// table pointers are uniform over the block), covering n in [nb[k], nb[k + 1]).
struct LaneArgs {
  uint32_t lanes, stride, ncuts;
  uint32_t nb[FS_LANES + 1];
  uint32_t terms[FS_LANES + 1];    // setup program words of each lane (entry FS_LANES unused:
  uint32_t poff[FS_LANES + 1];     // fs_pick reads FS_LANES + 1 entries) and their offsets
  uint32_t cuts[4];                // loop boundaries: iterate launches of one lane are cut there
  uint32_t tbase, thalf;           // first n of the multiplier table, half length of its pair table
};

// This thread's lane (a scalar: the same for the whole block) and global
// index; false when it holds no candidate.
__device__ __forceinline__ bool fs_lane_of(const LaneArgs& a, const uint32_t* ctr, uint32_t* lane, uint32_t* i) {
  const uint32_t b0 = blockIdx.x * FS_BLOCK;
  *lane = __builtin_amdgcn_readfirstlane(b0 / a.stride);
  *i = b0 + threadIdx.x;
  return *lane < a.lanes && *i - *lane * a.stride < fs_live(ctr, a.stride);
}
// Entry `lane` of a per-lane argument array, selected without indexing (so a
// scalar lane keeps the value in scalar registers, as the scalar program and
// table loads need).
template <class T>
__device__ __forceinline__ T fs_pick(const T* v, uint32_t lane) {
  T r = v[0];
#pragma unroll
  for (uint32_t k = 1; k <= FS_LANES; ++k)
    if (lane == k) r = v[k];
  return r;
}

// Every lane: setup (its whole program: the residue becomes nb[k]! mod p),
// then its n-range as one launch per loop interval would run it.  A lane with
// a short program starts iterating while longer ones still set up.  The first
// candidate's setup residue is also written to *res0 for the power-table
// cross-check (the iterate then moves the residue on).
// The paired loops run 8 multipliers per block here (16 in k_iterate_*): their
// two scalar register sets of pairs then take 32 SGPRs instead of 64, which
// keeps the kernel at 80 (FP64) / 90 (hybrid) SGPRs and 8 waves per SIMD
// rather than 106 with spills and 7.  Every live wave of a launch that fills
// most of the GPU once is then resident from the start.
template <int PATH>
__global__ __launch_bounds__(FS_BLOCK) void k_lanes(fs_cands c, uint32_t* ctr, fs_factor* fac, uint32_t fcap, LaneArgs a,
                                                    const uint64_t* __restrict__ progs, const double* __restrict__ ntab,
                                                    const double* __restrict__ pairs, uint64_t* res0) {
  uint32_t lane, i;
  if (!fs_lane_of(a, ctr, &lane, &i)) return;
  const uint64_t p = c.p[i];
  const uint32_t terms = fs_pick(a.terms, lane);
  const uint64_t* prog = progs + fs_pick(a.poff, lane);
  uint64_t s = PATH == FS_PATH_FP64 ? fs_cand_setup_fp64(p, c.q[i], c.r2[i], 0, (const double*)prog, 0, terms)
                                    : fs_cand_setup_hybrid(p, c.q[i], c.r2[i], 0, prog, terms, 0, terms);
  if (i == 0) *res0 = s;
  const uint32_t hi = fs_pick(a.nb, lane + 1);
  uint32_t lo = fs_pick(a.nb, lane);
  while (lo < hi) {
    uint32_t b = hi;
    for (uint32_t k = 0; k < a.ncuts; ++k)
      if (a.cuts[k] > lo && a.cuts[k] < b) b = a.cuts[k];
    const uint32_t j = lo + 1 - a.tbase;
    const double* nt = ntab + j;
    const double* pt = pairs + (size_t)(j & 1) * 2 * a.thalf + 2 * (size_t)(j >> 1);
    if (PATH == FS_PATH_FP64)
      s = fs_cand_iter_fp64<FS_UNROLL_DEFAULT, 8>(p, s, lo, b, nt, pt, true, FS_REC(p));
    else
      s = fs_cand_iter_hybrid<FS_UNROLL_DEFAULT, 8>(p, s, lo, b, nt, pt, true, FS_REC(p));
    lo = b;
  }
  c.res[i] = s;
  c.nnext[i] = hi;
}

// ---- check: checksum fold + count + validation (check.cl:50-109)

__global__ __launch_bounds__(FS_BLOCK) void k_check(fs_cands c, uint32_t* ctr, uint64_t* sum,
                                                    uint32_t nmax) {
  // Grid-stride over the live candidates; the wrapping sum is reduced per
  // block and added with one atomic per block (the reference keeps one slot
  // per workgroup, check.cl:84-98).
  __shared__ uint64_t s_part[FS_BLOCK / 64];
  const uint32_t live = fs_live(ctr, c.cap);
  uint64_t v = 0;
  bool bad = false;
  for (uint32_t i = blockIdx.x * FS_BLOCK + threadIdx.x; i < live; i += gridDim.x * FS_BLOCK) {
    bool ok;
    v += fs_cand_fold(c.p[i], c.q[i], c.r2[i], c.res[i], nmax, ok);
    bad |= !ok || c.nnext[i] != nmax;
  }
  if (bad) atomicOr(&ctr[FS_CTR_VALIDATE], 1u);
  v = fs_wave_sum(v);
  if ((threadIdx.x & 63) == 0) s_part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t t = 0;
    for (uint32_t w = 0; w < FS_BLOCK / 64; ++w) t += s_part[w];
    if (t) atomicAdd((unsigned long long*)&sum[1], (unsigned long long)t);
    if (blockIdx.x == 0) {
      atomicAdd((unsigned long long*)&sum[0], (unsigned long long)live);
      if (!ctr[FS_CTR_PRIMEOVF]) atomicMax(&ctr[FS_CTR_MAXCOUNT], ctr[FS_CTR_COUNT]);
    }
  }
}

// ---- verification of nmin! on the segment's first candidate

struct VerifyPair { uint64_t slow, pow; };

// Block product (Montgomery form) of one value per thread.
__device__ uint64_t fs_block_prod(uint64_t v, uint64_t p, uint64_t q, uint64_t* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t s = FS_BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fs_mont_mul(sh[threadIdx.x], sh[threadIdx.x + s], p, q);
    __syncthreads();
  }
  return sh[0];
}

__global__ __launch_bounds__(FS_BLOCK) void k_verifyslow(fs_cands c, const uint32_t* ctr,
                                                         VerifyPair* ver, uint32_t nmin) {
  __shared__ uint64_t sh[FS_BLOCK];
  if (fs_live(ctr, c.cap) == 0) return;
  const uint64_t p = c.p[0], q = c.q[0], r2 = c.r2[0];
  uint64_t acc = fs_mont_mul(1, r2, p, q);
  const uint32_t gs = gridDim.x * FS_BLOCK;
  for (uint64_t n = 2 + blockIdx.x * FS_BLOCK + threadIdx.x; n <= nmin; n += gs)
    acc = fs_mont_mul(acc, fs_mont_mul(n, r2, p, q), p, q);
  const uint64_t b = fs_block_prod(acc, p, q, sh);
  if (threadIdx.x == 0) ver[blockIdx.x].slow = b;
  // slots of blocks not launched hold 1 (Montgomery form) for verifyreduce
  const uint64_t one = fs_mont_mul(1, r2, p, q);
  for (uint32_t k = gridDim.x + blockIdx.x * FS_BLOCK + threadIdx.x; k < FS_VERIFY_BLOCKS; k += gs) ver[k].slow = one;
}

__global__ __launch_bounds__(FS_BLOCK) void k_verifypow(fs_cands c, const uint32_t* ctr,
                                                        const uint64_t* bases, const uint32_t* exps,
                                                        uint32_t cnt, VerifyPair* ver) {
  __shared__ uint64_t sh[FS_BLOCK];
  if (fs_live(ctr, c.cap) == 0) return;
  const uint64_t p = c.p[0], q = c.q[0], r2 = c.r2[0];
  uint64_t acc = fs_mont_mul(1, r2, p, q);
  const uint32_t gs = gridDim.x * FS_BLOCK;
  for (uint32_t k = blockIdx.x * FS_BLOCK + threadIdx.x; k < cnt; k += gs) {
    const uint64_t b = fs_mont_mul(bases[k], r2, p, q);
    uint64_t a = fs_mont_mul(1, r2, p, q), x = b;
    for (uint32_t e = exps[k]; e; e >>= 1) {  // right-to-left binary power
      if (e & 1) a = fs_mont_mul(a, x, p, q);
      x = fs_mont_mul(x, x, p, q);
    }
    acc = fs_mont_mul(acc, a, p, q);
  }
  const uint64_t b = fs_block_prod(acc, p, q, sh);
  if (threadIdx.x == 0) ver[blockIdx.x].pow = b;
  const uint64_t one = fs_mont_mul(1, r2, p, q);
  for (uint32_t k = gridDim.x + blockIdx.x * FS_BLOCK + threadIdx.x; k < FS_VERIFY_BLOCKS; k += gs) ver[k].pow = one;
}

__global__ __launch_bounds__(FS_BLOCK) void k_verifyreduce(fs_cands c, const uint32_t* ctr,
                                                           VerifyPair* ver) {
  __shared__ uint64_t sh[FS_BLOCK];
  if (fs_live(ctr, c.cap) == 0) return;
  const uint64_t p = c.p[0], q = c.q[0], r2 = c.r2[0];
  const uint64_t one = fs_mont_mul(1, r2, p, q);
  const uint32_t k = blockIdx.x * FS_BLOCK + threadIdx.x;
  const bool in = k < FS_VERIFY_BLOCKS;
  const uint64_t a = fs_block_prod(in ? ver[k].slow : one, p, q, sh);
  __syncthreads();
  const uint64_t b = fs_block_prod(in ? ver[k].pow : one, p, q, sh);
  if (threadIdx.x == 0) {
    ver[FS_VERIFY_BLOCKS + blockIdx.x].slow = a;
    ver[FS_VERIFY_BLOCKS + blockIdx.x].pow = b;
  }
}

// res0: the setup residue of the first candidate, or null for c.res[0].
__global__ __launch_bounds__(FS_BLOCK) void k_verifyresult(fs_cands c, uint32_t* ctr, VerifyPair* ver,
                                                           const uint64_t* res0) {
  __shared__ uint64_t sh[FS_BLOCK];
  if (fs_live(ctr, c.cap) == 0) return;
  const uint64_t p = c.p[0], q = c.q[0], r2 = c.r2[0];
  const uint64_t one = fs_mont_mul(1, r2, p, q);
  const bool in = threadIdx.x < FS_VERIFY_L2;
  const uint64_t a = fs_block_prod(in ? ver[FS_VERIFY_BLOCKS + threadIdx.x].slow : one, p, q, sh);
  __syncthreads();
  const uint64_t b = fs_block_prod(in ? ver[FS_VERIFY_BLOCKS + threadIdx.x].pow : one, p, q, sh);
  if (threadIdx.x == 0) {
    // slow product == power table product == the setup kernel's residue
    const uint64_t plain = fs_mont_mul(a, 1, p, q);
    if (a != b || plain != (res0 ? *res0 : c.res[0])) atomicOr(&ctr[FS_CTR_POWERTABLE], 1u);
  }
}

// ---- factor verification on the device (verify, verifyprime.c:137-164)
//
// The reference re-checks every factor p | n! +- 1 on the CPU with an O(n)
// product (an OpenMP loop, cl_sieve.cpp:491-515).  Here each factor's product
// 2 * 3 * ... * n mod p is cut into ranges of FS_VRANGE consecutive multipliers,
// one range per thread, with integer Montgomery products and a Montgomery
// counter (not the FP64 or hybrid arithmetic of the sieve, so the check stays
// independent of it); a second kernel multiplies each factor's range products.
// first[i] is the index of factor i's first range (first[nf] = the range
// count); factors with no range (n < 2, n >= p, even p, p < 3) are decided on
// the host.

#define FS_VRANGE 4096u

__device__ __forceinline__ uint32_t fs_vrange_owner(const uint32_t* first, uint32_t nf, uint32_t t) {
  uint32_t lo = 0, hi = nf - 1;  // last i with first[i] <= t
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (first[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ __launch_bounds__(FS_BLOCK) void k_verify_ranges(const fs_factor* __restrict__ f,
                                                            const uint32_t* __restrict__ first, uint32_t nf,
                                                            uint32_t nrg, uint64_t* __restrict__ part) {
  const uint32_t t = blockIdx.x * FS_BLOCK + threadIdx.x;
  if (t >= nrg) return;
  const uint32_t i = fs_vrange_owner(first, nf, t);
  const uint64_t p = f[i].p;
  const uint32_t n = (uint32_t)(f[i].nc < 0 ? -(int64_t)f[i].nc : f[i].nc);
  const uint64_t lo = 2 + (uint64_t)(t - first[i]) * FS_VRANGE;
  const uint64_t hi = min(lo + FS_VRANGE - 1, (uint64_t)n);
  const uint64_t q = fs_invert(p), one = fs_r64_mod(p), r2 = fs_r2_from_one(one, p, q);
  uint64_t rm = one, im = fs_mont_mul(lo, r2, p, q);  // Montgomery forms of 1 and lo (lo < p)
  for (uint64_t m = lo; m <= hi; ++m) {
    rm = fs_mont_mul(rm, im, p, q);
    im = fs_add_mod(im, one, p);
  }
  part[t] = fs_mont_mul(rm, 1, p, q);
}

__global__ __launch_bounds__(FS_BLOCK) void k_verify_combine(const fs_factor* __restrict__ f,
                                                             const uint32_t* __restrict__ first, uint32_t nf,
                                                             const uint64_t* __restrict__ part,
                                                             uint64_t* __restrict__ res) {
  const uint32_t i = blockIdx.x * FS_BLOCK + threadIdx.x;
  if (i >= nf) return;
  const uint32_t b = first[i], e = first[i + 1];
  if (b == e) return;
  const uint64_t p = f[i].p;
  const uint64_t q = fs_invert(p), r2 = fs_r2_from_one(fs_r64_mod(p), p, q);
  uint64_t r = part[b];
  for (uint32_t t = b + 1; t < e; ++t) r = fs_mul_mod(r, part[t], p, q, r2);
  res[i] = r;  // n! mod p
}

// ================================================================ C ABI: runtime

static int g_device = -1;
static std::mutex g_init_mu;
static int g_inject = 0;

// Tables of fs_iterate_noalloc: one file-scope device array, a copy on every
// GPU the library runs on (held until the process ends).  Layout as a
// multiplier table of kScrCap entries (NTab below): kScrCap multipliers, then
// the pair table's two halves of kScrCap / 2 + 1 pairs of two doubles.
static constexpr uint32_t kScrCap = FS_ITERATE_NOALLOC_MAX_N + 64;
static constexpr size_t kScrDoubles = (size_t)kScrCap + 4 * ((size_t)kScrCap / 2 + 1);
static __device__ __align__(256) double scratch[kScrDoubles];
// Device address of `scratch` per GPU, from hipGetSymbolAddress with that GPU
// current (the array's name in host code is not a device address).
static constexpr int kMaxDevices = 64;
static double* g_scratch_ptr[kMaxDevices] = {};
static hipError_t g_scratch_err[kMaxDevices] = {};

extern "C" int fs_init(int device_index) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0) {
    g_last_error = "no HIP device available";
    return FS_E_NODEV;
  }
  if (device_index < 0 || device_index >= n) {
    g_last_error = "device index out of range";
    return FS_E_NODEV;
  }
  HIPCHK(hipSetDevice(device_index), "hipSetDevice");
  g_device = device_index;
  if (device_index < kMaxDevices && !g_scratch_ptr[device_index] && g_scratch_err[device_index] == hipSuccess) {
    void* ptr = nullptr;
    const hipError_t se = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (se == hipSuccess && ptr) g_scratch_ptr[device_index] = (double*)ptr;
    else g_scratch_err[device_index] = se != hipSuccess ? se : hipErrorInvalidSymbol;  // reported by the
                                                                                    // entry point that needs it
  }
  return FS_OK;
}

static int ensure_init() {
  if (g_device >= 0) {
    HIPCHK(hipSetDevice(g_device), "hipSetDevice");
    return FS_OK;
  }
  return fs_init(0);
}

extern "C" int fs_device_info(char* name, size_t name_len, uint32_t* cus, uint64_t* total,
                              uint64_t* freeb) {
  int rc = ensure_init();
  if (rc) return rc;
  hipDeviceProp_t pr;
  HIPCHK(hipGetDeviceProperties(&pr, g_device), "hipGetDeviceProperties");
  if (name && name_len) {
    snprintf(name, name_len, "%s (%s)", pr.name, pr.gcnArchName);
  }
  if (cus) *cus = (uint32_t)pr.multiProcessorCount;
  size_t f = 0, t = 0;
  HIPCHK(hipMemGetInfo(&f, &t), "hipMemGetInfo");
  if (total) *total = t;
  if (freeb) *freeb = f;
  return FS_OK;
}

extern "C" int fs_inject_fault(int idx) {
  if (idx != 0 && idx != FS_CTR_POWERTABLE && idx != FS_CTR_VALIDATE && idx != FS_CTR_PRIMEOVF &&
      idx != FS_CTR_FACTOROVF) {
    g_last_error = "fs_inject_fault: unsupported counter";
    return FS_E_ARG;
  }
  g_inject = idx;
  return FS_OK;
}


// The injected fault `idx`, once: 1 (and the injection cleared) when it was set.
int fs_take_injected_fault(int idx) {
  if (g_inject != idx || idx == 0) return 0;
  g_inject = 0;
  return 1;
}

extern "C" int fs_malloc(void** d, size_t bytes) {
  int rc = ensure_init();
  if (rc) return rc;
  hipError_t e = hipMalloc(d, bytes ? bytes : 1);
  if (e != hipSuccess) {
    set_error("hipMalloc", e);
    return FS_E_NOMEM;
  }
  return FS_OK;
}

extern "C" int fs_free(void* d) {
  if (!d) return FS_OK;
  HIPCHK(hipFree(d), "hipFree");
  return FS_OK;
}

extern "C" int fs_memcpy_htod(void* dst, const void* src, size_t bytes, fs_stream s, int blocking) {
  if (!bytes) return FS_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, S(s)), "hipMemcpyAsync(HtoD)");
  if (blocking) HIPCHK(hipStreamSynchronize(S(s)), "hipStreamSynchronize");
  return FS_OK;
}

extern "C" int fs_memcpy_dtoh(void* dst, const void* src, size_t bytes, fs_stream s, int blocking) {
  if (!bytes) return FS_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, S(s)), "hipMemcpyAsync(DtoH)");
  if (blocking) HIPCHK(hipStreamSynchronize(S(s)), "hipStreamSynchronize");
  return FS_OK;
}

extern "C" int fs_memcpy_dtod(void* dst, const void* src, size_t bytes, fs_stream s) {
  if (!bytes) return FS_OK;
  HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, S(s)), "hipMemcpyAsync(DtoD)");
  return FS_OK;
}

extern "C" int fs_memset(void* d, int v, size_t bytes, fs_stream s) {
  if (!bytes) return FS_OK;
  HIPCHK(hipMemsetAsync(d, v, bytes, S(s)), "hipMemsetAsync");
  return FS_OK;
}

extern "C" int fs_cands_alloc(fs_cands* c, uint32_t cap) {
  memset(c, 0, sizeof *c);
  if (cap == 0) {
    g_last_error = "fs_cands_alloc: cap must be > 0";
    return FS_E_ARG;
  }
  int rc;
  if ((rc = fs_malloc((void**)&c->p, (size_t)cap * 8)) || (rc = fs_malloc((void**)&c->q, (size_t)cap * 8)) ||
      (rc = fs_malloc((void**)&c->r2, (size_t)cap * 8)) ||
      (rc = fs_malloc((void**)&c->res, (size_t)cap * 8)) ||
      (rc = fs_malloc((void**)&c->nnext, (size_t)cap * 4))) {
    fs_cands_free(c);
    return rc;
  }
  c->cap = cap;
  return FS_OK;
}

extern "C" int fs_cands_free(fs_cands* c) {
  fs_free(c->p);
  fs_free(c->q);
  fs_free(c->r2);
  fs_free(c->res);
  fs_free(c->nnext);
  memset(c, 0, sizeof *c);
  return FS_OK;
}

extern "C" int fs_stream_create(fs_stream* s) {
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t h;
  HIPCHK(hipStreamCreateWithFlags(&h, hipStreamNonBlocking), "hipStreamCreate");
  *s = (fs_stream)h;
  return FS_OK;
}
static int ntab_release(hipStream_t st);
extern "C" int fs_stream_destroy(fs_stream s) {
  if (!s) return FS_OK;
  // the stream's multiplier table goes with it (its last kernels finish first)
  int rc = ntab_release(S(s));
  HIPCHK(hipStreamDestroy(S(s)), "hipStreamDestroy");
  return rc;
}
extern "C" int fs_stream_sync(fs_stream s) {
  HIPCHK(hipStreamSynchronize(S(s)), "hipStreamSynchronize");
  return FS_OK;
}
extern "C" int fs_event_create(fs_event* e) {
  int rc = ensure_init();
  if (rc) return rc;
  hipEvent_t h;
  HIPCHK(hipEventCreate(&h), "hipEventCreate");
  *e = (fs_event)h;
  return FS_OK;
}
extern "C" int fs_event_destroy(fs_event e) {
  if (e) HIPCHK(hipEventDestroy((hipEvent_t)e), "hipEventDestroy");
  return FS_OK;
}
extern "C" int fs_event_record(fs_event e, fs_stream s) {
  HIPCHK(hipEventRecord((hipEvent_t)e, S(s)), "hipEventRecord");
  return FS_OK;
}
extern "C" int fs_event_query(fs_event e) {
  hipError_t r = hipEventQuery((hipEvent_t)e);
  if (r == hipSuccess) return 1;
  if (r == hipErrorNotReady) return 0;
  set_error("hipEventQuery", r);
  return FS_E_HIP;
}
extern "C" int fs_event_sync(fs_event e) {
  HIPCHK(hipEventSynchronize((hipEvent_t)e), "hipEventSynchronize");
  return FS_OK;
}
extern "C" int fs_event_elapsed_ms(fs_event a, fs_event b, float* ms) {
  HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b), "hipEventElapsedTime");
  return FS_OK;
}
extern "C" int fs_graph_begin(fs_stream s) {
  HIPCHK(hipStreamBeginCapture(S(s), hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
  return FS_OK;
}
extern "C" int fs_graph_end(fs_stream s, fs_graph* g) {
  hipGraph_t graph;
  HIPCHK(hipStreamEndCapture(S(s), &graph), "hipStreamEndCapture");
  hipGraphExec_t ex;
  hipError_t e = hipGraphInstantiate(&ex, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) {
    set_error("hipGraphInstantiate", e);
    return FS_E_HIP;
  }
  *g = (fs_graph)ex;
  return FS_OK;
}
extern "C" int fs_graph_launch(fs_graph g, fs_stream s) {
  HIPCHK(hipGraphLaunch((hipGraphExec_t)g, S(s)), "hipGraphLaunch");
  return FS_OK;
}
extern "C" int fs_graph_destroy(fs_graph g) {
  if (g) HIPCHK(hipGraphExecDestroy((hipGraphExec_t)g), "hipGraphExecDestroy");
  return FS_OK;
}

// ================================================================ C ABI: kernels

static inline uint32_t grid_for(uint32_t n) { return (n + FS_BLOCK - 1) / FS_BLOCK; }
// Workgroups for a per-candidate kernel: the live count when the host knows it.
static inline uint32_t cand_grid(const fs_cands* c) {
  const uint32_t n = (c->grid_hint && c->grid_hint < c->cap) ? c->grid_hint : c->cap;
  return n ? grid_for(n) : 1;
}

#define LAUNCHCHK(what)                              \
  do {                                               \
    hipError_t e__ = hipGetLastError();              \
    if (e__ != hipSuccess) {                         \
      set_error(what, e__);                          \
      return FS_E_HIP;                               \
    }                                                \
  } while (0)

static int check_cands(const fs_cands* c, const char* who) {
  if (!c || !c->p || !c->q || !c->r2 || !c->res || !c->nnext || c->cap == 0) {
    g_last_error = std::string(who) + ": invalid candidate buffers";
    return FS_E_ARG;
  }
  return FS_OK;
}

extern "C" int fs_clearn(uint32_t* d_ctr, fs_stream s) {
  if (!d_ctr) return FS_E_ARG;
  k_clearn<<<1, 64, 0, S(s)>>>(d_ctr);
  LAUNCHCHK("clearn");
  return FS_OK;
}

extern "C" int fs_clearresult(uint32_t* d_ctr, uint64_t* d_sum, fs_stream s) {
  if (!d_ctr || !d_sum) return FS_E_ARG;
  k_clearresult<<<1, 64, 0, S(s)>>>(d_ctr, d_sum);
  LAUNCHCHK("clearresult");
  return FS_OK;
}

int fs_clearall(uint32_t* d_ctr, uint64_t* d_sum, fs_stream s) {
  if (!d_ctr || !d_sum) return FS_E_ARG;
  k_clearall<<<1, 64, 0, S(s)>>>(d_ctr, d_sum);
  LAUNCHCHK("clearall");
  return FS_OK;
}

extern "C" int fs_getsegprimes(uint64_t low, uint64_t high, const fs_cands* c, uint32_t* d_ctr,
                               fs_stream s) {
  return fs_getsegprimes_lanes(low, high, c, d_ctr, 1, 0, s);
}

int fs_getsegprimes_lanes(uint64_t low, uint64_t high, const fs_cands* c, uint32_t* d_ctr, uint32_t lanes,
                          uint32_t stride, fs_stream s) {
  int rc = check_cands(c, "getsegprimes");
  if (rc) return rc;
  if (low > high) {
    g_last_error = "getsegprimes: low > high";
    return FS_E_ARG;
  }
  if (low == high || low == UINT64_MAX) return FS_OK;
  const uint64_t base = low | 1;  // first odd number >= low
  if (base >= high) return FS_OK;
  const uint64_t threads = (high - base + 127) / 128;
  const uint64_t blocks = (threads + FS_BLOCK - 1) / FS_BLOCK;
  if (blocks > 0x7fffffffull) {
    g_last_error = "getsegprimes: range too large for one launch";
    return FS_E_ARG;
  }
  // a range of fewer than 2,048 tiles shares each tile's tests among up to 16
  // blocks, so the chip holds at least ~2,048 workgroups
  const uint32_t split = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(16, 2048 / blocks));
  if (lanes < 1 || (lanes > 1 && (stride < c->cap || (uint64_t)(lanes - 1) * stride + c->cap > 0xFFFFFFFFull))) {
    g_last_error = "getsegprimes: bad lane layout";
    return FS_E_ARG;
  }
  k_getsegprimes<<<(uint32_t)(blocks * split), FS_BLOCK, 0, S(s)>>>(base, high, split, *c, d_ctr, lanes, stride);
  LAUNCHCHK("getsegprimes");
  if (g_inject == FS_CTR_PRIMEOVF) {
    k_set_flag<<<1, 64, 0, S(s)>>>(d_ctr, FS_CTR_PRIMEOVF);
    g_inject = 0;
  }
  return FS_OK;
}

extern "C" int fs_setup(const fs_cands* c, const uint32_t* d_ctr, const void* d_prog,
                        uint32_t prog_terms, uint32_t start, uint32_t end, uint32_t nmin, int path,
                        fs_stream s) {
  int rc = check_cands(c, "setup");
  if (rc) return rc;
  if (end > prog_terms) end = prog_terms;
  if (start > end) start = end;
  if (path != FS_PATH_MONT && ((start | end) & 3)) {
    g_last_error = "setup: FP64/hybrid program chunks must start and end on 4-term blocks";
    return FS_E_ARG;
  }
  if (!d_prog && end > start) {
    g_last_error = "setup: null program";
    return FS_E_ARG;
  }
  const uint32_t g = cand_grid(c);
  switch (path) {
    case FS_PATH_FP64:
      k_setup_fp64<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, (const double*)d_prog, start, end, nmin);
      break;
    case FS_PATH_HYBRID:  // d_prog: prog_terms multipliers, then the 2-exponent word
      k_setup_hybrid<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, (const uint64_t*)d_prog, prog_terms, start, end,
                                              nmin);
      break;
    case FS_PATH_MONT:  // d_prog: bases[prog_terms] then (exp | curbit << 32)[prog_terms]
      k_setup_mont<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, (const uint64_t*)d_prog, prog_terms, start,
                                            end, nmin);
      break;
    default:
      g_last_error = "setup: unknown path";
      return FS_E_ARG;
  }
  LAUNCHCHK("setup");
  return FS_OK;
}

// Multiplier tables of the FP64 iterate kernel, one per stream: nt[j] = base + j
// as doubles, and the pair table of the paired step (k_fill_ntab), refilled on
// the stream (stream order protects earlier launches).
struct NTab {
  double* d = nullptr;   // cap multipliers, then the pair table (2 halves of cap/2 + 1 pairs)
  uint32_t base = 0, len = 0, cap = 0;
  double* pairs() const { return d + cap; }
  uint32_t half() const { return cap / 2 + 1; }
  // pair table position of the pairs (a, a+1), (a+2, a+3), ... for a >= base
  const double* pairs_at(uint32_t a) const {
    const uint32_t j = a - base;
    return pairs() + (size_t)(j & 1) * 2 * half() + 2 * (size_t)(j >> 1);
  }
};
static std::mutex g_ntab_mu;
static std::map<hipStream_t, NTab> g_ntab;

static int ntab_release(hipStream_t st) {
  double* d = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_ntab_mu);
    auto it = g_ntab.find(st);
    if (it == g_ntab.end()) return FS_OK;
    d = it->second.d;
    g_ntab.erase(it);
  }
  if (d) {
    HIPCHK(hipStreamSynchronize(st), "hipStreamSynchronize(ntab)");
    HIPCHK(hipFree(d), "hipFree(ntab)");
  }
  return FS_OK;
}

static int ntab_for(hipStream_t st, uint32_t first, uint32_t count, const double** out,
                    const double** pout = nullptr) {
  std::lock_guard<std::mutex> lk(g_ntab_mu);
  NTab& t = g_ntab[st];
  if (t.d && first >= t.base && (uint64_t)first + count + 16 <= (uint64_t)t.base + t.len) {
    *out = t.d + (first - t.base);
    if (pout) *pout = t.pairs_at(first);
    return FS_OK;
  }
  // 16 entries of padding: the kernels read one unrolled block ahead (8
  // multipliers, or 8 pairs at 16 steps per block)
  const uint32_t want = (count < 65536 ? 65536 : count) + 16;
  if (want > t.cap) {
    if (t.d) HIPCHK(hipFree(t.d), "hipFree(ntab)");
    t.d = nullptr;
    hipError_t e = hipMalloc(&t.d, ((size_t)want + 4 * ((size_t)want / 2 + 1)) * sizeof(double));
    if (e != hipSuccess) {
      set_error("hipMalloc(ntab)", e);
      t.cap = t.len = 0;
      return FS_E_NOMEM;
    }
    t.cap = want;
  }
  const uint32_t len = want;  // values past 2^31 are never used as multipliers
  k_fill_ntab<<<(len + 255) / 256, 256, 0, st>>>(t.d, t.pairs(), t.half(), first, len);
  HIPCHK(hipGetLastError(), "fill ntab");
  t.base = first;
  t.len = len;
  *out = t.d;
  if (pout) *pout = t.pairs_at(first);
  return FS_OK;
}

extern "C" int fs_iterate(const fs_cands* c, uint32_t* d_ctr, fs_factor* d_factor, uint32_t fcap,
                          uint32_t n0, uint32_t n1, int path, fs_stream s) {
  int rc = check_cands(c, "iterate");
  if (rc) return rc;
  if (n0 > n1 || n1 > 0x7fffffffu) {
    g_last_error = "iterate: need startN <= endN < 2^31";
    return FS_E_ARG;
  }
  const uint32_t g = cand_grid(c);
  switch (path) {
    case FS_PATH_FP64: {
      const double *nt = nullptr, *pt = nullptr;
      if ((rc = ntab_for(S(s), n0 + 1, n1 - n0, &nt, &pt))) return rc;
      k_iterate_fp64<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_factor, fcap, n0, n1, nt, pt);
      break;
    }
    case FS_PATH_HYBRID: {
      const double *nt = nullptr, *pt = nullptr;
      if ((rc = ntab_for(S(s), n0 + 1, n1 - n0, &nt, &pt))) return rc;
      k_iterate_hybrid<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_factor, fcap, n0, n1, nt, pt);
      break;
    }
    case FS_PATH_MONT:
      k_iterate_mont<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_factor, fcap, n0, n1);
      break;
    default:
      g_last_error = "iterate: unknown path";
      return FS_E_ARG;
  }
  LAUNCHCHK("iterate");
  if (g_inject == FS_CTR_FACTOROVF) {
    k_set_flag<<<1, 64, 0, S(s)>>>(d_ctr, FS_CTR_FACTOROVF);
    g_inject = 0;
  }
  return FS_OK;
}

// The last slot of the verify buffer (past the FS_VERIFY_L2 partial products)
// holds the setup residue of the first candidate when lanes run setup and
// iterate in one launch.
static uint64_t* fs_verify_res0(void* d_verify) { return &((VerifyPair*)d_verify)[FS_VERIFY_BLOCKS + 63].slow; }

static int lane_args(const fs_cands* c, const fs_lane_set* ls, int path, LaneArgs* a, const char* who) {
  int rc = check_cands(c, who);
  if (rc) return rc;
  if (path != FS_PATH_FP64 && path != FS_PATH_HYBRID) {
    g_last_error = std::string(who) + ": lanes run on the FP64 and hybrid paths only";
    return FS_E_ARG;
  }
  if (ls->lanes < 1 || ls->lanes > FS_LANES || ls->stride == 0 || ls->stride % FS_BLOCK ||
      (uint64_t)ls->lanes * ls->stride > c->cap) {
    g_last_error = std::string(who) + ": bad lane layout";
    return FS_E_ARG;
  }
  memset(a, 0, sizeof *a);
  a->lanes = ls->lanes;
  a->stride = ls->stride;
  for (uint32_t k = 0; k <= ls->lanes; ++k) a->nb[k] = ls->nb[k];
  for (uint32_t k = 0; k < ls->lanes; ++k) {
    if (k && a->nb[k] < a->nb[k - 1]) {
      g_last_error = std::string(who) + ": lane n-ranges out of order";
      return FS_E_ARG;
    }
    a->poff[k] = ls->poff[k];
    a->terms[k] = ls->terms[k];
  }
  if (a->nb[ls->lanes] < a->nb[ls->lanes - 1] || a->nb[ls->lanes] > 0x7fffffffu) {
    g_last_error = std::string(who) + ": need n < 2^31";
    return FS_E_ARG;
  }
  return FS_OK;
}

static uint32_t lane_grid(const fs_lane_set* ls) { return (uint32_t)((uint64_t)ls->lanes * ls->stride / FS_BLOCK); }

int fs_lanes(const fs_cands* c, uint32_t* d_ctr, fs_factor* d_factor, uint32_t fcap, const fs_lane_set* ls,
             const uint32_t* cuts, uint32_t ncuts, void* d_verify, int path, fs_stream s) {
  LaneArgs a;
  int rc = lane_args(c, ls, path, &a, "lanes");
  if (rc) return rc;
  for (uint32_t k = 0; k < ls->lanes; ++k)
    if ((ls->terms[k] & 3) || (ls->poff[k] & 3) || !ls->progs) {
      g_last_error = "lanes: bad setup program";
      return FS_E_ARG;
    }
  if (ncuts > 4 || !d_verify) {
    g_last_error = "lanes: at most 4 loop boundaries, and a verify buffer";
    return FS_E_ARG;
  }
  a.ncuts = ncuts;
  for (uint32_t k = 0; k < ncuts; ++k) a.cuts[k] = cuts[k];
  const uint32_t n0 = ls->nb[0], n1 = ls->nb[ls->lanes];
  const double *nt = nullptr, *pt = nullptr, *tab = nullptr, *pairs = nullptr;
  if ((rc = ntab_for(S(s), n0 + 1, n1 - n0, &nt, &pt))) return rc;
  {
    std::lock_guard<std::mutex> lk(g_ntab_mu);
    const NTab& t = g_ntab[S(s)];
    tab = t.d;
    pairs = t.pairs();
    a.tbase = t.base;
    a.thalf = t.half();
  }
  const uint64_t* pg = (const uint64_t*)ls->progs;
  uint64_t* res0 = fs_verify_res0(d_verify);
  if (path == FS_PATH_FP64)
    k_lanes<FS_PATH_FP64><<<lane_grid(ls), FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_factor, fcap, a, pg, tab, pairs, res0);
  else
    k_lanes<FS_PATH_HYBRID><<<lane_grid(ls), FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_factor, fcap, a, pg, tab, pairs, res0);
  LAUNCHCHK("lanes");
  if (g_inject == FS_CTR_FACTOROVF) {
    k_set_flag<<<1, 64, 0, S(s)>>>(d_ctr, FS_CTR_FACTOROVF);
    g_inject = 0;
  }
  return FS_OK;
}

int fs_verifyresult_res0(const fs_cands* c, uint32_t* d_ctr, void* d_verify, fs_stream s) {
  int rc = check_cands(c, "verifyresult");
  if (rc) return rc;
  k_verifyresult<<<1, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, (VerifyPair*)d_verify, fs_verify_res0(d_verify));
  LAUNCHCHK("verifyresult");
  if (g_inject == FS_CTR_POWERTABLE) {
    k_set_flag<<<1, 64, 0, S(s)>>>(d_ctr, FS_CTR_POWERTABLE);
    g_inject = 0;
  }
  return FS_OK;
}

extern "C" int fs_iterate_noalloc(const fs_cands* c, uint32_t* d_ctr, fs_factor* d_factor, uint32_t fcap,
                                  uint32_t n0, uint32_t n1, int path, fs_stream s) {
  int rc = check_cands(c, "iterate_noalloc");
  if (rc) return rc;
  if (n0 > n1 || n1 > 0x7fffffffu) {
    g_last_error = "iterate_noalloc: need startN <= endN < 2^31";
    return FS_E_ARG;
  }
  if (n1 - n0 > FS_ITERATE_NOALLOC_MAX_N) {
    g_last_error = "iterate_noalloc: endN - startN exceeds FS_ITERATE_NOALLOC_MAX_N (2^20)";
    return FS_E_ARG;
  }
  if (path != FS_PATH_FP64 && path != FS_PATH_HYBRID && path != FS_PATH_MONT) {
    g_last_error = "iterate_noalloc: unknown path";
    return FS_E_ARG;
  }
  if ((rc = ensure_init())) return rc;
  double* base = g_device < kMaxDevices ? g_scratch_ptr[g_device] : nullptr;
  if (!base) {
    set_error("hipGetSymbolAddress(scratch)",
              g_device < kMaxDevices && g_scratch_err[g_device] != hipSuccess ? g_scratch_err[g_device]
                                                                              : hipErrorInvalidSymbol);
    return FS_E_HIP;
  }
  const uint32_t g = cand_grid(c);
  if (path != FS_PATH_MONT) {
    // multipliers n0+1 .. and their pairs, 32 entries past the range: the
    // loops read one unrolled block ahead (at most 16 multipliers or 16 pair
    // words); count + 32 <= kScrCap keeps every entry inside the array
    const uint32_t len = (n1 - n0) + 32;
    k_fill_ntab<<<(len + 255) / 256, 256, 0, S(s)>>>(base, base + kScrCap, kScrCap / 2 + 1, n0 + 1, len);
    LAUNCHCHK("iterate_noalloc tables");
  }
  switch (path) {
    case FS_PATH_FP64:
      k_iterate_fp64<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_factor, fcap, n0, n1, base, base + kScrCap);
      break;
    case FS_PATH_HYBRID:
      k_iterate_hybrid<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_factor, fcap, n0, n1, base, base + kScrCap);
      break;
    default:
      k_iterate_mont<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_factor, fcap, n0, n1);
      break;
  }
  LAUNCHCHK("iterate_noalloc");
  return FS_OK;
}

extern "C" int fs_check(const fs_cands* c, uint32_t* d_ctr, uint64_t* d_sum, uint32_t nmax, fs_stream s) {
  int rc = check_cands(c, "check");
  if (rc) return rc;
  k_check<<<std::min<uint32_t>(cand_grid(c), 2048), FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_sum, nmax);
  LAUNCHCHK("check");
  if (g_inject == FS_CTR_VALIDATE) {
    k_set_flag<<<1, 64, 0, S(s)>>>(d_ctr, FS_CTR_VALIDATE);
    g_inject = 0;
  }
  return FS_OK;
}

extern "C" size_t fs_verify_bytes(void) { return (FS_VERIFY_BLOCKS + 64) * sizeof(VerifyPair); }

extern "C" int fs_verifyslow(const fs_cands* c, const uint32_t* d_ctr, void* d_verify, uint32_t nmin,
                             fs_stream s) {
  int rc = check_cands(c, "verifyslow");
  if (rc) return rc;
  // about 16 multipliers per thread, at most FS_VERIFY_BLOCKS blocks
  const uint32_t g = (uint32_t)std::min<uint64_t>(FS_VERIFY_BLOCKS, std::max<uint64_t>(1, ((uint64_t)nmin + 4095) / 4096));
  k_verifyslow<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, (VerifyPair*)d_verify, nmin);
  LAUNCHCHK("verifyslow");
  return FS_OK;
}

extern "C" int fs_verifypow(const fs_cands* c, const uint32_t* d_ctr, const uint64_t* d_bases,
                            const uint32_t* d_exps, uint32_t count, void* d_verify, fs_stream s) {
  int rc = check_cands(c, "verifypow");
  if (rc) return rc;
  // one power per thread, at most FS_VERIFY_BLOCKS blocks
  const uint32_t g = std::min<uint32_t>(FS_VERIFY_BLOCKS, std::max<uint32_t>(1, (count + FS_BLOCK - 1) / FS_BLOCK));
  k_verifypow<<<g, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, d_bases, d_exps, count, (VerifyPair*)d_verify);
  LAUNCHCHK("verifypow");
  return FS_OK;
}

extern "C" int fs_verifyreduce(const fs_cands* c, const uint32_t* d_ctr, void* d_verify, fs_stream s) {
  int rc = check_cands(c, "verifyreduce");
  if (rc) return rc;
  k_verifyreduce<<<FS_VERIFY_L2, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, (VerifyPair*)d_verify);
  LAUNCHCHK("verifyreduce");
  return FS_OK;
}

extern "C" int fs_verifyresult(const fs_cands* c, uint32_t* d_ctr, void* d_verify, fs_stream s) {
  int rc = check_cands(c, "verifyresult");
  if (rc) return rc;
  k_verifyresult<<<1, FS_BLOCK, 0, S(s)>>>(*c, d_ctr, (VerifyPair*)d_verify, nullptr);
  LAUNCHCHK("verifyresult");
  if (g_inject == FS_CTR_POWERTABLE) {
    k_set_flag<<<1, 64, 0, S(s)>>>(d_ctr, FS_CTR_POWERTABLE);
    g_inject = 0;
  }
  return FS_OK;
}

// Factor verification on the device: the contract of fs_verify_factors.
extern "C" int64_t fs_verify_factors_gpu(const fs_factor* f, uint64_t n, fs_stream s) {
  if (n == 0) return -1;
  if (!f) {
    g_last_error = "fs_verify_factors_gpu: null factor list";
    return FS_E_VERIFY;
  }
  if (n >= 0x7FFFFFFFull) {
    g_last_error = "fs_verify_factors_gpu: too many factors";
    return FS_E_VERIFY;
  }
  int rc = ensure_init();
  if (rc) return rc;
  const uint32_t nf = (uint32_t)n;
  std::vector<uint32_t> first(nf + 1);
  std::vector<char> direct(nf, 0);
  uint64_t nrg = 0;
  for (uint32_t i = 0; i < nf; ++i) {
    first[i] = (uint32_t)nrg;
    const uint64_t p = f[i].p;
    const uint32_t fn = (uint32_t)(f[i].nc < 0 ? -(int64_t)f[i].nc : f[i].nc);
    if (p < 3 || (p & 1) == 0 || fn < 2 || fn >= p) {
      direct[i] = 1;
      continue;
    }
    nrg += ((uint64_t)fn - 2) / FS_VRANGE + 1;
    if (nrg >= 0xFFFFFFFFull) {
      g_last_error = "fs_verify_factors_gpu: too many multiplier ranges";
      return FS_E_VERIFY;
    }
  }
  first[nf] = (uint32_t)nrg;
  std::vector<uint64_t> res(nf, 0);
  if (nrg > 0) {
    fs_factor* d_f = nullptr;
    uint32_t* d_first = nullptr;
    uint64_t *d_part = nullptr, *d_res = nullptr;
    auto release = [&]() {
      (void)hipFree(d_f);
      (void)hipFree(d_first);
      (void)hipFree(d_part);
      (void)hipFree(d_res);
    };
    hipError_t e = hipMalloc((void**)&d_f, (size_t)nf * sizeof(fs_factor));
    if (e == hipSuccess) e = hipMalloc((void**)&d_first, ((size_t)nf + 1) * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&d_part, (size_t)nrg * 8);
    if (e == hipSuccess) e = hipMalloc((void**)&d_res, (size_t)nf * 8);
    if (e != hipSuccess) {
      release();
      set_error("hipMalloc(verify)", e);
      return FS_E_NOMEM;
    }
    e = hipMemcpyAsync(d_f, f, (size_t)nf * sizeof(fs_factor), hipMemcpyHostToDevice, S(s));
    if (e == hipSuccess) e = hipMemcpyAsync(d_first, first.data(), ((size_t)nf + 1) * 4, hipMemcpyHostToDevice, S(s));
    if (e == hipSuccess) {
      k_verify_ranges<<<(uint32_t)((nrg + FS_BLOCK - 1) / FS_BLOCK), FS_BLOCK, 0, S(s)>>>(d_f, d_first, nf,
                                                                                         (uint32_t)nrg, d_part);
      e = hipGetLastError();
    }
    if (e == hipSuccess) {
      k_verify_combine<<<(nf + FS_BLOCK - 1) / FS_BLOCK, FS_BLOCK, 0, S(s)>>>(d_f, d_first, nf, d_part, d_res);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(res.data(), d_res, (size_t)nf * 8, hipMemcpyDeviceToHost, S(s));
    if (e == hipSuccess) e = hipStreamSynchronize(S(s));
    release();
    if (e != hipSuccess) {
      set_error("verify kernels", e);
      return FS_E_HIP;
    }
  }
  for (uint32_t i = 0; i < nf; ++i) {
    const uint64_t p = f[i].p;
    const int32_t c = f[i].nc < 0 ? -1 : 1;
    const uint32_t fn = (uint32_t)(f[i].nc < 0 ? -(int64_t)f[i].nc : f[i].nc);
    // p | n! + c: n! = -c (mod p), i.e. 1 for n!-1 and p-1 for n!+1
    const bool ok = direct[i] ? fs_verify_factor(p, fn, c) != 0 : (c == -1 ? res[i] == 1 : res[i] == p - 1);
    if (!ok) return (int64_t)i;
  }
  return -1;
}
