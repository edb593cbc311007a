// fs_batch.hip -- many sieve jobs in one persistent launch.
//
// A batch is run as device launches on one stream (optionally captured into a
// hipGraph):
//   k_bclear       zero the per-job counters and the tile queue
//   k_bprimegen    candidate generation for every job at once: workgroup b
//                  belongs to the job whose block range holds b and appends
//                  into that job's candidate region (bounded)
// then, once per round (one round unless the n-ranges are cut into chunks):
//   k_bmake_tiles  one workgroup turns the per-job candidate counts into a tile
//                  list (job, first candidate, n0, n1) in cost order
//   k_bsieve       persistent workgroups pull tiles from an atomic queue; for
//                  each candidate of a tile: setup (nmin!), iterate over the
//                  tile's n-range with +-1 capture, and the checksum fold, all
//                  in registers; between rounds the residue stays in the
//                  candidate region
// A job may also be split over several lanes per prime: its n-range is cut
// into subranges that run as independent entries, each with the setup program
// of its own first n.  The serving loop (fs_server.cpp) drives the same kernels
// with n-chunked tiles, one round per quantum.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <vector>
#include "fs_arith.hpp"
#include "fs_kern.hpp"
#include "fs_internal.hpp"
#include "fs_batch_internal.hpp"
#include "../../include/factsieve.h"

#define BCHK(call, what)                          \
  do {                                            \
    hipError_t e__ = (call);                      \
    if (e__ != hipSuccess) {                      \
      fs_set_hip_error(what, (int)e__);           \
      return FS_E_HIP;                            \
    }                                             \
  } while (0)

// ================================================================== kernels

// Clears the counters of the jobs that start in this launch and the global words.
__global__ void k_bclear(const fs_bjob* __restrict__ jobs, uint32_t* jcnt, uint32_t* jflag, uint64_t* jsum,
                         const uint32_t* __restrict__ params, uint32_t* glob) {
  const uint32_t nj = params[FS_BP_NJ];
  for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nj; t += gridDim.x * blockDim.x) {
    if (jobs[t].flags & FS_BJ_NEW) {
      jcnt[t] = 0;
      jflag[t] = 0;
      jsum[t] = 0;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x < FS_BG_WORDS) glob[threadIdx.x] = 0;
}

// Job of primegen virtual block b: last job with pg_first <= b.
__device__ __forceinline__ uint32_t job_of_block(const fs_bjob* jobs, uint32_t nj, uint32_t b) {
  uint32_t lo = 0, hi = nj - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi + 1) >> 1;
    if (jobs[mid].pg_first <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Candidate generation of every new job: the workgroups stride over the
// FS_BP_PGBLOCKS virtual blocks; virtual block b belongs to the job whose
// block range holds it and appends into that job's candidate region (bounded).
__global__ __launch_bounds__(FS_BLOCK) void k_bprimegen(const fs_bjob* __restrict__ jobs,
                                                        const uint32_t* __restrict__ params, fs_cands c,
                                                        uint32_t* jcnt, uint32_t* jflag) {
  __shared__ uint64_t s_tab[FS_NMASK];
  __shared__ uint64_t s_mask[FS_BLOCK];
  __shared__ uint32_t s_excl[FS_BLOCK];
  __shared__ uint32_t s_wsum[FS_BLOCK / 64 + 1];
  const uint32_t nj = params[FS_BP_NJ], total = params[FS_BP_PGBLOCKS];
  if (blockIdx.x >= total || nj == 0) return;  // uniform per workgroup
  fs_sieve_table_init(s_tab);
  for (uint32_t b = blockIdx.x; b < total; b += gridDim.x) {
    const uint32_t j = job_of_block(jobs, nj, b);
    const fs_bjob J = jobs[j];
    if (J.flags & FS_BJ_NEW) {  // uniform per workgroup
      fs_primegen_block(
          J.pmin | 1, J.pmax, b - J.pg_first, 0, 1, s_tab, s_mask, s_excl, s_wsum,
          [&](uint32_t n) { return atomicAdd(&jcnt[j], n); },
          [&](uint32_t idx, uint64_t N, uint64_t q, uint64_t r2) {
            if (idx < J.cand_cap) {
              const uint32_t g = J.cand_off + idx;
              c.p[g] = N;
              c.q[g] = q;
              c.r2[g] = r2;
              c.res[g] = 0;
              c.nnext[g] = J.nmin;
            } else {
              atomicOr(&jflag[j], FS_BJF_CANDOVF);
            }
          });
    }
    __syncthreads();  // the shared masks are rewritten by the next virtual block
  }
}

// Tile list of round r: the params-given entries of `order` (cost order),
// covering n in [n0, n1) of each job as given by chunk[] (n0 = chunk[2k],
// n1 = chunk[2k+1] for job order[k]); pref: entries+1 words of scratch.
__global__ __launch_bounds__(FS_BLOCK) void k_bmake_tiles(const fs_bjob* __restrict__ jobs,
                                                          const uint32_t* __restrict__ order,
                                                          const uint32_t* __restrict__ chunk,
                                                          const uint32_t* __restrict__ params, uint32_t r,
                                                          const uint32_t* jcnt, uint32_t tile, uint4* tiles,
                                                          uint32_t tiles_cap, uint32_t* pref, uint32_t* glob) {
  const uint32_t first = params[FS_BP_ROUND0 + 2 * r], nj = params[FS_BP_ROUND0 + 2 * r + 1];
  order += first;
  chunk += 2 * first;
  if (threadIdx.x == 0) {
    glob[FS_BG_HEAD] = 0;  // the previous round's persistent kernel has finished (stream order)
    uint32_t acc = 0;
    for (uint32_t k = 0; k < nj; ++k) {
      const fs_bjob& J = jobs[order[k]];
      const uint32_t live = min(jcnt[J.parent], J.cand_cap);
      pref[k] = acc;
      acc += (live + tile - 1) / tile;
    }
    pref[nj] = acc;
    glob[FS_BG_NTILES] = min(acc, tiles_cap);
    if (acc > tiles_cap) atomicOr(&glob[FS_BG_TILEOVF], 1u);
  }
  __syncthreads();
  for (uint32_t k = 0; k < nj; ++k) {
    const uint32_t j = order[k], b = pref[k], e = min(pref[k + 1], tiles_cap);
    for (uint32_t t = b + threadIdx.x; t < e; t += FS_BLOCK)
      tiles[t] = make_uint4(j, (t - b) * tile, chunk[2 * k], chunk[2 * k + 1]);
  }
}

// Persistent sieve: each workgroup repeatedly takes the next tile; every queue
// index is taken once, and a workgroup leaves when the index passes the tile
// count, so all workgroups exit.  A job split over several lanes per prime
// appears as n-subrange entries that share the candidates, counters, sums and
// factor tag of their parent; each entry's setup program starts at its own
// nmin, and only the entry that ends at the job's nmax folds the checksum.
// U: steps per +-1 branch of the iterate loops.
template <int U>
__global__ __launch_bounds__(FS_BLOCK) void k_bsieve(const fs_bjob* __restrict__ jobs,
                                                     const uint4* __restrict__ tiles, fs_cands c,
                                                     const uint64_t* __restrict__ prog,
                                                     const double* __restrict__ ntab, uint32_t pair_len,
                                                     const double* __restrict__ ptab, fs_factor* fac, uint32_t fcap,
                                                     const uint32_t* jcnt, uint32_t* jflag, uint64_t* jsum,
                                                     uint32_t* glob, uint32_t tile) {
  __shared__ uint32_t s_t;
  const uint32_t ntiles = glob[FS_BG_NTILES];
  for (;;) {
    if (threadIdx.x == 0) s_t = atomicAdd(&glob[FS_BG_HEAD], 1u);
    __syncthreads();
    const uint32_t t = s_t;
    __syncthreads();
    if (t >= ntiles) break;
    const uint4 T = tiles[t];
    const uint32_t j = T.x, begin = T.y, n0 = T.z, n1 = T.w;
    const fs_bjob J = jobs[j];
    const uint32_t jo = J.parent;  // owner of counters, sums and factor tags
    const uint32_t live = min(jcnt[jo], J.cand_cap);
    const bool first = (n0 == J.nmin), last = (n1 == J.nmax), fold = last && !(J.flags & FS_BJ_NOFOLD);
    const uint64_t* jp = prog + J.prog_off;
    const double* nt = ntab + n0 + 1;
    // pairs (n0+1, n0+2), (n0+3, n0+4), ... while the table covers n1 + 16
    const uint32_t a0 = n0 + 1;
    const bool pairs = (n1 + 16 <= pair_len);
    const double* pt = ptab + (pairs ? (size_t)(a0 & 1) * 2 * (pair_len / 2 + 1) + 2 * (size_t)(a0 >> 1) : 0);
    for (uint32_t k = threadIdx.x; k < tile; k += FS_BLOCK) {
      const uint32_t i = begin + k;
      uint64_t v = 0;
      if (i < live) {
        const uint32_t g = J.cand_off + i;
        const uint64_t p = c.p[g], q = c.q[g], r2 = c.r2[g];
        auto rec = [&](int32_t nc) {
          fs_append_factor(&glob[FS_BG_NF], &glob[FS_BG_FOVF], fac, fcap, p, nc, (int32_t)jo);
        };
        uint64_t s;
        if (!first && c.nnext[g] != n0) atomicOr(&jflag[jo], FS_BJF_VALIDATE);
        if (J.path == FS_PATH_FP64) {
          s = first ? fs_cand_setup_fp64(p, q, r2, 0, (const double*)jp, 0, J.prog_terms) : c.res[g];
          s = fs_cand_iter_fp64<U, (U == 8 ? 16 : U)>(p, s, n0, n1, nt, pt, pairs, rec);
        } else if (J.path == FS_PATH_HYBRID) {
          s = first ? fs_cand_setup_hybrid(p, q, r2, 0, jp, J.prog_terms, 0, J.prog_terms) : c.res[g];
          // paired hybrid blocks of U steps (16 would hold 64 scalar registers of
          // pair words and push the kernel to 5 waves per SIMD)
          s = fs_cand_iter_hybrid<U, U>(p, s, n0, n1, nt, pt, pairs, rec);
        } else {
          s = first ? fs_cand_setup_mont(p, q, r2, 0, jp, J.prog_terms, 0, J.prog_terms) : c.res[g];
          s = fs_cand_iter_mont(p, q, r2, s, n0, n1, rec);
        }
        if (fold) {
          bool ok;
          v = fs_cand_fold(p, q, r2, s, J.nmax, ok);
          if (!ok) atomicOr(&jflag[jo], FS_BJF_VALIDATE);
        } else if (!last) {
          c.res[g] = s;
          c.nnext[g] = n1;
        } else if (s >= p) {  // a subrange's end residue must be canonical
          atomicOr(&jflag[jo], FS_BJF_VALIDATE);
        }
      }
      if (fold) {
        v = fs_wave_sum(v);
        if ((threadIdx.x & 63) == 0 && v != 0) atomicAdd((unsigned long long*)&jsum[jo], (unsigned long long)v);
      }
    }
  }
}

__global__ void k_bfill_ntab(double* t, uint32_t count, double* pt, uint32_t pair_len) {
  const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < count) t[j] = (double)j;
  if (j < pair_len) {
    const uint64_t a = j;
    double* e = pt + (size_t)(j & 1) * 2 * (pair_len / 2 + 1) + 2 * (size_t)(j >> 1);
    e[0] = (double)(a * (a + 1));
    e[1] = (double)(a + 1);
  }
}

// ============================================================ launch helpers

int fs_batch_launch(const fs_batch_dev& d, uint32_t jobs_cap, uint32_t rounds, uint32_t pg_grid, uint32_t tile,
                    uint32_t blocks, uint32_t unroll, hipStream_t st) {
  k_bclear<<<std::min<uint32_t>((jobs_cap + 255) / 256, 64) + 1, 256, 0, st>>>(d.jobs, d.jcnt, d.jflag, d.jsum,
                                                                               d.params, d.glob);
  BCHK(hipGetLastError(), "batch clear");
  if (pg_grid) {
    k_bprimegen<<<pg_grid, FS_BLOCK, 0, st>>>(d.jobs, d.params, d.c, d.jcnt, d.jflag);
    BCHK(hipGetLastError(), "batch primegen");
  }
  for (uint32_t r = 0; r < rounds; ++r) {
    k_bmake_tiles<<<1, FS_BLOCK, 0, st>>>(d.jobs, d.order, d.chunk, d.params, r, d.jcnt, tile, d.tiles, d.tiles_cap,
                                          d.pref, d.glob);
    BCHK(hipGetLastError(), "batch tiles");
    if (unroll == 8)
      k_bsieve<8><<<blocks, FS_BLOCK, 0, st>>>(d.jobs, d.tiles, d.c, d.prog, d.ntab, d.pair_len, d.ntab + d.ntab_len,
                                               d.fac, d.fcap, d.jcnt, d.jflag, d.jsum, d.glob, tile);
    else if (unroll == 2)
      k_bsieve<2><<<blocks, FS_BLOCK, 0, st>>>(d.jobs, d.tiles, d.c, d.prog, d.ntab, d.pair_len, d.ntab + d.ntab_len,
                                               d.fac, d.fcap, d.jcnt, d.jflag, d.jsum, d.glob, tile);
    else
      k_bsieve<4><<<blocks, FS_BLOCK, 0, st>>>(d.jobs, d.tiles, d.c, d.prog, d.ntab, d.pair_len, d.ntab + d.ntab_len,
                                               d.fac, d.fcap, d.jcnt, d.jflag, d.jsum, d.glob, tile);
    BCHK(hipGetLastError(), "batch sieve");
  }
  return FS_OK;
}

int fs_batch_fill_ntab(double* t, uint32_t count, hipStream_t st) {
  k_bfill_ntab<<<(count + 255) / 256, 256, 0, st>>>(t, count, t + count, fs_bpair_len(count));
  BCHK(hipGetLastError(), "batch ntab");
  return FS_OK;
}
