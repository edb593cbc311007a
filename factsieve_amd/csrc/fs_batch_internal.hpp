// fs_batch_internal.hpp -- device records and launch helpers shared by the
// batched form (fs_batch_host.cpp) and the serving loop (fs_server.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <vector>
#include "../../include/factsieve.h"

#ifndef FS_BLOCK
#define FS_BLOCK 256
#endif

// One job of a batch as the kernels see it (64 bytes).
struct fs_bjob {
  uint64_t pmin, pmax;     // p in [pmin, pmax)
  uint32_t nmin, nmax;     // n in [nmin, nmax)
  uint32_t path;           // FS_PATH_FP64 / _HYBRID / _MONT
  uint32_t prog_off;       // setup program offset in 8-byte words
  uint32_t prog_terms;     // setup program term count
  uint32_t cand_off;       // candidate region [cand_off, cand_off + cand_cap)
  uint32_t cand_cap;
  uint32_t pg_first;       // first candidate-generation workgroup of this job
  uint32_t flags;          // FS_BJ_*
  uint32_t parent;         // job whose candidates, counters, sums and factor tag it uses (itself
                           // unless it is a later n-subrange of a job split over several lanes)
};
#define FS_BJ_NEW 1u       // counters cleared and candidates generated in this launch
#define FS_BJ_NOFOLD 2u    // an n-subrange that ends before the job's nmax: no checksum fold

// per-job flag bits
#define FS_BJF_CANDOVF 1u
#define FS_BJF_VALIDATE 2u

// launch parameters read by the kernels from device memory (so that a captured
// graph of the launches can be replayed for batches and quanta of any size)
#define FS_BP_NJ 0         // job-table entries (clear and candidate generation)
#define FS_BP_PGBLOCKS 1   // candidate-generation workgroups in all (virtual blocks)
#define FS_BP_ROUND0 2     // round r: [FS_BP_ROUND0 + 2r] first order entry, [+1] entry count
#define FS_BP_WORDS(rounds) (FS_BP_ROUND0 + 2 * (rounds))

// global words
#define FS_BG_HEAD 0       // tile queue head
#define FS_BG_NTILES 1     // tiles in the list
#define FS_BG_NF 2         // factors appended
#define FS_BG_FOVF 3       // factor buffer overflow
#define FS_BG_TILEOVF 4    // tile list overflow
#define FS_BG_WORDS 8

struct fs_batch_dev {
  fs_bjob* jobs = nullptr;
  uint32_t* order = nullptr;   // job visit order for the tile list
  uint32_t* chunk = nullptr;   // (n0, n1) per order entry
  uint32_t* jcnt = nullptr;
  uint32_t* jflag = nullptr;
  uint64_t* jsum = nullptr;
  uint32_t* glob = nullptr;
  uint32_t* pref = nullptr;
  uint4* tiles = nullptr;
  uint32_t tiles_cap = 0;
  fs_cands c{};
  uint64_t* prog = nullptr;
  double* ntab = nullptr;      // ntab[k] = k for k < ntab_len, then the pair table
  uint32_t ntab_len = 0;
  uint32_t pair_len = 0;       // the pair table covers multipliers a < pair_len (fs_bpair_len)
  fs_factor* fac = nullptr;
  uint32_t fcap = 0;
  uint32_t* params = nullptr;  // FS_BP_* words
  uint32_t params_cap = 0;
};

// Enqueue clear -> candidate generation (jobs flagged FS_BJ_NEW, grid-stride
// over the FS_BP_PGBLOCKS virtual blocks with pg_grid workgroups) -> for each
// of `rounds` rounds: tile list (the entries FS_BP_ROUND0 + 2r of d.params
// name in d.order / d.chunk) -> persistent sieve.  Every count the kernels
// need is read from d.params, so the launches depend only on the buffers and
// on (jobs_cap, rounds, pg_grid, tile, blocks, unroll).  unroll: 2, 4 or 8
// steps per +-1 branch in the iterate loops.
int fs_batch_launch(const fs_batch_dev& d, uint32_t jobs_cap, uint32_t rounds, uint32_t pg_grid, uint32_t tile,
                    uint32_t blocks, uint32_t unroll, hipStream_t st);
// Pair table of the paired FP64 step, after the ntab_len multipliers: for
// a < pair_len, the entry (a*(a+1), a+1) at index a >> 1 of half a & 1, each
// half holding pair_len / 2 + 1 entries of two doubles.  Only moduli with
// n1 <= 2^22 take paired steps, so the table stops there.
inline uint32_t fs_bpair_len(uint32_t ntab_len) { return ntab_len < (1u << 22) + 64 ? ntab_len : (1u << 22) + 64; }
inline size_t fs_bntab_doubles(uint32_t ntab_len) {
  return (size_t)ntab_len + 4 * ((size_t)fs_bpair_len(ntab_len) / 2 + 1);
}
int fs_batch_fill_ntab(double* t, uint32_t count, hipStream_t st);
extern "C" void fs_set_hip_error(const char* what, int err);

// A captured and instantiated launch sequence, replayed while its key (the
// buffers and the launch shape) stays the same.
struct fs_batch_graph {
  std::vector<uint64_t> key;
  hipGraphExec_t exec = nullptr;
};

// Grow-only device buffers of one batched-form client (the batch entry point
// or a server) with its own stream.
struct fs_batch_bufs {
  fs_batch_dev d;
  uint32_t jobs_cap = 0, cands_cap = 0, prog_cap = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<fs_batch_graph> graphs;  // most recent last
  uint64_t graph_captures = 0, graph_replays = 0;
  int ensure(uint32_t nj, uint32_t ncands, uint32_t nprog, uint32_t ntiles, uint32_t nmax_max, uint32_t fcap,
             uint32_t nparams = FS_BP_WORDS(1));
  // Enqueue fs_batch_launch on st for the buffers in d (which may differ from
  // this->d in the factor and global-word pointers), as a replay of a cached
  // graph when graph is true (capturing one on a miss), else directly.
  int enqueue(const fs_batch_dev& d, uint32_t rounds, uint32_t pg_grid, uint32_t tile, uint32_t blocks,
              uint32_t unroll, bool graph);
  void drop_graphs();
  void release();
};
uint32_t fs_device_cus();
// Upload of a job table / tile order / chunks / programs and the launch, then
// a wait for the stream.  rounds: entry counts of consecutive rounds of
// order/chunk (empty = one round over all of order).
int fs_batch_run_device(fs_batch_bufs& E, const std::vector<fs_bjob>& bj, const std::vector<uint32_t>& order,
                        const std::vector<uint32_t>& chunk, const std::vector<uint64_t>& words, uint32_t pg_blocks,
                        uint32_t ncands, uint32_t tile, uint32_t blocks, uint32_t fcap, bool graph, float* dev_ms,
                        const std::vector<uint32_t>& rounds = {}, uint32_t unroll = 8);
// Launch parameters of a batch: nj entries, pg_blocks virtual blocks, and the
// rounds' (first entry, count) pairs.
std::vector<uint32_t> fs_batch_params(uint32_t nj, uint32_t pg_blocks, const std::vector<uint32_t>& rounds);
// Read back per-job counters, flags, sums and the factor records.
int fs_batch_read(fs_batch_bufs& E, uint32_t nj, std::vector<uint32_t>& cnt, std::vector<uint32_t>& flag,
                  std::vector<uint64_t>& sum, std::vector<fs_factor>& fac, uint32_t* glob_out);
uint32_t fs_pg_blocks(uint64_t pmin, uint64_t pmax);
