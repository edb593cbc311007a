// fs_internal.hpp -- declarations shared by the library's translation units
// (not part of the C ABI).
#pragma once
#include <stddef.h>
#include <stdint.h>
#include <memory>
#include <vector>
#include "../../include/factsieve.h"

extern "C" void fs_set_error_msg(const char* msg);

// Primes <= limit (cached for the last limit asked).
std::vector<uint32_t> fs_primes_upto(uint32_t limit);
// Horner program of nmin!: multipliers < bound, 0 = square.
std::vector<uint64_t> fs_horner_program(uint32_t nmin, uint64_t bound);
// Setup program words of (nmin, path) as fs_setup_program writes them, from a
// small shared cache; holders keep their program alive past cache eviction.
struct fs_program {
  uint32_t terms;
  std::vector<uint64_t> words;
};
std::shared_ptr<const fs_program> fs_program_get(uint32_t nmin, int path);
// The cached program of (nmin, path), or null when it is not cached.
std::shared_ptr<const fs_program> fs_program_cached(uint32_t nmin, int path);
// Sort by (p, |n|); verify on CPU threads (index of first failure or -1);
// drop composite p and fold kept factors into *checksum (returns kept count).
void fs_sort_factors(fs_factor* f, size_t n);
// prime: null, or n flags that receive 1/0 for the primality of each
// distinct p (first of a run of equal p; -1 elsewhere), computed in the same
// parallel pass for the composite filter.
int64_t fs_verify_all(const fs_factor* f, size_t n, int threads, signed char* prime = nullptr);
// CPU or GPU verifier by mode (FS_VERIFY_*) and the amount of work (the GPU
// above 2^24 multiplications in FS_VERIFY_AUTO); the contract of
// fs_verify_factors.  fs_verify_uses_gpu tells which one fs_verify_select runs.
bool fs_verify_uses_gpu(const fs_factor* f, size_t n, int mode);
int64_t fs_verify_select(const fs_factor* f, size_t n, int threads, int mode, fs_stream s,
                         signed char* prime = nullptr);
// prime: null, or the flags fs_verify_all filled (-1 entries are computed here).
size_t fs_filter_factors(fs_factor* f, size_t n, uint64_t* checksum, bool verbose,
                         const signed char* prime = nullptr);
uint32_t fs_auto_segment_candidates();
// Range checks of setupSearch (cl_sieve.cpp:557-589); FS_OK or FS_E_ARG.
int fs_validate_ranges(uint64_t pmin, uint64_t pmax, uint32_t nmin, uint32_t nmax);
// Upper estimate of the candidates in [lo, lo + len).
uint64_t fs_cand_capacity(uint64_t lo, uint64_t len);
// The fault injected with fs_inject_fault for counter idx, taken once: 1 when
// it was set (and it is then cleared), else 0.
int fs_take_injected_fault(int idx);
// fs_getsegprimes that stores candidate idx (< c->cap) at idx + k * stride for
// every lane k < lanes (stride >= c->cap; the lanes' room must hold
// (lanes - 1) * stride + c->cap entries).  lanes = 1 is fs_getsegprimes.
int fs_getsegprimes_lanes(uint64_t low, uint64_t high, const fs_cands* c, uint32_t* d_ctr, uint32_t lanes,
                          uint32_t stride, fs_stream s);

// Lanes per prime in one launch: lane k is the copy of a segment's candidates
// (their count read on the device, at most stride) at k * stride (stride a
// multiple of the 256-thread block) and covers n in [nb[k], nb[k + 1]),
// starting from the setup program of nb[k]!
// (terms[k] terms, as fs_setup takes it, at word poff[k] of the device buffer
// progs; offsets are multiples of 4 words).  FP64 and hybrid paths.
#define FS_LANES 4
struct fs_lane_set {
  uint32_t lanes, stride;
  uint32_t nb[FS_LANES + 1];
  const void* progs;
  uint32_t poff[FS_LANES];
  uint32_t terms[FS_LANES];
};
// Every lane's setup (its whole program) and then its n-range in one launch,
// cut internally at the (at most 4) loop boundaries `cuts`; appends factors as
// fs_iterate does.  The first candidate's setup residue goes to a slot of the
// verify buffer d_verify, which fs_verifyresult_res0 compares.
int fs_lanes(const fs_cands* c, uint32_t* d_ctr, fs_factor* d_factor, uint32_t fcap, const fs_lane_set* ls,
             const uint32_t* cuts, uint32_t ncuts, void* d_verify, int path, fs_stream s);
// fs_verifyresult against the setup residue fs_lanes kept (c.res[0] has moved on).
int fs_verifyresult_res0(const fs_cands* c, uint32_t* d_ctr, void* d_verify, fs_stream s);
// fs_clearn and fs_clearresult in one launch.
int fs_clearall(uint32_t* d_ctr, uint64_t* d_sum, fs_stream s);
