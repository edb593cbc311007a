// fs_server.cpp -- continuous-batching serving loop for sieve workunits.
//
// Workunits (fs_job) are submitted at any time into an incoming queue that
// has its own lock, so a submit never waits for device work.  Each call of
// fs_server_step is one scheduling quantum, pipelined one quantum deep:
//   1. admit queued workunits while a job slot and a candidate region of the
//      device pool are free (admission by device memory);
//   2. plan and enqueue quantum k: one batched launch over every active
//      workunit (new ones generate their candidates and run setup, all of them
//      advance their residues by the next n-chunk, residues staying resident
//      in their candidate regions, and those reaching nmax fold their
//      checksum), then an asynchronous read-back of the per-slot counters into
//      pinned host memory;
//   3. harvest quantum k-1 while quantum k runs: attribute its factors to
//      their workunits, verify and filter the factors of the workunits it
//      finished, publish their results.
// A workunit that ends in quantum k gives its slot and region to quantum k+1
// at planning time: the stream orders quantum k+1's writes after quantum k's
// reads, and quantum k's counters are copied out before quantum k+1 clears
// them.  Quanta alternate between two factor buffers, two global-word blocks
// and two pinned staging areas (parity k & 1), so harvesting k-1 never reads
// what k writes.  The launch sequence of a quantum is a captured graph,
// replayed for every quantum of the same parity (its sizes are read from
// device memory).
// A workunit's result does not depend on what it shared the device with: the
// kernels touch only its own region, counters and factor tags.  Overflows
// restart the affected workunits from scratch (a bounded number of times).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <atomic>
#include <chrono>
#include <deque>
#include <map>
#include <mutex>
#include <unordered_map>
#include <unordered_set>
#include <vector>
#include "fs_batch_internal.hpp"
#include "fs_internal.hpp"
#include "../../include/factsieve.h"

#define SCHK(call, what)                        \
  do {                                          \
    hipError_t e__ = (call);                    \
    if (e__ != hipSuccess) {                    \
      fs_set_hip_error(what, (int)e__);         \
      return FS_E_HIP;                          \
    }                                           \
  } while (0)

namespace {

constexpr uint32_t kTile = 256;
constexpr uint32_t kBytesPerCand = 8 * 4 + 4;  // p, q, r2, res, nnext
constexpr uint32_t kMaxRestarts = 3;           // overflow restarts of one workunit before it fails
constexpr uint32_t kInitialFactorCap = 1u << 20;
// n-values per quantum by default.  On the 64-workunit mix at 70 % load (n-ranges
// lognormal around 2e4) a quantum of 49,152 gave p50 / p99 latency of 14-15 / 69 ms
// against 15-18 / 100-124 ms with 4,096-8,192: a median workunit then needs one
// quantum, the longest three, and each quantum's fixed host and launch work is
// paid fewer times; 65,536 lengthened the quanta (p50 17 ms) without a better p99.
constexpr uint32_t kDefaultChunk = 49152;

struct WU {
  fs_job job{};
  int path = 0;
  std::shared_ptr<const fs_program> prog;
  uint32_t cap = 0, off = 0, slot = 0;
  uint32_t nplan = 0;      // first n of the next quantum's chunk
  uint32_t gen = 0;        // bumped at every restart: quanta planned before it are ignored
  uint32_t count = 0;      // candidates, from its first quantum
  uint32_t restarts = 0;
  bool admitted = false;   // holds a slot and a region
  bool counted = false;    // counted as active: admitted and its result not yet published
  std::vector<fs_factor> raw;
};

struct SlotPlan {
  uint32_t slot;
  int64_t id;
  uint32_t gen;
  uint32_t n0, n1;
  bool fresh, last;
};

// First-fit allocator of candidate regions [off, off + len) in the pool.
struct Regions {
  std::map<uint32_t, uint32_t> free_;  // off -> len
  void init(uint32_t total) {
    free_.clear();
    if (total) free_[0] = total;
  }
  bool alloc(uint32_t len, uint32_t* off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= len) {
        *off = it->first;
        const uint32_t rest = it->second - len, roff = it->first + len;
        free_.erase(it);
        if (rest) free_[roff] = rest;
        return true;
      }
    }
    return false;
  }
  void release(uint32_t off, uint32_t len) {
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

// Pinned host memory of one parity: the uploads of a quantum and the
// read-back of its per-slot counters and global words.
struct Staging {
  fs_bjob* jobs = nullptr;
  uint32_t* order = nullptr;
  uint32_t* chunk = nullptr;
  uint32_t* params = nullptr;
  uint64_t* words = nullptr;
  size_t words_cap = 0;
  uint32_t* cnt = nullptr;
  uint32_t* flag = nullptr;
  uint64_t* sum = nullptr;
  uint32_t* glob = nullptr;
};

struct Quantum {
  bool live = false;
  std::vector<SlotPlan> plan;
};

struct Done {
  fs_job_result res{};
  std::vector<fs_factor> kept;
};

struct Server {
  fs_server_options o{};
  fs_batch_bufs E;
  hipStream_t hst = nullptr;  // harvest copies and the GPU verifier, beside the quanta's stream
  hipEvent_t ev[2] = {};      // end of a quantum's read-back, by parity
  Staging stg[2];
  fs_factor* fac[2] = {};
  uint32_t fcap[2] = {};
  uint32_t* glob[2] = {};
  uint32_t fcap_target = kInitialFactorCap;
  uint32_t pool = 0, pg_grid = 0, blocks = 0;

  // incoming side (qmu): submissions, finished results, ids not yet polled
  std::mutex qmu;
  std::deque<std::pair<int64_t, fs_job>> incoming;
  std::unordered_map<int64_t, Done> done;
  std::unordered_set<int64_t> live;
  int64_t next_id = 0;

  // device side (smu, held by fs_server_step): admitted and waiting work
  std::mutex smu;
  std::deque<int64_t> pending;                  // admission order (restarts at the front)
  std::unordered_map<int64_t, WU> wus;          // queued-for-admission and admitted workunits
  std::vector<int64_t> slots;                   // slot -> workunit id or -1
  Regions regions;
  Quantum q[2];
  uint64_t quanta = 0;                          // quanta enqueued so far (the next one has parity quanta & 1)

  std::atomic<uint32_t> n_queued{0}, n_active{0}, n_done{0};
  std::atomic<uint64_t> bytes_in_use{0};
};

uint32_t region_len(const fs_job& j) {
  const uint64_t c = fs_cand_capacity(j.pmin, j.pmax - j.pmin);
  return (uint32_t)std::min<uint64_t>((c + kTile - 1) / kTile * kTile, 0x7fffffffull);
}

template <class T>
int pinned(T** p, size_t n) {
  if (*p) (void)hipHostFree(*p);
  *p = nullptr;
  SCHK(hipHostMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T), hipHostMallocDefault), "server pinned buffer");
  return FS_OK;
}

void free_server(Server* s) {
  if (s->E.st) (void)hipStreamSynchronize(s->E.st);
  if (s->hst) (void)hipStreamSynchronize(s->hst);
  for (int k = 0; k < 2; ++k) {
    Staging& g = s->stg[k];
    void* ps[] = {g.jobs, g.order, g.chunk, g.params, g.words, g.cnt, g.flag, g.sum, g.glob};
    for (void* p : ps)
      if (p) (void)hipHostFree(p);
    if (s->fac[k]) (void)hipFree(s->fac[k]);
    if (s->glob[k]) (void)hipFree(s->glob[k]);
    if (s->ev[k]) (void)hipEventDestroy(s->ev[k]);
  }
  if (s->hst) (void)hipStreamDestroy(s->hst);
  s->E.release();
  delete s;
}

// Publish a finished workunit's result (smu held; takes qmu).
void publish(Server* s, int64_t id, WU& w, const fs_job_result& r, std::vector<fs_factor>&& kept) {
  {
    std::lock_guard<std::mutex> lk(s->qmu);
    Done& d = s->done[id];
    d.res = r;
    d.kept = std::move(kept);
  }
  if (w.counted) s->n_active.fetch_sub(1);
  s->wus.erase(id);
  s->n_done.fetch_add(1);
}

void fail(Server* s, int64_t id, WU& w, int status) {
  fs_job_result r;
  memset(&r, 0, sizeof r);
  r.status = status;
  r.path = w.path;
  publish(s, id, w, r, {});
}

// Give a workunit's slot and region back (smu held).
void vacate(Server* s, WU& w) {
  if (!w.admitted) return;
  s->slots[w.slot] = -1;
  s->regions.release(w.off, w.cap);
  s->bytes_in_use.fetch_sub((uint64_t)w.cap * kBytesPerCand);
  w.admitted = false;
}

// Restart a workunit from nothing, at the front of the admission queue, with
// a candidate region of at least `cap` (smu held).  It fails instead after
// kMaxRestarts restarts or when it cannot fit the pool.
void restart(Server* s, int64_t id, WU& w, uint32_t cap) {
  vacate(s, w);
  if (w.counted) s->n_active.fetch_sub(1);
  w.counted = false;
  ++w.gen;
  w.raw.clear();
  w.count = 0;
  if (++w.restarts > kMaxRestarts || cap > s->pool) {
    fs_set_error_msg(cap > s->pool ? "server: workunit candidates exceed the device pool"
                                   : "server: workunit restarted too often after buffer overflows");
    fail(s, id, w, FS_E_OVERFLOW);
    return;
  }
  w.cap = cap;
  s->pending.push_front(id);
  s->n_queued.fetch_add(1);
}

// Plan and enqueue one quantum over every admitted workunit (smu held).
int launch_quantum(Server* s, uint32_t par) {
  Quantum& Q = s->q[par];
  Q.plan.clear();
  const uint32_t nj = (uint32_t)s->slots.size();
  Staging& g = s->stg[par];
  size_t nwords = 0;
  uint32_t pg = 0, nmax_max = 0;
  for (uint32_t k = 0; k < nj; ++k) {
    if (s->slots[k] >= 0) {
      const WU& w = s->wus[s->slots[k]];
      if (w.nplan == w.job.nmin) nwords += (w.prog->words.size() + 7) / 8 * 8;
      nmax_max = std::max(nmax_max, w.job.nmax);
    }
  }
  // buffers: grown only with the stream idle (a running quantum reads them)
  const bool grow = nwords > s->E.prog_cap || nmax_max + 64 > s->E.d.ntab_len;
  if (grow) SCHK(hipStreamSynchronize(s->E.st), "server sync");
  int rc;
  if ((rc = s->E.ensure(nj, s->pool, (uint32_t)std::max<size_t>(nwords, 1), s->pool / kTile + 64, nmax_max, 1,
                        FS_BP_WORDS(1))))
    return rc;
  if (nwords > g.words_cap) {
    if ((rc = pinned(&g.words, nwords + nwords / 2))) return rc;
    g.words_cap = nwords + nwords / 2;
  }
  size_t wpos = 0;
  uint32_t norder = 0;
  for (uint32_t k = 0; k < nj; ++k) {
    fs_bjob& b = g.jobs[k];
    memset(&b, 0, sizeof b);
    b.pg_first = pg;
    b.parent = k;
    if (s->slots[k] < 0) continue;
    const int64_t id = s->slots[k];
    WU& w = s->wus[id];
    const bool first = (w.nplan == w.job.nmin);
    b.pmin = w.job.pmin;
    b.pmax = w.job.pmax;
    b.nmin = w.job.nmin;
    b.nmax = w.job.nmax;
    b.path = (uint32_t)w.path;
    b.prog_terms = w.prog->terms;
    b.prog_off = (uint32_t)wpos;
    if (first) {  // candidate generation and setup run in this quantum
      memcpy(g.words + wpos, w.prog->words.data(), w.prog->words.size() * 8);
      wpos += w.prog->words.size();
      while (wpos % 8) g.words[wpos++] = 0;
      b.flags = FS_BJ_NEW;
      pg += fs_pg_blocks(w.job.pmin, w.job.pmax);
    }
    b.cand_off = w.off;
    b.cand_cap = w.cap;
    const uint32_t n0 = w.nplan;
    const uint32_t n1 = (w.job.nmax - n0 > s->o.n_chunk) ? n0 + s->o.n_chunk : w.job.nmax;
    g.order[norder] = k;
    g.chunk[2 * norder] = n0;
    g.chunk[2 * norder + 1] = n1;
    ++norder;
    Q.plan.push_back({k, id, w.gen, n0, n1, first, n1 == w.job.nmax});
    w.nplan = n1;
  }
  Q.live = !Q.plan.empty();
  if (!Q.live) return FS_OK;
  const std::vector<uint32_t> params = fs_batch_params(nj, pg, {norder});
  memcpy(g.params, params.data(), params.size() * 4);
  fs_batch_dev dv = s->E.d;
  dv.fac = s->fac[par];
  dv.fcap = s->fcap[par];
  dv.glob = s->glob[par];
  hipStream_t st = s->E.st;
  SCHK(hipMemcpyAsync(dv.jobs, g.jobs, nj * sizeof(fs_bjob), hipMemcpyHostToDevice, st), "server upload");
  SCHK(hipMemcpyAsync(dv.order, g.order, norder * 4, hipMemcpyHostToDevice, st), "server upload");
  SCHK(hipMemcpyAsync(dv.chunk, g.chunk, 2 * norder * 4, hipMemcpyHostToDevice, st), "server upload");
  if (wpos) SCHK(hipMemcpyAsync(dv.prog, g.words, wpos * 8, hipMemcpyHostToDevice, st), "server upload");
  SCHK(hipMemcpyAsync(dv.params, g.params, params.size() * 4, hipMemcpyHostToDevice, st), "server upload");
  if ((rc = s->E.enqueue(dv, 1, s->pg_grid, kTile, s->blocks, 8, s->o.use_graph != 0))) return rc;
  // counters out before the next quantum clears the slots it reuses
  SCHK(hipMemcpyAsync(g.cnt, dv.jcnt, nj * 4, hipMemcpyDeviceToHost, st), "server read-back");
  SCHK(hipMemcpyAsync(g.flag, dv.jflag, nj * 4, hipMemcpyDeviceToHost, st), "server read-back");
  SCHK(hipMemcpyAsync(g.sum, dv.jsum, nj * 8, hipMemcpyDeviceToHost, st), "server read-back");
  SCHK(hipMemcpyAsync(g.glob, dv.glob, FS_BG_WORDS * 4, hipMemcpyDeviceToHost, st), "server read-back");
  SCHK(hipEventRecord(s->ev[par], st), "server event");
  // workunits that end in this quantum hand their slot and region to the next
  for (const SlotPlan& sp : Q.plan)
    if (sp.last) vacate(s, s->wus[sp.id]);
  return FS_OK;
}

// Harvest the quantum of parity `par` (smu held); returns the number of
// workunits it finished, or an error.
int harvest(Server* s, uint32_t par, int* finished) {
  Quantum& Q = s->q[par];
  *finished = 0;
  if (!Q.live) return FS_OK;
  Q.live = false;
  SCHK(hipEventSynchronize(s->ev[par]), "server harvest");
  Staging& g = s->stg[par];
  const uint32_t nf_all = g.glob[FS_BG_NF];
  const bool fovf = g.glob[FS_BG_FOVF] != 0 || nf_all > s->fcap[par] || fs_take_injected_fault(FS_CTR_FACTOROVF);
  const bool covf_injected = fs_take_injected_fault(FS_CTR_PRIMEOVF);
  if (g.glob[FS_BG_TILEOVF]) {
    fs_set_error_msg("server: tile list overflow");
    return FS_E_OVERFLOW;
  }
  std::vector<fs_factor> fac(std::min(nf_all, s->fcap[par]));
  if (!fac.empty()) {
    SCHK(hipMemcpyAsync(fac.data(), s->fac[par], fac.size() * sizeof(fs_factor), hipMemcpyDeviceToHost, s->hst),
         "server factors");
    SCHK(hipStreamSynchronize(s->hst), "server factors");
  }
  if (fovf) {
    // factors of this quantum were lost: its workunits start again, with a
    // larger factor buffer from now on
    s->fcap_target = std::max<uint32_t>(s->fcap_target, std::min<uint64_t>((uint64_t)nf_all + nf_all / 4 + 4096,
                                                                            0x7fffffffu / sizeof(fs_factor)));
    if (s->fcap_target <= s->fcap[par]) s->fcap_target = s->fcap[par] * 2;
  }
  if (s->fcap[par] < s->fcap_target) {  // this parity's buffer is idle until the quantum after next
    SCHK(hipFree(s->fac[par]), "server factors");
    s->fac[par] = nullptr;
    SCHK(hipMalloc((void**)&s->fac[par], (size_t)s->fcap_target * sizeof(fs_factor)), "server factors");
    s->fcap[par] = s->fcap_target;
  }
  std::vector<int> plan_of(s->slots.size(), -1);
  for (size_t k = 0; k < Q.plan.size(); ++k) plan_of[Q.plan[k].slot] = (int)k;
  auto current = [&](const SlotPlan& sp) -> WU* {
    auto it = s->wus.find(sp.id);
    return (it != s->wus.end() && it->second.gen == sp.gen) ? &it->second : nullptr;
  };
  if (fovf) {
    for (const SlotPlan& sp : Q.plan)
      if (WU* w = current(sp)) restart(s, sp.id, *w, w->cap);
    return FS_OK;
  }
  for (const fs_factor& f : fac) {
    const uint32_t slot = (uint32_t)f.pad;
    if (slot < plan_of.size() && plan_of[slot] >= 0)
      if (WU* w = current(Q.plan[plan_of[slot]])) w->raw.push_back(f);
  }
  for (const SlotPlan& sp : Q.plan) {
    WU* wp = current(sp);
    if (!wp) continue;
    WU& w = *wp;
    if (sp.fresh) {
      const uint32_t cnt = g.cnt[sp.slot];
      if ((g.flag[sp.slot] & FS_BJF_CANDOVF) || cnt > w.cap || covf_injected) {
        // region too small: start again with one that holds the real count
        restart(s, sp.id, w, (uint32_t)std::min<uint64_t>(((uint64_t)cnt + cnt / 8 + 4096 + kTile - 1) / kTile * kTile,
                                                          0xffffffffu));
        continue;
      }
      w.count = cnt;
    }
    if (!sp.last) continue;
    // finished: the checksum fold happened in this quantum
    fs_job_result r;
    memset(&r, 0, sizeof r);
    r.path = w.path;
    r.primecount = w.count;
    r.steps = (uint64_t)w.count * (w.job.nmax - w.job.nmin);
    r.raw_factors = w.raw.size();
    r.status = (g.flag[sp.slot] & FS_BJF_VALIDATE) ? FS_E_VALIDATE : FS_OK;
    uint64_t ck = g.sum[sp.slot];
    std::vector<fs_factor> kept;
    if (r.status == FS_OK) {
      for (auto& x : w.raw) x.pad = 0;
      fs_sort_factors(w.raw.data(), w.raw.size());
      if (fs_verify_select(w.raw.data(), w.raw.size(), s->o.verify_threads, FS_VERIFY_AUTO, (fs_stream)s->hst) !=
          -1) {
        r.status = FS_E_VERIFY;
      } else {
        const size_t n = fs_filter_factors(w.raw.data(), w.raw.size(), &ck, false);
        kept.assign(w.raw.begin(), w.raw.begin() + n);
        r.factorcount = n;
        r.checksum = ck;
      }
    }
    publish(s, sp.id, w, r, std::move(kept));
    ++*finished;
  }
  return FS_OK;
}

}  // namespace

extern "C" void fs_default_server_options(fs_server_options* o) {
  memset(o, 0, sizeof *o);
  o->mem_budget = 0;
  o->max_slots = 256;
  o->n_chunk = kDefaultChunk;
  o->verify_threads = 16;
  o->use_graph = 1;
}

extern "C" int fs_server_create(const fs_server_options* opt, fs_server* out) {
  int rc = fs_init(0);
  if (rc) return rc;
  Server* s = new Server();
  if (opt) s->o = *opt;
  else fs_default_server_options(&s->o);
  if (s->o.max_slots == 0) s->o.max_slots = 256;
  if (s->o.n_chunk == 0) s->o.n_chunk = kDefaultChunk;
  uint64_t freeb = 0;
  fs_device_info(nullptr, 0, nullptr, nullptr, &freeb);
  uint64_t budget = s->o.mem_budget ? s->o.mem_budget : freeb / 2;
  budget = std::min<uint64_t>(budget, (uint64_t)0x7fffffff * kBytesPerCand);
  s->pool = (uint32_t)(budget / kBytesPerCand) / kTile * kTile;
  if (s->pool < kTile) {
    delete s;
    fs_set_error_msg("server: memory budget too small");
    return FS_E_ARG;
  }
  s->regions.init(s->pool);
  const uint32_t nj = s->o.max_slots;
  s->slots.assign(nj, -1);
  s->pg_grid = fs_device_cus() * 8;
  s->blocks = fs_device_cus() * 8;
  // the pool and the tables up front: admission never allocates device memory
  if ((rc = s->E.ensure(nj, s->pool, 4096, s->pool / kTile + 64, 1u << 16, 1, FS_BP_WORDS(1)))) {
    free_server(s);
    return rc;
  }
  auto fail_create = [&](int code) {
    free_server(s);
    return code;
  };
  if (hipStreamCreateWithFlags(&s->hst, hipStreamNonBlocking) != hipSuccess) {
    fs_set_error_msg("server: stream creation failed");
    return fail_create(FS_E_HIP);
  }
  for (int k = 0; k < 2; ++k) {
    Staging& g = s->stg[k];
    if ((rc = pinned(&g.jobs, nj)) || (rc = pinned(&g.order, nj)) || (rc = pinned(&g.chunk, 2 * nj)) ||
        (rc = pinned(&g.params, FS_BP_WORDS(1))) || (rc = pinned(&g.words, 4096)) || (rc = pinned(&g.cnt, nj)) ||
        (rc = pinned(&g.flag, nj)) || (rc = pinned(&g.sum, nj)) || (rc = pinned(&g.glob, FS_BG_WORDS)))
      return fail_create(rc);
    g.words_cap = 4096;
    if (hipEventCreateWithFlags(&s->ev[k], hipEventDisableTiming) != hipSuccess ||
        hipMalloc((void**)&s->fac[k], (size_t)kInitialFactorCap * sizeof(fs_factor)) != hipSuccess ||
        hipMalloc((void**)&s->glob[k], FS_BG_WORDS * 4) != hipSuccess) {
      fs_set_error_msg("server: device allocation failed");
      return fail_create(FS_E_NOMEM);
    }
    s->fcap[k] = kInitialFactorCap;
  }
  *out = (fs_server)s;
  return FS_OK;
}

extern "C" int64_t fs_server_submit(fs_server h, const fs_job* job) {
  Server* s = (Server*)h;
  if (!s || !job) return FS_E_ARG;
  if (fs_validate_ranges(job->pmin, job->pmax, job->nmin, job->nmax) != FS_OK) return FS_E_ARG;
  if (region_len(*job) > s->pool) {
    fs_set_error_msg("server: workunit larger than the device pool");
    return FS_E_BUSY;
  }
  std::lock_guard<std::mutex> lk(s->qmu);
  const int64_t id = s->next_id++;
  s->incoming.emplace_back(id, *job);
  s->live.insert(id);
  s->n_queued.fetch_add(1);
  return id;
}

extern "C" int fs_server_step(fs_server h) {
  Server* s = (Server*)h;
  if (!s) return FS_E_ARG;
  std::lock_guard<std::mutex> lk(s->smu);
  // 1. take the submissions, then admit in order while slots and memory last
  {
    std::lock_guard<std::mutex> q(s->qmu);
    while (!s->incoming.empty()) {
      const auto& in = s->incoming.front();
      WU& w = s->wus[in.first];
      w.job = in.second;
      w.path = fs_select_path(in.second.pmax);
      w.cap = region_len(in.second);
      s->pending.push_back(in.first);
      s->incoming.pop_front();
    }
  }
  while (!s->pending.empty()) {
    const int64_t id = s->pending.front();
    WU& w = s->wus[id];
    if (!w.prog && !(w.prog = fs_program_get(w.job.nmin, w.path))) {
      s->pending.pop_front();
      s->n_queued.fetch_sub(1);
      fail(s, id, w, FS_E_NOMEM);  // host allocation failed: the workunit ends with an error
      continue;
    }
    auto slot = std::find(s->slots.begin(), s->slots.end(), (int64_t)-1);
    if (slot == s->slots.end()) break;
    uint32_t off;
    if (!s->regions.alloc(w.cap, &off)) break;  // wait for memory (FIFO admission)
    s->pending.pop_front();
    s->n_queued.fetch_sub(1);
    *slot = id;
    w.slot = (uint32_t)(slot - s->slots.begin());
    w.off = off;
    w.admitted = true;
    w.nplan = w.job.nmin;
    w.count = 0;
    w.counted = true;
    s->n_active.fetch_add(1);
    s->bytes_in_use.fetch_add((uint64_t)w.cap * kBytesPerCand);
  }
  // 2. enqueue this quantum, 3. harvest the previous one while it runs
  const uint32_t par = (uint32_t)(s->quanta & 1);
  int rc;
  if (s->q[par].live) {  // (only when an error left a quantum unharvested)
    int fin = 0;
    if ((rc = harvest(s, par, &fin))) return rc;
  }
  if ((rc = launch_quantum(s, par))) return rc;
  if (s->q[par].live) ++s->quanta;
  int finished = 0;
  if ((rc = harvest(s, par ^ 1u, &finished))) return rc;
  return finished;
}

extern "C" int fs_server_poll(fs_server h, int64_t id, fs_job_result* out, fs_factor* factors, uint64_t cap,
                              uint64_t* nfactors) {
  Server* s = (Server*)h;
  if (!s) return FS_E_ARG;
  std::lock_guard<std::mutex> lk(s->qmu);
  if (!s->live.count(id)) return FS_E_ARG;
  auto it = s->done.find(id);
  if (it == s->done.end()) return 0;
  Done& d = it->second;
  if (out) *out = d.res;
  if (nfactors) *nfactors = d.kept.size();
  if (factors) memcpy(factors, d.kept.data(), std::min<uint64_t>(cap, d.kept.size()) * sizeof(fs_factor));
  s->done.erase(it);
  s->live.erase(id);
  return 1;
}

extern "C" int fs_server_stats(fs_server h, uint32_t* queued, uint32_t* active, uint32_t* done,
                               uint64_t* bytes_in_use) {
  Server* s = (Server*)h;
  if (!s) return FS_E_ARG;
  if (queued) *queued = s->n_queued.load();
  if (active) *active = s->n_active.load();
  if (done) *done = s->n_done.load();
  if (bytes_in_use) *bytes_in_use = s->bytes_in_use.load();
  return FS_OK;
}

extern "C" int fs_server_destroy(fs_server h) {
  Server* s = (Server*)h;
  if (!s) return FS_OK;
  { std::lock_guard<std::mutex> lk(s->smu); }  // a step still running on another thread ends first
  free_server(s);
  return FS_OK;
}
