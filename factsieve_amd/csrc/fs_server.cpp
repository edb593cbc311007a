// fs_server.cpp -- continuous-batching serving loop for sieve workunits.
//
// Workunits (fs_job) are submitted at any time and queue on the host.  Each
// call of fs_server_step is one scheduling quantum:
//   1. admit queued workunits while a job slot and a candidate region of the
//      device pool are free (admission by device memory);
//   2. one batched launch over every active workunit: new ones generate their
//      candidates and run setup, all of them advance their residues by the next
//      n-chunk (residues stay resident in their candidate regions between
//      quanta), and those reaching nmax fold their checksum;
//   3. harvest: factors are attributed to their workunits, finished workunits
//      are verified and filtered on the CPU, and their slot and region freed.
// A workunit's result does not depend on what it shared the device with: the
// kernels touch only its own region, counters and factor tags.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <mutex>
#include <vector>
#include "fs_batch_internal.hpp"
#include "fs_internal.hpp"
#include "../../include/factsieve.h"

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct WU {
  fs_job job{};
  int state = 0;  // 0 queued, 1 active, 2 done
  int path = 0;
  uint32_t terms = 0;
  std::shared_ptr<const fs_program> prog;
  uint32_t cap = 0, off = 0, slot = 0;
  uint32_t nnext = 0;
  bool fresh = false;
  uint32_t count = 0;
  std::vector<fs_factor> raw;
  fs_job_result res{};
  std::vector<fs_factor> kept;
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
  uint32_t largest() const {
    uint32_t m = 0;
    for (auto& kv : free_) m = std::max(m, kv.second);
    return m;
  }
};

struct Server {
  fs_server_options o{};
  fs_batch_bufs E;
  std::mutex mu;
  std::map<int64_t, WU> wus;
  std::deque<int64_t> queue;
  std::vector<int64_t> slots;  // slot -> workunit id or -1
  Regions regions;
  uint32_t pool = 0;           // candidate pool size (records)
  int64_t next_id = 0;
  uint32_t done = 0;
  static constexpr uint32_t kTile = 256;
  static constexpr uint32_t kBytesPerCand = 8 * 4 + 4;  // p, q, r2, res, nnext
};

uint32_t region_len(const fs_job& j) {
  const uint64_t c = fs_cand_capacity(j.pmin, j.pmax - j.pmin);
  return (uint32_t)std::min<uint64_t>((c + Server::kTile - 1) / Server::kTile * Server::kTile, 0x7fffffffull);
}

}  // namespace

extern "C" void fs_default_server_options(fs_server_options* o) {
  memset(o, 0, sizeof *o);
  o->mem_budget = 0;
  o->max_slots = 256;
  o->n_chunk = 8192;
  o->verify_threads = 16;
}

extern "C" int fs_server_create(const fs_server_options* opt, fs_server* out) {
  int rc = fs_init(0);
  if (rc) return rc;
  Server* s = new Server();
  if (opt) s->o = *opt;
  else fs_default_server_options(&s->o);
  if (s->o.max_slots == 0) s->o.max_slots = 256;
  if (s->o.n_chunk == 0) s->o.n_chunk = 8192;
  uint64_t freeb = 0;
  fs_device_info(nullptr, 0, nullptr, nullptr, &freeb);
  uint64_t budget = s->o.mem_budget ? s->o.mem_budget : freeb / 2;
  budget = std::min<uint64_t>(budget, (uint64_t)0x7fffffff * Server::kBytesPerCand);
  s->pool = (uint32_t)(budget / Server::kBytesPerCand) / Server::kTile * Server::kTile;
  if (s->pool < Server::kTile) {
    delete s;
    fs_set_error_msg("server: memory budget too small");
    return FS_E_ARG;
  }
  s->regions.init(s->pool);
  s->slots.assign(s->o.max_slots, -1);
  // allocate the pool up front: admission never allocates device memory
  if ((rc = s->E.ensure(s->o.max_slots, s->pool, 4096, s->pool / Server::kTile + 64, 1u << 16, 1u << 20))) {
    delete s;
    return rc;
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
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t id = s->next_id++;
  WU& w = s->wus[id];
  w.job = *job;
  w.path = fs_select_path(job->pmax);
  w.cap = region_len(*job);
  s->queue.push_back(id);
  return id;
}

extern "C" int fs_server_step(fs_server h) {
  Server* s = (Server*)h;
  if (!s) return FS_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  // 1. admission
  while (!s->queue.empty()) {
    const int64_t id = s->queue.front();
    WU& w = s->wus[id];
    if (!w.prog && !(w.prog = fs_program_get(w.job.nmin, w.path))) {
      memset(&w.res, 0, sizeof w.res);  // host allocation failed: the workunit ends with an error
      w.res.status = FS_E_NOMEM;
      w.res.path = w.path;
      w.state = 2;
      ++s->done;
      s->queue.pop_front();
      continue;
    }
    w.terms = w.prog->terms;
    auto slot = std::find(s->slots.begin(), s->slots.end(), (int64_t)-1);
    if (slot == s->slots.end()) break;
    uint32_t off;
    if (!s->regions.alloc(w.cap, &off)) break;  // wait for memory (FIFO admission)
    s->queue.pop_front();
    *slot = id;
    w.slot = (uint32_t)(slot - s->slots.begin());
    w.off = off;
    w.state = 1;
    w.fresh = true;
    w.nnext = w.job.nmin;
  }
  // 2. one quantum over all active workunits
  const uint32_t nj = (uint32_t)s->slots.size();
  std::vector<fs_bjob> bj(nj);
  std::vector<uint32_t> order, chunk;
  std::vector<uint64_t> words;
  uint32_t pg = 0;
  for (uint32_t k = 0; k < nj; ++k) {
    fs_bjob& b = bj[k];
    memset(&b, 0, sizeof b);
    b.pg_first = pg;
    b.parent = k;
    if (s->slots[k] < 0) continue;
    WU& w = s->wus[s->slots[k]];
    b.pmin = w.job.pmin;
    b.pmax = w.job.pmax;
    b.nmin = w.job.nmin;
    b.nmax = w.job.nmax;
    b.path = (uint32_t)w.path;
    b.prog_off = (uint32_t)words.size();
    b.prog_terms = w.terms;
    if (w.nnext == w.job.nmin) {  // setup runs in this quantum
      words.insert(words.end(), w.prog->words.begin(), w.prog->words.end());
      while (words.size() % 8) words.push_back(0);
    }
    b.cand_off = w.off;
    b.cand_cap = w.cap;
    b.flags = w.fresh ? FS_BJ_NEW : 0;
    if (w.fresh) pg += fs_pg_blocks(w.job.pmin, w.job.pmax);
    order.push_back(k);
    const uint32_t n0 = w.nnext;
    const uint32_t n1 = (w.job.nmax - n0 > s->o.n_chunk) ? n0 + s->o.n_chunk : w.job.nmax;
    chunk.push_back(n0);
    chunk.push_back(n1);
  }
  if (order.empty()) return 0;
  float ms = 0;
  int rc = fs_batch_run_device(s->E, bj, order, chunk, words, pg, s->pool, Server::kTile,
                               fs_device_cus() * 8, s->E.d.fcap, false, &ms);
  if (rc) return rc;
  std::vector<uint32_t> cnt, flag;
  std::vector<uint64_t> sum;
  std::vector<fs_factor> fac;
  uint32_t glob[FS_BG_WORDS];
  if ((rc = fs_batch_read(s->E, nj, cnt, flag, sum, fac, glob))) return rc;
  if (glob[FS_BG_FOVF]) {
    fs_set_error_msg("server: factor buffer overflow in one quantum");
    return FS_E_OVERFLOW;
  }
  // 3. harvest
  for (const fs_factor& f : fac) {
    if ((uint32_t)f.pad < nj && s->slots[f.pad] >= 0) s->wus[s->slots[f.pad]].raw.push_back(f);
  }
  int finished = 0;
  for (size_t k = 0; k < order.size(); ++k) {
    const uint32_t slot = order[k];
    WU& w = s->wus[s->slots[slot]];
    if (w.fresh && (flag[slot] & FS_BJF_CANDOVF)) {
      // region too small: give it back and queue the workunit again, larger
      s->regions.release(w.off, w.cap);
      w.cap = std::min<uint32_t>(s->pool, (cnt[slot] + cnt[slot] / 8 + 4096 + Server::kTile - 1) / Server::kTile *
                                              Server::kTile);
      w.raw.clear();
      w.state = 0;
      s->queue.push_front(s->slots[slot]);
      s->slots[slot] = -1;
      continue;
    }
    if (w.fresh) w.count = cnt[slot];
    w.fresh = false;
    w.nnext = chunk[2 * k + 1];
    if (w.nnext < w.job.nmax) continue;
    // finished: checksum fold happened in this quantum
    fs_job_result& r = w.res;
    memset(&r, 0, sizeof r);
    r.path = w.path;
    r.primecount = w.count;
    r.steps = (uint64_t)w.count * (w.job.nmax - w.job.nmin);
    r.raw_factors = w.raw.size();
    r.status = (flag[slot] & FS_BJF_VALIDATE) ? FS_E_VALIDATE : FS_OK;
    uint64_t ck = sum[slot];
    if (r.status == FS_OK) {
      for (auto& x : w.raw) x.pad = 0;
      fs_sort_factors(w.raw.data(), w.raw.size());
      if (fs_verify_select(w.raw.data(), w.raw.size(), s->o.verify_threads, FS_VERIFY_AUTO, (fs_stream)s->E.st) !=
          -1) {
        r.status = FS_E_VERIFY;
      } else {
        const size_t kept = fs_filter_factors(w.raw.data(), w.raw.size(), &ck, false);
        w.kept.assign(w.raw.begin(), w.raw.begin() + kept);
        r.factorcount = kept;
        r.checksum = ck;
      }
    }
    w.raw.clear();
    w.raw.shrink_to_fit();
    w.state = 2;
    s->regions.release(w.off, w.cap);
    s->slots[slot] = -1;
    ++s->done;
    ++finished;
  }
  return finished;
}

extern "C" int fs_server_poll(fs_server h, int64_t id, fs_job_result* out, fs_factor* factors, uint64_t cap,
                              uint64_t* nfactors) {
  Server* s = (Server*)h;
  if (!s) return FS_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->wus.find(id);
  if (it == s->wus.end()) return FS_E_ARG;
  WU& w = it->second;
  if (w.state != 2) return 0;
  if (out) *out = w.res;
  if (nfactors) *nfactors = w.kept.size();
  if (factors)
    memcpy(factors, w.kept.data(), std::min<uint64_t>(cap, w.kept.size()) * sizeof(fs_factor));
  s->wus.erase(it);
  return 1;
}

extern "C" int fs_server_stats(fs_server h, uint32_t* queued, uint32_t* active, uint32_t* done,
                               uint64_t* bytes_in_use) {
  Server* s = (Server*)h;
  if (!s) return FS_E_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  uint32_t act = 0;
  uint64_t used = 0;
  for (int64_t id : s->slots) {
    if (id < 0) continue;
    ++act;
    used += (uint64_t)s->wus[id].cap * Server::kBytesPerCand;
  }
  if (queued) *queued = (uint32_t)s->queue.size();
  if (active) *active = act;
  if (done) *done = s->done;
  if (bytes_in_use) *bytes_in_use = used;
  return FS_OK;
}

extern "C" int fs_server_destroy(fs_server h) {
  Server* s = (Server*)h;
  if (!s) return FS_OK;
  s->E.release();
  delete s;
  return FS_OK;
}
