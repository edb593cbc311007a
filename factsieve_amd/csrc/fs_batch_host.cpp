// fs_batch_host.cpp -- host side of the batched persistent form: job planning,
// setup-program cache, device buffers, launch (plain or hipGraph), retries on
// buffer overflow, per-job drain, and the autotuner of the launch shape.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>
#include <algorithm>
#include <chrono>
#include <map>
#include <atomic>
#include <mutex>
#include <vector>
#include "fs_batch_internal.hpp"
#include "fs_internal.hpp"
#include "../../include/factsieve.h"

#define HCHK(call, what)                        \
  do {                                          \
    hipError_t e__ = (call);                    \
    if (e__ != hipSuccess) {                    \
      fs_set_hip_error(what, (int)e__);         \
      return FS_E_HIP;                          \
    }                                           \
  } while (0)

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Relative cost of one p*n step and one setup term per path (VALU operations
// per step of the respective kernels), used to order tiles longest first.
double step_cost(int path) { return path == FS_PATH_FP64 ? 7.0 : path == FS_PATH_HYBRID ? 12.25 : 45.0; }
double setup_cost(int path) { return path == FS_PATH_FP64 ? 6.55 : path == FS_PATH_HYBRID ? 18.0 : 90.0; }

}  // namespace

namespace {
template <class T>
int grow(T** p, size_t want, const char* what) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  hipError_t e = hipMalloc((void**)p, want * sizeof(T) + 64);
  if (e != hipSuccess) {
    fs_set_hip_error(what, (int)e);
    return FS_E_NOMEM;
  }
  return FS_OK;
}
}  // namespace

int fs_batch_bufs::ensure(uint32_t nj, uint32_t ncands, uint32_t nprog, uint32_t ntiles, uint32_t nmax_max,
                          uint32_t fcap, uint32_t nparams) {
  int rc;
  if (!st) {
    HCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking), "batch stream");
    HCHK(hipEventCreate(&ev0), "batch event");
    HCHK(hipEventCreate(&ev1), "batch event");
  }
  if (nj > jobs_cap || !d.jobs) {
    const uint32_t n = std::max<uint32_t>(nj, 64);
    if ((rc = grow(&d.jobs, n, "batch jobs")) || (rc = grow(&d.order, n, "batch order")) ||
        (rc = grow(&d.chunk, 2 * n, "batch chunk")) || (rc = grow(&d.jcnt, n, "batch jcnt")) ||
        (rc = grow(&d.jflag, n, "batch jflag")) || (rc = grow(&d.jsum, n, "batch jsum")) ||
        (rc = grow(&d.pref, n + 1, "batch pref")))
      return rc;
    jobs_cap = n;
  }
  if (!d.glob && (rc = grow(&d.glob, FS_BG_WORDS, "batch glob"))) return rc;
  if (ncands > cands_cap || !d.c.p) {
    if (d.c.p) fs_cands_free(&d.c);
    if ((rc = fs_cands_alloc(&d.c, std::max<uint32_t>(ncands, 4096)))) return rc;
    cands_cap = d.c.cap;
  }
  if (nprog > prog_cap || !d.prog) {
    if ((rc = grow(&d.prog, std::max<uint32_t>(nprog, 4096), "batch program"))) return rc;
    prog_cap = std::max<uint32_t>(nprog, 4096);
  }
  if (ntiles > d.tiles_cap || !d.tiles) {
    const uint32_t n = std::max<uint32_t>(ntiles, 4096);
    if ((rc = grow(&d.tiles, n, "batch tiles"))) return rc;
    d.tiles_cap = n;
  }
  const uint32_t nt_need = nmax_max + 64;
  if (nt_need > d.ntab_len || !d.ntab) {
    const uint32_t n = std::max<uint32_t>(nt_need, 1u << 16);
    if ((rc = grow(&d.ntab, fs_bntab_doubles(n), "batch ntab")) || (rc = fs_batch_fill_ntab(d.ntab, n, st)))
      return rc;
    d.ntab_len = n;
    d.pair_len = fs_bpair_len(n);
  }
  if (fcap > d.fcap || !d.fac) {
    if ((rc = grow(&d.fac, fcap, "batch factors"))) return rc;
    d.fcap = fcap;
  }
  if (nparams > d.params_cap || !d.params) {
    const uint32_t n = std::max<uint32_t>(nparams, 64);
    if ((rc = grow(&d.params, n, "batch params"))) return rc;
    d.params_cap = n;
  }
  return FS_OK;
}

void fs_batch_bufs::drop_graphs() {
  for (auto& g : graphs)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  graphs.clear();
}

int fs_batch_bufs::enqueue(const fs_batch_dev& dv, uint32_t rounds, uint32_t pg_grid, uint32_t tile,
                           uint32_t blocks, uint32_t unroll, bool graph) {
  if (!graph) return fs_batch_launch(dv, jobs_cap, rounds, pg_grid, tile, blocks, unroll, st);
  // everything a launch bakes into its kernel arguments and grids
  std::vector<uint64_t> key = {
      (uint64_t)(uintptr_t)dv.jobs, (uint64_t)(uintptr_t)dv.order, (uint64_t)(uintptr_t)dv.chunk,
      (uint64_t)(uintptr_t)dv.jcnt, (uint64_t)(uintptr_t)dv.jflag, (uint64_t)(uintptr_t)dv.jsum,
      (uint64_t)(uintptr_t)dv.glob, (uint64_t)(uintptr_t)dv.pref, (uint64_t)(uintptr_t)dv.tiles, dv.tiles_cap,
      (uint64_t)(uintptr_t)dv.c.p, (uint64_t)(uintptr_t)dv.c.q, (uint64_t)(uintptr_t)dv.c.r2,
      (uint64_t)(uintptr_t)dv.c.res, (uint64_t)(uintptr_t)dv.c.nnext, dv.c.cap, dv.c.grid_hint,
      (uint64_t)(uintptr_t)dv.prog, (uint64_t)(uintptr_t)dv.ntab, dv.ntab_len, dv.pair_len,
      (uint64_t)(uintptr_t)dv.fac, dv.fcap, (uint64_t)(uintptr_t)dv.params, jobs_cap, rounds, pg_grid, tile,
      blocks, unroll};
  for (size_t k = 0; k < graphs.size(); ++k) {
    if (graphs[k].key == key) {
      HCHK(hipGraphLaunch(graphs[k].exec, st), "batch graph launch");
      ++graph_replays;
      return FS_OK;
    }
  }
  hipGraph_t g;
  HCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "batch capture");
  int rc = fs_batch_launch(dv, jobs_cap, rounds, pg_grid, tile, blocks, unroll, st);
  hipError_t ce = hipStreamEndCapture(st, &g);
  if (rc) return rc;
  HCHK(ce, "batch end capture");
  hipGraphExec_t ge;
  hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HCHK(ie, "batch graph instantiate");
  if (graphs.size() >= 4) {  // keep the most recent few shapes
    (void)hipGraphExecDestroy(graphs.front().exec);
    graphs.erase(graphs.begin());
  }
  graphs.push_back({key, ge});
  ++graph_captures;
  HCHK(hipGraphLaunch(ge, st), "batch graph launch");
  return FS_OK;
}

std::vector<uint32_t> fs_batch_params(uint32_t nj, uint32_t pg_blocks, const std::vector<uint32_t>& rounds) {
  std::vector<uint32_t> p(FS_BP_WORDS(rounds.size()));
  p[FS_BP_NJ] = nj;
  p[FS_BP_PGBLOCKS] = pg_blocks;
  uint32_t first = 0;
  for (size_t r = 0; r < rounds.size(); ++r) {
    p[FS_BP_ROUND0 + 2 * r] = first;
    p[FS_BP_ROUND0 + 2 * r + 1] = rounds[r];
    first += rounds[r];
  }
  return p;
}

void fs_batch_bufs::release() {
  if (st) (void)hipStreamSynchronize(st);
  drop_graphs();
  if (d.c.p) fs_cands_free(&d.c);
  void* ps[] = {d.jobs, d.order, d.chunk, d.jcnt, d.jflag, d.jsum, d.glob, d.pref, d.tiles, d.prog, d.ntab, d.fac,
                d.params};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  d = fs_batch_dev();
  jobs_cap = cands_cap = prog_cap = 0;
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (st) (void)hipStreamDestroy(st);
  st = nullptr;
  ev0 = ev1 = nullptr;
}

int fs_batch_run_device(fs_batch_bufs& E, const std::vector<fs_bjob>& bj, const std::vector<uint32_t>& order,
                        const std::vector<uint32_t>& chunk, const std::vector<uint64_t>& words, uint32_t pg_blocks,
                        uint32_t ncands, uint32_t tile, uint32_t blocks, uint32_t fcap, bool graph, float* dev_ms,
                        const std::vector<uint32_t>& rounds, uint32_t unroll) {
  const uint32_t nj = (uint32_t)bj.size();
  std::vector<uint32_t> norder = rounds;
  if (norder.empty()) norder.push_back((uint32_t)order.size());
  uint32_t nmax_max = 0, ntiles = 0, k = 0;
  for (uint32_t r : norder) {  // the tile list holds one round at a time
    uint32_t t = 0;
    for (uint32_t e = 0; e < r; ++e, ++k) {
      const fs_bjob& b = bj[order[k]];
      nmax_max = std::max(nmax_max, b.nmax);
      t += (b.cand_cap + tile - 1) / tile;
    }
    ntiles = std::max(ntiles, t);
  }
  if (k != order.size() || 2 * order.size() != chunk.size()) {
    fs_set_error_msg("batch: rounds do not cover the tile order");
    return FS_E_ARG;
  }
  const std::vector<uint32_t> params = fs_batch_params(nj, pg_blocks, norder);
  int rc = E.ensure(std::max<uint32_t>(nj, (uint32_t)order.size()), ncands, (uint32_t)words.size(), ntiles,
                    nmax_max, fcap, (uint32_t)params.size());
  if (rc) return rc;
  fs_batch_dev& d = E.d;
  HCHK(hipMemcpyAsync(d.jobs, bj.data(), nj * sizeof(fs_bjob), hipMemcpyHostToDevice, E.st), "batch upload");
  if (!order.empty()) {
    HCHK(hipMemcpyAsync(d.order, order.data(), order.size() * 4, hipMemcpyHostToDevice, E.st), "batch upload");
    HCHK(hipMemcpyAsync(d.chunk, chunk.data(), chunk.size() * 4, hipMemcpyHostToDevice, E.st), "batch upload");
  }
  if (!words.empty())
    HCHK(hipMemcpyAsync(d.prog, words.data(), words.size() * 8, hipMemcpyHostToDevice, E.st), "batch upload");
  HCHK(hipMemcpyAsync(d.params, params.data(), params.size() * 4, hipMemcpyHostToDevice, E.st), "batch upload");
  HCHK(hipStreamSynchronize(E.st), "batch upload sync");
  HCHK(hipEventRecord(E.ev0, E.st), "batch event");
  // one workgroup per virtual block: the dispatcher balances jobs of unequal
  // cost better than a grid-stride over a fixed grid
  const uint32_t pg_grid = pg_blocks;
  if ((rc = E.enqueue(d, (uint32_t)norder.size(), pg_grid, tile, blocks, unroll, graph))) return rc;
  HCHK(hipEventRecord(E.ev1, E.st), "batch event");
  HCHK(hipStreamSynchronize(E.st), "batch sync");
  if (dev_ms) HCHK(hipEventElapsedTime(dev_ms, E.ev0, E.ev1), "batch time");
  return FS_OK;
}

int fs_batch_read(fs_batch_bufs& E, uint32_t nj, std::vector<uint32_t>& cnt, std::vector<uint32_t>& flag,
                  std::vector<uint64_t>& sum, std::vector<fs_factor>& fac, uint32_t* glob) {
  fs_batch_dev& d = E.d;
  cnt.resize(nj);
  flag.resize(nj);
  sum.resize(nj);
  HCHK(hipMemcpy(glob, d.glob, FS_BG_WORDS * 4, hipMemcpyDeviceToHost), "batch read");
  if (nj) {
    HCHK(hipMemcpy(cnt.data(), d.jcnt, nj * 4, hipMemcpyDeviceToHost), "batch read");
    HCHK(hipMemcpy(flag.data(), d.jflag, nj * 4, hipMemcpyDeviceToHost), "batch read");
    HCHK(hipMemcpy(sum.data(), d.jsum, nj * 8, hipMemcpyDeviceToHost), "batch read");
  }
  const uint32_t nf = std::min(glob[FS_BG_NF], d.fcap);
  fac.resize(nf);
  if (nf) HCHK(hipMemcpy(fac.data(), d.fac, (size_t)nf * sizeof(fs_factor), hipMemcpyDeviceToHost), "batch read");
  if (glob[FS_BG_TILEOVF]) {
    fs_set_error_msg("batch: tile list overflow");
    return FS_E_OVERFLOW;
  }
  return FS_OK;
}

uint32_t fs_pg_blocks(uint64_t pmin, uint64_t pmax) {
  const uint64_t base = pmin | 1;
  if (base >= pmax) return 0;
  const uint64_t threads = (pmax - base + 127) / 128;
  return (uint32_t)((threads + FS_BLOCK - 1) / FS_BLOCK);
}

fs_batch_bufs& fs_batch_engine() {
  static fs_batch_bufs b;
  return b;
}

uint32_t fs_device_cus() {
  static std::atomic<uint32_t> cached{0};  // callers on several host threads may race to fill it
  uint32_t cus = cached.load();
  if (!cus) {
    if (fs_device_info(nullptr, 0, &cus, nullptr, nullptr) != FS_OK || !cus) cus = 256;
    cached.store(cus);
  }
  return cus;
}

extern "C" void fs_default_batch_options(fs_batch_options* o) {
  memset(o, 0, sizeof *o);
  o->tile_candidates = 256;
  o->persistent_blocks = 0;
  o->use_graph = 1;
  o->lpt_order = 1;
  o->verify_threads = 16;
  o->lanes_per_prime = 1;
  o->n_chunk = 0;
  o->unroll = 8;
}

namespace {

// One n-subrange of a job: [n0, n1) from the setup program of n0.
struct Sub {
  uint32_t n0, n1;
  std::shared_ptr<const fs_program> prog;
  double cost;
};

struct Plan {
  int status = FS_OK;
  int path = 0;
  std::vector<Sub> sub;  // one entry unless the job is split over several lanes per prime
  uint32_t cap = 0;
  uint32_t pg_blocks = 0;
};

// Shortest n-subrange a job is cut into when split over several lanes.
constexpr uint32_t kMinSubrange = 1024;


// One batch through the device; results per job (before the host filter).
struct DevOut {
  std::vector<uint32_t> cnt, flag;
  std::vector<uint64_t> sum;
  std::vector<fs_factor> fac;
};

int run_batch(fs_batch_bufs& E, const fs_job* jobs, uint32_t nj, std::vector<Plan>& plan, const fs_batch_options& o,
              bool lpt, uint32_t fcap, DevOut& out, float* dev_ms) {
  const uint32_t tile = o.tile_candidates ? o.tile_candidates : 256;
  std::vector<fs_bjob> bj(nj);
  std::vector<uint64_t> words;
  uint32_t cand_off = 0, pg = 0;
  // Entries: job j at index j (its first subrange), later subranges appended.
  struct Ent {
    uint32_t idx, job, sub;
  };
  std::vector<Ent> ents;
  auto add_prog = [&](fs_bjob& b, const Sub& s) {
    b.prog_off = (uint32_t)words.size();
    b.prog_terms = s.prog->terms;
    words.insert(words.end(), s.prog->words.begin(), s.prog->words.end());
    while (words.size() % 8) words.push_back(0);  // 64-byte aligned programs
  };
  for (uint32_t j = 0; j < nj; ++j) {
    fs_bjob& b = bj[j];
    memset(&b, 0, sizeof b);
    b.pg_first = pg;
    b.parent = j;
    if (plan[j].status != FS_OK) continue;
    const Plan& P = plan[j];
    b.pmin = jobs[j].pmin;
    b.pmax = jobs[j].pmax;
    b.nmin = P.sub[0].n0;
    b.nmax = P.sub[0].n1;
    b.path = (uint32_t)P.path;
    add_prog(b, P.sub[0]);
    b.cand_off = cand_off;
    b.cand_cap = P.cap;
    b.flags = FS_BJ_NEW | (P.sub.size() > 1 ? FS_BJ_NOFOLD : 0);
    cand_off += P.cap;
    pg += P.pg_blocks;
    ents.push_back({j, j, 0});
  }
  for (uint32_t j = 0; j < nj; ++j) {  // later subranges: no candidate generation of their own
    if (plan[j].status != FS_OK) continue;
    const Plan& P = plan[j];
    for (uint32_t k = 1; k < P.sub.size(); ++k) {
      fs_bjob b = bj[j];
      b.pg_first = pg;
      b.nmin = P.sub[k].n0;
      b.nmax = P.sub[k].n1;
      add_prog(b, P.sub[k]);
      b.flags = (k + 1 < P.sub.size()) ? FS_BJ_NOFOLD : 0;
      b.parent = j;
      ents.push_back({(uint32_t)bj.size(), j, k});
      bj.push_back(b);
    }
  }
  // Rounds: a whole subrange per entry, or (n_chunk, unsplit jobs) chunk r of
  // each entry whose range reaches that far; each round in cost order.
  std::vector<uint32_t> order, chunk, rounds;
  const uint32_t C = o.n_chunk;
  for (uint32_t r = 0;; ++r) {
    std::vector<std::pair<double, uint32_t>> cur;  // (cost, entry)
    std::vector<std::pair<uint32_t, uint32_t>> rng(ents.size());
    for (uint32_t e = 0; e < ents.size(); ++e) {
      const Plan& P = plan[ents[e].job];
      const Sub& s = P.sub[ents[e].sub];
      const bool chunked = C && P.sub.size() == 1;
      if (!chunked && r > 0) continue;
      const uint64_t n0 = chunked ? (uint64_t)s.n0 + (uint64_t)r * C : s.n0;
      if (r > 0 && n0 >= s.n1) continue;
      const uint32_t n1 = chunked ? (uint32_t)std::min<uint64_t>(n0 + C, s.n1) : s.n1;
      rng[e] = {(uint32_t)n0, n1};
      const double frac = s.n1 > s.n0 ? (double)(n1 - n0) / (double)(s.n1 - s.n0) : 1.0;
      cur.push_back({s.cost * frac, e});
    }
    if (cur.empty()) break;
    if (lpt)
      std::stable_sort(cur.begin(), cur.end(), [](const std::pair<double, uint32_t>& a,
                                                  const std::pair<double, uint32_t>& b) { return a.first > b.first; });
    for (auto& c : cur) {
      order.push_back(ents[c.second].idx);
      chunk.push_back(rng[c.second].first);
      chunk.push_back(rng[c.second].second);
    }
    rounds.push_back((uint32_t)cur.size());
  }
  if (order.empty()) rounds.clear();
  const uint32_t blocks = o.persistent_blocks ? o.persistent_blocks : fs_device_cus() * 8;
  const uint32_t unroll = (o.unroll == 2 || o.unroll == 4) ? o.unroll : 8;
  int rc = fs_batch_run_device(E, bj, order, chunk, words, pg, cand_off, tile, blocks, fcap, o.use_graph != 0, dev_ms,
                               rounds.empty() ? std::vector<uint32_t>{0} : rounds, unroll);
  if (rc) return rc;
  uint32_t glob[FS_BG_WORDS];
  if ((rc = fs_batch_read(E, nj, out.cnt, out.flag, out.sum, out.fac, glob))) return rc;
  if (glob[FS_BG_FOVF] || glob[FS_BG_NF] > E.d.fcap) {
    out.fac.assign(glob[FS_BG_NF], fs_factor{});
    fs_set_error_msg("batch: factor buffer overflow");
    return FS_E_OVERFLOW;
  }
  return FS_OK;
}

int plan_jobs(const fs_job* jobs, uint32_t nj, const fs_batch_options& o, std::vector<Plan>& plan) {
  const uint32_t tile = o.tile_candidates ? o.tile_candidates : 256;
  const uint32_t lanes = std::max<uint32_t>(1, std::min<uint32_t>(o.lanes_per_prime, 8));
  plan.assign(nj, Plan());
  for (uint32_t j = 0; j < nj; ++j) {
    Plan& P = plan[j];
    const fs_job& J = jobs[j];
    if (fs_validate_ranges(J.pmin, J.pmax, J.nmin, J.nmax) != FS_OK) {
      P.status = FS_E_ARG;
      continue;
    }
    P.path = fs_select_path(J.pmax);
    const uint32_t len = J.nmax - J.nmin;
    const uint32_t L = std::max<uint32_t>(1, std::min<uint32_t>(lanes, len / kMinSubrange));
    for (uint32_t k = 0; k < L; ++k) {
      Sub s;
      s.n0 = J.nmin + (uint32_t)((uint64_t)len * k / L);
      s.n1 = J.nmin + (uint32_t)((uint64_t)len * (k + 1) / L);
      s.prog = fs_program_get(s.n0, P.path);
      if (!s.prog) break;
      s.cost = (double)(s.n1 - s.n0) * step_cost(P.path) + (double)s.prog->terms * setup_cost(P.path);
      P.sub.push_back(s);
    }
    if (P.sub.size() != L) {
      P.status = FS_E_NOMEM;
      continue;
    }
    const uint64_t c = fs_cand_capacity(J.pmin, J.pmax - J.pmin);
    if (c > 0x7fffffffull) {
      fs_set_error_msg("batch: job p-range too large for one batch");
      P.status = FS_E_ARG;
      continue;
    }
    P.cap = (uint32_t)((c + tile - 1) / tile * tile);
    P.pg_blocks = fs_pg_blocks(J.pmin, J.pmax);
  }
  return FS_OK;
}

}  // namespace

// Sorts each listed job's factors and verifies all of them in one pass (the
// CPU threads, or the GPU when the pass needs more than 2^24 multiplications;
// factors of different jobs share it).  ok[j] = 0 for a job with a factor that
// fails verification, or for every listed job when the verifier itself fails.
static void verify_jobs(std::vector<std::vector<fs_factor>>& per, const std::vector<uint32_t>& list, int threads,
                        hipStream_t st, std::vector<char>& ok) {
  std::vector<fs_factor> all;
  for (uint32_t j : list) {
    for (auto& x : per[j]) x.pad = 0;
    fs_sort_factors(per[j].data(), per[j].size());
    all.insert(all.end(), per[j].begin(), per[j].end());
  }
  const int64_t r = fs_verify_select(all.data(), all.size(), threads, FS_VERIFY_AUTO, (fs_stream)st);
  if (r == -1) return;
  for (uint32_t j : list) {  // a failure: find the job(s) it belongs to
    const int64_t bad = r < -1 ? r : fs_verify_all(per[j].data(), per[j].size(), threads);
    if (bad >= 0) {
      char buf[160];
      snprintf(buf, sizeof buf, "CPU factor verification failed for p=%llu", (unsigned long long)per[j][bad].p);
      fs_set_error_msg(buf);
      ok[j] = 0;
    } else if (bad < -1) {
      ok[j] = 0;  // the verifier's own error stays in fs_last_error
    }
  }
}

extern "C" int fs_batch(const fs_job* jobs, uint32_t nj, const fs_batch_options* opt, fs_job_result* out,
                        fs_factor* factors, uint64_t cap, uint64_t* per_job_nf, double* seconds) {
  static std::mutex mu;
  std::lock_guard<std::mutex> lk(mu);
  const double t0 = now_s();
  fs_batch_options o;
  if (opt) o = *opt;
  else fs_default_batch_options(&o);
  if (o.tile_candidates == 0 || o.tile_candidates % FS_BLOCK) {
    fs_set_error_msg("batch: tile_candidates must be a positive multiple of 256");
    return FS_E_ARG;
  }
  if (nj == 0) return FS_OK;
  int rc = fs_init(0) == FS_OK ? FS_OK : FS_E_NODEV;
  if (rc) return rc;
  std::vector<Plan> plan;
  plan_jobs(jobs, nj, o, plan);
  fs_batch_bufs& E = fs_batch_engine();
  DevOut dev;
  float ms = 0;
  uint32_t fcap = 1u << 20;
  for (int attempt = 0;; ++attempt) {
    rc = run_batch(E, jobs, nj, plan, o, o.lpt_order != 0, fcap, dev, &ms);
    if (rc == FS_E_OVERFLOW && !dev.fac.empty() && attempt < 3 && dev.fac.size() < (1u << 28)) {
      fcap = (uint32_t)dev.fac.size() + dev.fac.size() / 4 + 1024;  // factor buffer too small: grow, rerun
      continue;
    }
    if (rc) return rc;
    bool again = false;
    for (uint32_t j = 0; j < nj; ++j) {
      if (plan[j].status == FS_OK && (dev.flag[j] & FS_BJF_CANDOVF)) {
        plan[j].cap = (dev.cnt[j] + dev.cnt[j] / 8 + 4096 + o.tile_candidates - 1) / o.tile_candidates *
                      o.tile_candidates;
        again = true;
      }
    }
    if (!again) break;
    if (attempt >= 2) {
      fs_set_error_msg("batch: candidate buffer overflow");
      return FS_E_OVERFLOW;
    }
  }
  std::vector<std::vector<fs_factor>> per(nj);
  for (const fs_factor& f : dev.fac)
    if ((uint32_t)f.pad < nj) per[f.pad].push_back(f);
  std::vector<uint32_t> finished;
  for (uint32_t j = 0; j < nj; ++j)
    if (plan[j].status == FS_OK && !(dev.flag[j] & FS_BJF_VALIDATE)) finished.push_back(j);
  std::vector<char> verified(nj, 1);
  verify_jobs(per, finished, o.verify_threads, E.st, verified);
  uint64_t written = 0;
  for (uint32_t j = 0; j < nj; ++j) {
    fs_job_result& r = out[j];
    memset(&r, 0, sizeof r);
    r.status = plan[j].status;
    r.path = plan[j].path;
    if (per_job_nf) per_job_nf[j] = 0;
    if (r.status != FS_OK) continue;
    if (dev.flag[j] & FS_BJF_VALIDATE) {
      r.status = FS_E_VALIDATE;
      continue;
    }
    r.primecount = dev.cnt[j];
    r.steps = (uint64_t)dev.cnt[j] * (jobs[j].nmax - jobs[j].nmin);
    r.raw_factors = per[j].size();
    if (!verified[j]) {
      r.status = FS_E_VERIFY;
      continue;
    }
    uint64_t ck = dev.sum[j];
    const uint64_t kept = fs_filter_factors(per[j].data(), per[j].size(), &ck, false);
    per[j].resize(kept);
    r.checksum = ck;
    r.factorcount = kept;
    if (per_job_nf) per_job_nf[j] = kept;
    for (uint64_t k = 0; k < kept && factors && written < cap; ++k) factors[written++] = per[j][k];
  }
  if (seconds) *seconds = now_s() - t0;
  return FS_OK;
}

extern "C" int fs_batch_autotune(const fs_job* jobs, uint32_t nj, fs_batch_options* opt) {
  if (!opt || nj == 0) return FS_E_ARG;
  // sample: every job with its p-range cut to 1/8 (at least 2^20 integers)
  std::vector<fs_job> sample(jobs, jobs + nj);
  for (auto& j : sample) {
    const uint64_t len = j.pmax - j.pmin;
    const uint64_t cut = std::max<uint64_t>(len / 8, std::min<uint64_t>(len, 1u << 20));
    j.pmax = j.pmin + cut;
  }
  const uint32_t cus = fs_device_cus();
  uint32_t nlen_max = 0;
  for (uint32_t k = 0; k < nj; ++k) nlen_max = std::max(nlen_max, jobs[k].nmax - jobs[k].nmin);
  std::vector<fs_job_result> res(nj);
  // one warm run, then the fastest of three timed runs (single runs differ by
  // a few per cent, as much as some of the choices below)
  auto timed = [&](const fs_batch_options& o, double* sec) {
    int rc = fs_batch(sample.data(), nj, &o, res.data(), nullptr, 0, nullptr, sec);
    if (rc) return rc;
    double best_s = 1e30;
    for (int k = 0; k < 3; ++k) {
      if ((rc = fs_batch(sample.data(), nj, &o, res.data(), nullptr, 0, nullptr, sec))) return rc;
      best_s = std::min(best_s, *sec);
    }
    *sec = best_s;
    return FS_OK;
  };
  fs_batch_options best = *opt;
  if (best.tile_candidates == 0 || best.tile_candidates % FS_BLOCK) best.tile_candidates = 256;
  if (best.persistent_blocks == 0) best.persistent_blocks = cus * 8;  // the launch's default, made explicit
  double tbest = 1e30;
  int rc = timed(best, &tbest);
  if (rc) return rc;
  // coordinate descent: each dimension in turn, keeping the best value found;
  // a value replaces the current one only when it is faster by more than 2 %
  // (the sample's timing noise), so near-ties keep the default shape, which
  // the full-size batch prefers (a 1/8 sample has fewer tiles per workgroup:
  // with a 1 % margin, 4 workgroups per CU sometimes won on the sample and then
  // ran the full batch 2-4 % slower than 8)
  auto dim = [&](auto set, std::initializer_list<uint32_t> vals) -> int {
    for (uint32_t v : vals) {
      fs_batch_options o = best;
      set(o, v);
      double sec = 0;
      const int rc2 = timed(o, &sec);
      if (rc2) return rc2;
      if (sec < tbest * 0.98) {
        tbest = sec;
        best = o;
      }
    }
    return FS_OK;
  };
  if ((rc = dim([](fs_batch_options& o, uint32_t v) { o.tile_candidates = v; }, {256u, 512u})) ||
      (rc = dim([&](fs_batch_options& o, uint32_t v) { o.persistent_blocks = cus * v; }, {2u, 4u, 8u})) ||
      (rc = dim([](fs_batch_options& o, uint32_t v) { o.lpt_order = v; }, {0u, 1u})) ||
      (rc = dim([](fs_batch_options& o, uint32_t v) { o.unroll = v; }, {2u, 4u, 8u})) ||
      (rc = dim([](fs_batch_options& o, uint32_t v) { o.lanes_per_prime = v; }, {1u, 2u, 4u})))
    return rc;
  if (nlen_max > 4096) {
    const uint32_t c1 = std::max<uint32_t>(2048, nlen_max / 4), c2 = std::max<uint32_t>(1024, nlen_max / 16);
    if ((rc = dim([](fs_batch_options& o, uint32_t v) { o.n_chunk = v; }, {0u, c1, c2}))) return rc;
  }
  *opt = best;
  return FS_OK;
}
