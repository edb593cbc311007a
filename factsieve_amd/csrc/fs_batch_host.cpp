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
                          uint32_t fcap) {
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
    if ((rc = grow(&d.ntab, n, "batch ntab")) || (rc = fs_batch_fill_ntab(d.ntab, n, st))) return rc;
    d.ntab_len = n;
  }
  if (fcap > d.fcap || !d.fac) {
    if ((rc = grow(&d.fac, fcap, "batch factors"))) return rc;
    d.fcap = fcap;
  }
  return FS_OK;
}

void fs_batch_bufs::release() {
  if (d.c.p) fs_cands_free(&d.c);
  void* ps[] = {d.jobs, d.order, d.chunk, d.jcnt, d.jflag, d.jsum, d.glob, d.pref, d.tiles, d.prog, d.ntab, d.fac};
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
                        uint32_t ncands, uint32_t tile, uint32_t blocks, uint32_t fcap, bool graph, float* dev_ms) {
  const uint32_t nj = (uint32_t)bj.size();
  uint32_t nmax_max = 0, ntiles = 0;
  for (uint32_t k = 0; k < order.size(); ++k) {
    const fs_bjob& b = bj[order[k]];
    nmax_max = std::max(nmax_max, b.nmax);
    ntiles += (b.cand_cap + tile - 1) / tile;
  }
  int rc = E.ensure(nj, ncands, (uint32_t)words.size(), ntiles, nmax_max, fcap);
  if (rc) return rc;
  fs_batch_dev& d = E.d;
  HCHK(hipMemcpyAsync(d.jobs, bj.data(), nj * sizeof(fs_bjob), hipMemcpyHostToDevice, E.st), "batch upload");
  if (!order.empty()) {
    HCHK(hipMemcpyAsync(d.order, order.data(), order.size() * 4, hipMemcpyHostToDevice, E.st), "batch upload");
    HCHK(hipMemcpyAsync(d.chunk, chunk.data(), chunk.size() * 4, hipMemcpyHostToDevice, E.st), "batch upload");
  }
  if (!words.empty())
    HCHK(hipMemcpyAsync(d.prog, words.data(), words.size() * 8, hipMemcpyHostToDevice, E.st), "batch upload");
  HCHK(hipStreamSynchronize(E.st), "batch upload sync");
  HCHK(hipEventRecord(E.ev0, E.st), "batch event");
  if (graph) {
    hipGraph_t g;
    hipGraphExec_t ge;
    HCHK(hipStreamBeginCapture(E.st, hipStreamCaptureModeThreadLocal), "batch capture");
    rc = fs_batch_launch(d, nj, (uint32_t)order.size(), pg_blocks, tile, blocks, E.st);
    hipError_t ce = hipStreamEndCapture(E.st, &g);
    if (rc) return rc;
    HCHK(ce, "batch end capture");
    hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HCHK(ie, "batch graph instantiate");
    hipError_t le = hipGraphLaunch(ge, E.st);
    hipError_t se = hipStreamSynchronize(E.st);
    (void)hipGraphExecDestroy(ge);
    HCHK(le, "batch graph launch");
    HCHK(se, "batch graph sync");
  } else {
    if ((rc = fs_batch_launch(d, nj, (uint32_t)order.size(), pg_blocks, tile, blocks, E.st))) return rc;
  }
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
  static uint32_t cus = 0;
  if (!cus) {
    if (fs_device_info(nullptr, 0, &cus, nullptr, nullptr) != FS_OK || !cus) cus = 256;
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
}

namespace {

struct Plan {
  int status = FS_OK;
  int path = 0;
  uint32_t terms = 0;
  std::shared_ptr<const fs_program> prog;
  uint32_t cap = 0;
  uint32_t pg_blocks = 0;
  double cost = 0;
};


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
  std::vector<uint32_t> order, chunk;
  uint32_t cand_off = 0, pg = 0;
  for (uint32_t j = 0; j < nj; ++j) {
    fs_bjob& b = bj[j];
    memset(&b, 0, sizeof b);
    b.pg_first = pg;
    if (plan[j].status != FS_OK) continue;
    b.pmin = jobs[j].pmin;
    b.pmax = jobs[j].pmax;
    b.nmin = jobs[j].nmin;
    b.nmax = jobs[j].nmax;
    b.path = (uint32_t)plan[j].path;
    b.prog_off = (uint32_t)words.size();
    b.prog_terms = plan[j].terms;
    words.insert(words.end(), plan[j].prog->words.begin(), plan[j].prog->words.end());
    while (words.size() % 8) words.push_back(0);  // 64-byte aligned programs
    b.cand_off = cand_off;
    b.cand_cap = plan[j].cap;
    b.flags = FS_BJ_NEW;
    cand_off += plan[j].cap;
    pg += plan[j].pg_blocks;
    order.push_back(j);
  }
  if (lpt)
    std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return plan[a].cost > plan[b].cost; });
  for (uint32_t j : order) {
    chunk.push_back(jobs[j].nmin);
    chunk.push_back(jobs[j].nmax);
  }
  const uint32_t blocks = o.persistent_blocks ? o.persistent_blocks : fs_device_cus() * 8;
  int rc = fs_batch_run_device(E, bj, order, chunk, words, pg, cand_off, tile, blocks, fcap, o.use_graph != 0,
                               dev_ms);
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
  plan.assign(nj, Plan());
  for (uint32_t j = 0; j < nj; ++j) {
    Plan& P = plan[j];
    const fs_job& J = jobs[j];
    if (fs_validate_ranges(J.pmin, J.pmax, J.nmin, J.nmax) != FS_OK) {
      P.status = FS_E_ARG;
      continue;
    }
    P.path = fs_select_path(J.pmax);
    P.prog = fs_program_get(J.nmin, P.path);
    if (!P.prog) {
      P.status = FS_E_NOMEM;
      continue;
    }
    P.terms = P.prog->terms;
    const uint64_t c = fs_cand_capacity(J.pmin, J.pmax - J.pmin);
    if (c > 0x7fffffffull) {
      fs_set_error_msg("batch: job p-range too large for one batch");
      P.status = FS_E_ARG;
      continue;
    }
    P.cap = (uint32_t)((c + tile - 1) / tile * tile);
    P.pg_blocks = fs_pg_blocks(J.pmin, J.pmax);
    P.cost = (double)(J.nmax - J.nmin) * step_cost(P.path) + (double)P.terms * setup_cost(P.path);
  }
  return FS_OK;
}

}  // namespace

// Verify, filter and fold the factors of one job; returns kept factors.
// Sorts each listed job's factors and verifies all of them in one pass over
// the CPU threads (factors of different jobs share the pool).  ok[j] = 0 for a
// job with a factor that fails verification.
static void verify_jobs(std::vector<std::vector<fs_factor>>& per, const std::vector<uint32_t>& list, int threads,
                        std::vector<char>& ok) {
  std::vector<fs_factor> all;
  for (uint32_t j : list) {
    for (auto& x : per[j]) x.pad = 0;
    fs_sort_factors(per[j].data(), per[j].size());
    all.insert(all.end(), per[j].begin(), per[j].end());
  }
  if (fs_verify_all(all.data(), all.size(), threads) < 0) return;
  for (uint32_t j : list) {  // a failure: find the job(s) it belongs to
    const int64_t bad = fs_verify_all(per[j].data(), per[j].size(), threads);
    if (bad >= 0) {
      char buf[160];
      snprintf(buf, sizeof buf, "CPU factor verification failed for p=%llu", (unsigned long long)per[j][bad].p);
      fs_set_error_msg(buf);
      ok[j] = 0;
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
  verify_jobs(per, finished, o.verify_threads, verified);
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
  double best = 1e30;
  fs_batch_options bo = *opt;
  std::vector<fs_job_result> res(nj);
  for (uint32_t tile : {256u, 512u}) {
    for (uint32_t mult : {2u, 4u, 8u}) {
     for (uint32_t lpt : {0u, 1u}) {
      fs_batch_options o = *opt;
      o.tile_candidates = tile;
      o.persistent_blocks = cus * mult;
      o.lpt_order = lpt;
      double sec = 0;
      int rc = fs_batch(sample.data(), nj, &o, res.data(), nullptr, 0, nullptr, &sec);  // warm
      if (rc) return rc;
      rc = fs_batch(sample.data(), nj, &o, res.data(), nullptr, 0, nullptr, &sec);
      if (rc) return rc;
      if (sec < best) {
        best = sec;
        bo = o;
      }
     }
    }
  }
  *opt = bo;
  return FS_OK;
}
