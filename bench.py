"""Benchmark of the FactSieve GPU sieve on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c2|c1|c3]

A step is one whole plain sieve job of the workload through the C ABI
(fs_sieve: segmented candidate generation, nmin! setup, the n-iteration with
+-1 capture, the checksum fold, the result drain and CPU re-verification of
every factor).  The default workload is the production-shape single job
n in [1e5, 2e5), p in [1e12, 1e12 + 1e9) (36,191,004 candidates, 3.619e12
p*n residue steps per job).  `value` is p*n residue steps per second of whole
jobs.  One JSON line is printed.

roofline: the dominant kernel is the residue iteration.  Its work is counted
in VALU lane-operations per p*n step of the compiled 8-step loop, timed with
HIP events on the launch stream.  On the FP64 path (p < 2^53): 3.5 with paired
steps, used when nmax <= 2^22 and 64 nmax <= 31 p (one 6-operation product by
(n+1)(n+2) per two steps -- 2 multiplies, 2 FMAs, 1 add, 1 round -- and one
compare, |v'| <= the block's last n+2, which fires when the residue of n+1
(then v' = +-(n+2)) or of n+2 (then v' = +-1) is +-1), 2.5 when also
p nmax^2 <= 2^53 (exact product: 2 multiplies, 1 FMA, 1 round); otherwise 6.5
with single steps (one compare per two steps: |v[k+1]| <= m[k] covers +-1 at k
and k+1); 11.75 on the hybrid path of p > 2^53 (6 FP64 quotient ops,
2 v_mad_u64_u32, 2 v_mul_lo_u32, 1 add, and 6 filter ops per 8 steps), 7.0
with paired steps when nmax <= 2^22 and every modulus is in [2^45, 2^64 - 2^58)
(per two steps one product by (n+1)(n+2): 2 v_mad_u64_u32, 4 v_mul_lo_u32,
2 v_add3_u32 and 1 FMA for the quotient, 2 conversions, 1 multiply and 1 FMA
for the next quotient estimate, 1 filter minimum; 112 per 16 steps), 9.625
when every modulus is at least nmax * 2^34 (below 2^63; nmax * 2^39 above)
and the quotient comes from the high word of the residue alone (4 FP64 ops
instead of 6), and 4.5 on the FP64
path with single steps when p * nmax <= 2^53 and the product s*m is exact;
peak = 256 CUs x 64 lanes x 2.4 GHz = 39.3e12 lane-ops/s (the FP64
vector peak of 78.6 TFLOP/s counts each FMA lane-op as two flops).
int32_mul_equiv: the same steps/s expressed in the 11 32-bit multiplies a
64-bit Montgomery step needs, over the measured int32-multiply peak (full
VALU rate on gfx950).
cpu_baseline: the CPU oracle (a C/OpenMP port of the reference algorithm) on a
bounded p-subrange of the same workload, on this host's cores.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WORKLOADS = {
    # name: (pmin, pmax, nmin, nmax, description)
    "c2": (10 ** 12, 10 ** 12 + 10 ** 9, 100000, 200000,
           "plain single job at production shape: n in [1e5,2e5), p in [1e12,1e12+1e9)"),
    "c1": (10 ** 8, 10 ** 8 + 2 * 10 ** 6, 1000, 10000,
           "plain sieve at self-test scale: n in [1e3,1e4), p in [1e8,1e8+2e6)"),
    "c3": (2 ** 62, 2 ** 62 + 2 ** 30, 500000, 510000,
           "full-width modulus stress: n in [5e5,5.1e5), p in [2^62,2^62+2^30)"),
}
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12       # lane-ops/s, 2.4 GHz peak engine clock
# VALU lane-ops per p*n step by path, from the compiled 8-step loops: FP64 52
# per 8 steps, hybrid 94, Montgomery ~45 per step
OPS_PER_STEP = {1: 6.5, 2: 11.75, 3: 45}
OPS_PER_STEP_HYBRID_HI = 9.625                  # hybrid with the high-word quotient: 77 per 8 steps
OPS_PER_STEP_HYBRID_PAIR = 7.0                  # hybrid paired steps (nmax <= 2^22, 2^45 <= p < 2^64 - 2^58):
                                                # 8 x 13 + 7 filter minimums + 1 compare per 16 steps
OPS_PER_STEP_FP64_EXACT = 4.5                   # FP64 with an exact product (p * nmax <= 2^53): 36 per 8
OPS_PER_STEP_FP64_PAIR = 3.5                    # FP64 paired steps (64 nmax <= 31 p): 8 x (6 + 1 compare) per 16
OPS_PER_STEP_FP64_PAIR_EXACT = 2.5              # paired steps with an exact product (p nmax^2 <= 2^53): 8 x (4 + 1)
# VALU lane-ops per setup program term (PMC of k_setup_fp64 / k_setup_hybrid):
# counted in the iterate window of a job run as lanes per prime, whose one
# launch (k_lanes) sets up and iterates.  FP64 setups square in FP64 while every
# modulus is below 2^50 (6.04 per term) and with Montgomery products above (6.3).
SETUP_OPS_PER_TERM = {1: 6.3, 2: 18.1}
SETUP_OPS_PER_TERM_FP64_FPSQ = 6.04


def setup_ops_per_term(path, pmax):
    return SETUP_OPS_PER_TERM_FP64_FPSQ if path == 1 and pmax - 1 < 2 ** 50 else SETUP_OPS_PER_TERM[path]


def fp64_ops_at(pmin, pmax, n1):
    """VALU operations per step of the FP64 iterate loop every wave of a launch
    ending at n1 takes (the per-wave choice of fs_cand_iter_fp64)."""
    if n1 <= 2 ** 22 and 64 * n1 <= 31 * pmin:
        return OPS_PER_STEP_FP64_PAIR_EXACT if pmax * n1 * n1 <= 2 ** 53 else OPS_PER_STEP_FP64_PAIR
    return OPS_PER_STEP_FP64_EXACT if pmax * n1 <= 2 ** 53 else OPS_PER_STEP[1]


def fp64_ops(pmin, pmax, nmin, nmax):
    """Step-weighted VALU operations per step of a job's FP64 iterate launches:
    the driver ends launches where a cheaper loop stops applying (paired steps
    up to 31 pmin / 64 and 2^22, exact products up to 2^53 / pmax and
    isqrt(2^53 / pmax)), as for a one-segment job."""
    top = pmax - 1
    pair = min(31 * pmin // 64, 2 ** 22)
    cuts = sorted({c for c in (pair, min(pair, math.isqrt(2 ** 53 // top)), 2 ** 53 // top) if nmin < c < nmax})
    total, a = 0.0, nmin
    for b in cuts + [nmax]:
        total += (b - a) * fp64_ops_at(pmin, pmax, b)
        a = b
    return total / max(nmax - nmin, 1)


def hybrid_pair(pmin, pmax, nmax):
    """The hybrid iterate takes paired steps when nmax <= 2^22 and every
    modulus is in [2^45, 2^64 - 2^58)."""
    return nmax <= 2 ** 22 and pmin >= 2 ** 45 and pmax - 1 < 2 ** 64 - 2 ** 58


def hybrid_hi(pmin, pmax, nmax):
    """The hybrid iterate takes its quotient from the high word of the residue
    when every modulus p satisfies p >= nmax * 2^34 and p < 2^63, or
    p >= nmax * 2^39 and p < 2^64 - 2^58."""
    def mmax(p):
        if p >= 2 ** 64 - 2 ** 58:
            return 0
        return p >> 34 if p < 2 ** 63 else p >> 39
    top = pmax - 1
    lowest = mmax(pmin) if top < 2 ** 63 else min(mmax(pmin), mmax(max(pmin, 2 ** 63)), mmax(top))
    return lowest >= nmax
INT32_MUL_PER_STEP = 11                         # 64-bit Montgomery step: 4 + 3 + 4 partial products
INT32_MUL_PEAK = VALU_PEAK_TOPS * 1e12          # full-rate v_mad_u64_u32 / v_mul_*_u32 (measured)
# HBM bytes per candidate per iterate launch, measured with rocprofv3 --pmc FETCH_SIZE
# and WRITE_SIZE (k_iterate_fp64 on c2, launches of 18 M candidates: 23.6 B read + 12.0 B
# written; k_iterate_hybrid on c3: 20.1 + 12.0), each scaled by the counter's reading for
# a 1 GiB device copy (x2.0 fetch, x1.0 write).
TRAFFIC_BYTES_PER_CAND = {1: 35.6, 2: 32.1}


def cpu_baseline(pmin, nmin, nmax, target_s=15.0):
    """Time the oracle on a p-subrange sized for ~target_s seconds of CPU work."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol  # the oracle is used here only as the CPU baseline

    # the host cores this process may run on; a GPU machine's share of them is
    # what OMP_NUM_THREADS names there (16 per GPU), and the oracle runs on all of it
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    threads = min(avail, share) if share > 0 else avail
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    span = 2 * 10 ** 5
    t = time.time()
    ol.sieve(pmin, pmin + span, nmin, nmax, threads=threads)
    dt = time.time() - t
    span = int(min(10 ** 9, max(span, span * target_s / max(dt, 1e-3))))
    t = time.time()
    r = ol.sieve(pmin, pmin + span, nmin, nmax, threads=threads)
    dt = time.time() - t
    steps = r["primecount"] * (nmax - nmin)
    return {"value": steps / dt, "unit": "p*n steps/s", "cores": threads, "kind": "port",
            "cpu_model": model, "nproc": os.cpu_count(), "cores_available": avail,
            "omp_num_threads": share or None,
            "sample": f"oracle (C/OpenMP port of the reference algorithm) on p in [{pmin}, {pmin + span}), "
                      f"n in [{nmin}, {nmax}): {r['primecount']} candidates, {steps:.3e} p*n steps in {dt:.1f} s"}


def bench_batched(fs, njobs=64, seed=None):
    """Mixed-job config: the same 64 jobs as one batched persistent run and as
    64 plain single-job runs; results must agree job by job."""
    from factsieve_amd import workloads
    jobs = workloads.mixed_jobs(njobs, seed=workloads.DEFAULT_SEED if seed is None else seed)
    print("bench: autotuning the batch", file=sys.stderr, flush=True)
    opts = fs.batch_autotune(jobs)
    print(f"bench: autotuned {opts}", file=sys.stderr, flush=True)
    # Both forms are timed after one untimed pass over the same jobs (setup
    # programs cached by nmin, device buffers sized, code objects loaded).
    fs.batch(jobs, **opts)
    t = time.perf_counter()
    res, _ = fs.batch(jobs, **opts)
    t_batch = time.perf_counter() - t
    print(f"bench: batched {t_batch:.3f} s; plain jobs next", file=sys.stderr, flush=True)
    for j in jobs:
        fs.sieve(*j, verify_threads=16)
    t = time.perf_counter()
    plain = [fs.sieve(*j, verify_threads=16) for j in jobs]
    t_plain = time.perf_counter() - t
    same = all((r.factorcount, r.primecount, r.checksum, r.factors) ==
               (p.factorcount, p.primecount, p.checksum, p.factors) and r.status == 0 for r, p in zip(res, plain))
    steps = sum(r.steps for r in res)
    ops = batch_ops(fs, jobs, res)
    return {"workload": f"{njobs} mixed jobs (seed {workloads.DEFAULT_SEED if seed is None else seed}): n_start "
                        "U[1e3,5e5], n_len lognormal(median 2e4, sigma 1), p_start 2^U[32,63], p_len "
                        "lognormal(median 2^24, sigma 1)",
            "pn_steps": steps, "candidates": sum(r.primecount for r in res),
            "batched_steps_per_s": steps / t_batch, "plain_steps_per_s": steps / t_plain,
            "batched_s": t_batch, "plain_s": t_plain, "speedup": t_plain / t_batch,
            "bit_exact_vs_plain": same, "autotuned": opts,
            "roofline_wall": {"bound": "valu", "kernel": "k_bsieve (whole batch: wall time, candidate generation "
                                                        "and host work included)",
                              "achieved": ops / t_batch / 1e12, "peak": VALU_PEAK_TOPS, "unit": "TOP/s (VALU lane-ops)",
                              "frac": ops / t_batch / 1e12 / VALU_PEAK_TOPS, "lane_ops": ops},
            "paths": {k: sum(1 for r in res if r.path == v) for k, v in (("fp64", 1), ("hybrid", 2), ("mont", 3))}}, \
        {tuple(j): r for j, r in zip(jobs, res)}, {tuple(j): p.seconds for j, p in zip(jobs, plain)}


def batch_ops(fs, jobs, res):
    """VALU lane-operations of a batch by the same model as the plain legs: each
    job's steps at the loop its waves take for the whole n-range (the batch runs
    it unchunked) plus its candidates' setup program terms."""
    total = 0.0
    for (pmin, pmax, nmin, nmax), r in zip(jobs, res):
        if r.path == 1:
            per_step = fp64_ops_at(pmin, pmax, nmax)
        elif r.path == 2:
            per_step = (OPS_PER_STEP_HYBRID_PAIR if hybrid_pair(pmin, pmax, nmax) else
                        OPS_PER_STEP_HYBRID_HI if hybrid_hi(pmin, pmax, nmax) else OPS_PER_STEP[2])
        else:
            per_step = OPS_PER_STEP[3]
        total += r.steps * per_step
        if r.path in SETUP_OPS_PER_TERM:
            total += r.primecount * fs.setup_program_terms(nmin, r.path) * setup_ops_per_term(r.path, pmax)
    return total


def bench_graph_replay(fs, reps=300):
    """Per-call wall time of a tiny batch (one job of ~260 candidates and 100
    n) with its launches replayed from the cached graph against enqueued one
    by one: the launch cost a quantum saves with the graph."""
    job = [(10 ** 8, 10 ** 8 + 5000, 1000, 1100)]
    out = {}
    for graph in (1, 0):
        for _ in range(20):
            fs.batch(job, use_graph=graph)
        t = time.perf_counter()
        for _ in range(reps):
            fs.batch(job, use_graph=graph)
        out["graph_ms" if graph else "direct_ms"] = (time.perf_counter() - t) / reps * 1e3
    out["saving_us_per_launch"] = (out["direct_ms"] - out["graph_ms"]) * 1e3
    out["workload"] = "one job p in [1e8, 1e8+5000), n in [1000, 1100), %d calls each" % reps
    return out


def bench_serving(fs, reference, alone_s=None, rate_frac=0.7, nwu=64, n_chunk=49152, use_graph=1):
    """Poisson arrivals of mixed workunits into the continuous-batching server;
    p50/p99 arrival-to-result latency and aggregate steps/s.  With alone_s (each
    workunit's wall time run alone, one job at a time on the whole GPU) also the
    p50/p99 of each workunit's latency over that time (its slowdown)."""
    from factsieve_amd import workloads
    jobs = workloads.mixed_jobs(nwu)
    mean_steps = sum(r.steps for r in reference.values()) / max(len(reference), 1)
    cap = bench_serving.capacity_steps_per_s
    rate = rate_frac * cap / mean_steps
    arrivals = workloads.poisson_arrivals(rate, nwu)
    srv = fs.Server(n_chunk=n_chunk, use_graph=use_graph)
    ids, lat, results = {}, {}, {}
    steps_run, step_s = 0, 0.0
    t0 = time.perf_counter()
    nxt = 0
    try:
        while len(results) < nwu:
            now = time.perf_counter() - t0
            while nxt < nwu and arrivals[nxt] <= now:
                ids[srv.submit(jobs[nxt])] = nxt
                nxt += 1
            st = srv.stats()
            if st["active"] == 0 and st["queued"] == 0 and nxt < nwu:
                time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
                continue
            ts = time.perf_counter()
            srv.step()
            step_s += time.perf_counter() - ts
            steps_run += 1
            now = time.perf_counter() - t0
            for wid, k in ids.items():
                if k in results:
                    continue
                r = srv.poll(wid)
                if r is not None:
                    results[k] = r
                    lat[k] = now - arrivals[k]
        makespan = time.perf_counter() - t0
    finally:
        srv.close()
    steps = sum(r.steps for r in results.values())
    same = all((results[k].factorcount, results[k].primecount, results[k].checksum, results[k].factors) ==
               (reference[tuple(jobs[k])].factorcount, reference[tuple(jobs[k])].primecount,
                reference[tuple(jobs[k])].checksum, reference[tuple(jobs[k])].factors) for k in results)
    ls = sorted(lat.values())
    out = {"workunits": nwu, "arrival_rate_per_s": rate, "target_utilisation": rate_frac, "n_chunk": n_chunk,
            "graph": bool(use_graph), "steps": steps_run, "mean_step_ms": step_s / max(steps_run, 1) * 1e3,
            "p50_latency_s": ls[len(ls) // 2], "p99_latency_s": ls[min(len(ls) - 1, int(0.99 * len(ls)))],
            "makespan_s": makespan, "steps_per_s": steps / makespan, "bit_exact_vs_batched_alone": same}
    if alone_s:
        sl = sorted(lat[k] / alone_s[tuple(jobs[k])] for k in lat)
        out["p50_slowdown"] = sl[len(sl) // 2]
        out["p99_slowdown"] = sl[min(len(sl) - 1, int(0.99 * len(sl)))]
        out["largest_alone_s"] = max(alone_s[tuple(j)] for j in jobs)
    return out


def plain_leg(fs, workload, steps, warmup):
    """`steps` timed whole jobs of a plain workload after `warmup` untimed ones:
    (whole-job p*n steps/s, the wall time, the job results, roofline and kernel
    times of its iterate launches)."""
    pmin, pmax, nmin, nmax, desc = WORKLOADS[workload]
    print(f"bench: plain {workload}, {steps} timed jobs", file=sys.stderr, flush=True)
    for _ in range(warmup):
        fs.sieve(pmin, pmax, nmin, nmax, verify_threads=16)
    results = []
    t0 = time.perf_counter()
    for _ in range(steps):
        results.append(fs.sieve(pmin, pmax, nmin, nmax, verify_threads=16))
    elapsed = time.perf_counter() - t0
    ref = results[0]
    for r in results[1:]:
        assert (r.checksum, r.factorcount, r.primecount) == (ref.checksum, ref.factorcount, ref.primecount)
    steps_total = sum(r.steps for r in results)
    it_ms = sum(r.kernel_ms_iterate for r in results)
    launches = sum(r.iterate_launches for r in results)
    ops = OPS_PER_STEP[ref.path]
    if ref.path == 2 and hybrid_pair(pmin, pmax, nmax):
        ops = OPS_PER_STEP_HYBRID_PAIR
    elif ref.path == 2 and hybrid_hi(pmin, pmax, nmax):
        ops = OPS_PER_STEP_HYBRID_HI
    if ref.path == 1:
        ops = fp64_ops(pmin, pmax, nmin, nmax)
    window_ops = steps_total * ops
    if ref.lanes > 1:  # the lanes' launch also runs every lane's setup
        window_ops += sum(r.setup_mulmods for r in results) * setup_ops_per_term(ref.path, pmax)
        ops = window_ops / steps_total
    achieved = window_ops / (it_ms * 1e-3) / 1e12
    traffic = None
    if ref.path in TRAFFIC_BYTES_PER_CAND:
        traffic = TRAFFIC_BYTES_PER_CAND[ref.path] * ref.primecount * max(ref.lanes, 1) / max(ref.iterate_launches, 1)
    leg = {
        "workload": desc, "pmin": pmin, "pmax": pmax, "nmin": nmin, "nmax": nmax,
        "candidates": ref.primecount, "pn_steps_per_job": ref.steps, "factors": ref.factorcount,
        "checksum": f"{ref.checksum:016X}", "path": {1: "fp64", 2: "hybrid", 3: "montgomery"}[ref.path],
        "segments": ref.segments, "lanes_per_prime": ref.lanes,
        "steps_per_s": steps_total / elapsed, "ms_per_job": elapsed / steps * 1e3,
        "roofline": {"bound": "valu",
                     "kernel": ("k_lanes<%s>" % {1: "FP64", 2: "HYBRID"}[ref.path]) if ref.lanes > 1
                     else "k_iterate_" + {1: "fp64", 2: "hybrid", 3: "mont"}[ref.path],
                     "achieved": achieved, "peak": VALU_PEAK_TOPS, "unit": "TOP/s (VALU lane-ops)",
                     "frac": achieved / VALU_PEAK_TOPS, "traffic": traffic, "traffic_unit": "bytes per launch",
                     "ops_per_step": ops, "avg_launch_ms": it_ms / max(launches, 1),
                     "kernel_steps_per_s": steps_total / (it_ms * 1e-3)},
        "int32_mul_equiv": {"frac": steps_total / (it_ms * 1e-3) * INT32_MUL_PER_STEP / INT32_MUL_PEAK,
                            "muls_per_step": INT32_MUL_PER_STEP, "peak_ops_per_s": INT32_MUL_PEAK},
        "kernel_ms_per_job": {"iterate": it_ms / steps,
                              "setup": sum(r.kernel_ms_setup for r in results) / steps,
                              "primegen": sum(r.kernel_ms_primegen for r in results) / steps},
    }
    return leg, elapsed, ref


def valu_peaks(fs):
    """Measured issue rate of each VALU instruction kind the kernels use."""
    out = {}
    for kind, name in fs.VALU_PEAK_KINDS.items():
        rate, ms = fs.valu_peak(kind)
        out[name] = {"lane_ops_per_s": rate, "frac_of_2p4ghz_issue": rate / (VALU_PEAK_TOPS * 1e12), "ms": ms}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS) + ["c4"],
                    help="plain workload of the headline line (c4: only the batched leg, for profiling)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched and serving legs")
    ap.add_argument("--no-extra", action="store_true", help="skip the c1 and c3 plain legs")
    ap.add_argument("--valu-peak", action="store_true",
                    help="only measure the VALU issue rate of each instruction kind and exit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy of BYTES (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        import factsieve_amd as fs
        fs.init(0)
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        fs.check(fs.lib().fs_malloc(ctypes.byref(a), args.copy_calibration))
        fs.check(fs.lib().fs_malloc(ctypes.byref(b), args.copy_calibration))
        fs.check(fs.lib().fs_memset(a, 1, args.copy_calibration, None))
        fs.check(fs.lib().fs_memcpy_dtod(b, a, args.copy_calibration, None))
        fs.check(fs.lib().fs_stream_sync(None))
        print(json.dumps({"copy_bytes": args.copy_calibration}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the sieve shards over p across GPUs with a host-side checksum sum, not run here)",
              file=sys.stderr)
        sys.exit(2)

    import factsieve_amd as fs
    fs.init(0)
    if args.valu_peak:
        print(json.dumps({"valu_peak": valu_peaks(fs), "peak_2p4ghz_lane_ops_per_s": VALU_PEAK_TOPS * 1e12}),
              flush=True)
        return
    if args.workload == "c4":
        bat, _, _ = bench_batched(fs)
        print(json.dumps({"metric": "p*n residue steps/s (batched, 64 mixed jobs)",
                          "value": bat["batched_steps_per_s"], "unit": "p*n steps/s", "batched": bat}), flush=True)
        return
    leg, elapsed, ref = plain_leg(fs, args.workload, args.steps, args.warmup)
    value = leg["steps_per_s"]
    line = {
        "metric": "p*n residue steps/s (plain single job)",
        "value": value,
        "unit": "p*n steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {1: "f64 (exact integer residues)", 2: "u64+f64", 3: "u64"}[ref.path],
        "data": "synthetic (the workload is a numeric range; no data set)",
        "config": {k: leg[k] for k in ("workload", "pmin", "pmax", "nmin", "nmax", "candidates",
                                       "pn_steps_per_job", "factors", "checksum", "path", "segments",
                                       "lanes_per_prime")},
        "roofline": leg["roofline"],
        "int32_mul_equiv": leg["int32_mul_equiv"],
        "kernel_ms_per_job": leg["kernel_ms_per_job"],
    }
    line["config"]["parallelism"] = "single GPU"
    if not args.no_extra:
        for w, steps, warm in (("c1", 30, 3), ("c3", 3, 1)):
            if w != args.workload:
                line["plain_" + w] = plain_leg(fs, w, steps, warm)[0]
    if not args.no_batch:
        bat, ref_jobs, alone = bench_batched(fs)
        line["batched"] = bat
        bench_serving.capacity_steps_per_s = bat["batched_steps_per_s"]
        line["serving"] = bench_serving(fs, ref_jobs, alone)
        line["serving_without_graph"] = bench_serving(fs, ref_jobs, alone, use_graph=0)
        line["graph_replay"] = bench_graph_replay(fs)
    if not args.no_cpu_baseline:
        pmin, _, nmin, nmax, _ = WORKLOADS[args.workload]
        line["cpu_baseline"] = cpu_baseline(pmin, nmin, nmax)
        line["cpu_baseline"]["gpu_over_cpu"] = value / line["cpu_baseline"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
