"""GPU parity of the batched persistent form and the serving loop: every job's
factors, candidate count and checksum must equal the plain single-job sieve
(and the oracle on the small jobs), whatever the mix, tile shape, queue order,
graph capture or admission schedule."""
import pytest

import oracle_lib as ol
from jobs_lib import random_jobs, wide_jobs
from factsieve_amd import workloads

pytestmark = pytest.mark.gpu


def small_mix():
    jobs = []
    for pmin, pmax, nmin, nmax in workloads.mixed_jobs(20, seed=7):
        jobs.append((pmin, pmin + max(1, (pmax - pmin) // 256), nmin, nmin + max(1, (nmax - nmin) // 16)))
    jobs += [
        (10 ** 8, 10 ** 8 + 300000, 1000, 3000),        # FP64, factors present
        (127, 20000, 127, 30000),                       # p <= nmax: many factors
        (10 ** 9, 10 ** 9 + 100000, 5000, 5000),        # empty n-range
        (2 ** 63 + 1000, 2 ** 63 + 200000, 2000, 2600), # hybrid above 2^63
        (2 ** 64 - 100000, 2 ** 64 - 1, 500, 900),      # top of the range: Montgomery
        (2 ** 52 - 50000, 2 ** 52 + 50000, 3000, 3300), # FP64 across 2^52
        (2 ** 53 - 50000, 2 ** 53 + 50000, 3000, 3300), # hybrid across 2^53
    ]
    return jobs


def plain(gpu, job):
    r = gpu.sieve(*job, verify_threads=16)
    return (r.factorcount, r.primecount, r.checksum, r.factors)


def outcome(o):
    return (o.factorcount, o.primecount, o.checksum, o.factors)


def test_batch_matches_plain_on_mixed_jobs(gpu):
    jobs = small_mix()
    res, sec = gpu.batch(jobs)
    assert len(res) == len(jobs)
    for job, o in zip(jobs, res):
        print("job", job, flush=True)
        assert o.status == 0
        assert outcome(o) == plain(gpu, job)


def test_batch_small_jobs_match_oracle(gpu):
    jobs = small_mix()[-6:-1]
    res, _ = gpu.batch(jobs)
    for job, o in zip(jobs, res):
        want = ol.sieve(*job, threads=8)
        assert (o.factorcount, o.primecount, o.checksum) == (want["factorcount"], want["primecount"],
                                                             want["checksum"])
        assert o.factors == want["factors"]


@pytest.mark.parametrize("opts", [dict(tile_candidates=512), dict(persistent_blocks=64), dict(use_graph=0),
                                  dict(lpt_order=0), dict(unroll=2), dict(unroll=8), dict(n_chunk=700),
                                  dict(n_chunk=1500, use_graph=0), dict(lanes_per_prime=2),
                                  dict(lanes_per_prime=8, unroll=8, n_chunk=3000)])
def test_batch_shape_invariance(gpu, opts):
    jobs = small_mix()
    base, _ = gpu.batch(jobs)
    print("options", opts, flush=True)
    other, _ = gpu.batch(jobs, **opts)
    assert [outcome(o) for o in base] == [outcome(o) for o in other]


def pair_mix():
    # FP64 jobs with 128 nmax^2 <= 63 p (paired steps): with and without the
    # exact product, odd and even first multipliers, one straddling
    # 128 nmax^2 = 63 p
    return [
        (4 * 10 ** 6, 4 * 10 ** 6 + 400000, 10, 1000),           # exact pairs, factors present
        (4 * 10 ** 6 + 400000, 4 * 10 ** 6 + 800000, 11, 997),   # odd nmin and nmax
        (10 ** 10, 10 ** 10 + 200000, 20001, 49999),             # inexact pairs
        (1600000000 - 100000, 1600000000 + 100000, 19500, 20000),  # all paired
        (812698413 - 100000, 812698413 + 100000, 19500, 20000),  # paired and single waves
    ]


@pytest.mark.parametrize("opts", [dict(), dict(n_chunk=777), dict(n_chunk=1001, unroll=4), dict(unroll=2),
                                  dict(lanes_per_prime=2), dict(lanes_per_prime=4, n_chunk=1333)])
def test_batch_paired_steps_match_oracle(gpu, opts):
    """Paired FP64 steps in the batched kernel read an absolute pair table split
    by the parity of the first multiplier; n-chunks of odd length and lanes
    per prime start tiles at both parities.  Every job against the oracle."""
    jobs = pair_mix()
    print("options", opts, flush=True)
    res, _ = gpu.batch(jobs, **opts)
    for job, o in zip(jobs, res):
        want = ol.sieve(*job, threads=8)
        assert o.status == 0
        assert (o.factorcount, o.primecount, o.checksum) == (want["factorcount"], want["primecount"],
                                                             want["checksum"]), job
        assert o.factors == want["factors"]


def test_batch_hybrid_high_word_limits_match_oracle(gpu):
    """The batched kernel's per-wave choice of the hybrid high-word quotient at
    its limits: p >> 34 crossing n1 just above 2^53, the limit dropping from
    p >> 34 to p >> 39 at 2^63, and a job far inside each regime; every job
    against the oracle."""
    jobs = [((524300 << 34) - 100000, (524300 << 34) + 100000, 523800, 524300),
            (2 ** 63 - 3000, 2 ** 63 + 3000, 2 ** 24 - 40, 2 ** 24 + 40),
            (2 ** 55 + 1, 2 ** 55 + 200000, 300000, 301000),
            (2 ** 63 + 1, 2 ** 63 + 100000, 20000, 21000)]
    res, _ = gpu.batch(jobs)
    for job, o in zip(jobs, res):
        print("job", job, flush=True)
        want = ol.sieve(*job, threads=8)
        assert o.status == 0
        assert (o.factorcount, o.primecount, o.checksum) == (want["factorcount"], want["primecount"],
                                                             want["checksum"]), job
        assert o.factors == want["factors"]


@pytest.mark.parametrize("seed", [4, 5])
def test_batch_random_jobs_match_oracle(gpu, seed):
    """Seeded random jobs over every modulus size (every path and loop) in one
    batch, each against the oracle."""
    jobs = random_jobs(seed, 16)
    res, _ = gpu.batch(jobs)
    for job, o in zip(jobs, res):
        want = ol.sieve(*job, threads=8)
        assert o.status == 0
        assert (o.factorcount, o.primecount, o.checksum) == (want["factorcount"], want["primecount"],
                                                             want["checksum"]), job
        assert o.factors == want["factors"]


@pytest.mark.parametrize("seed", [11, 12])
def test_wide_random_jobs_plain_and_batch_match_oracle(gpu, seed):
    """Seeded random jobs with setup programs of every length (nmin up to 2e5)
    and moduli from 2^20 to 2^64, through the plain sieve and one batch, each
    against the oracle."""
    jobs = wide_jobs(seed, 12)
    res, _ = gpu.batch(jobs)
    for job, b in zip(jobs, res):
        print("wide", job, flush=True)
        o = ol.sieve(*job, threads=8)
        want = (o["factorcount"], o["primecount"], o["checksum"], o["factors"])
        r = gpu.sieve(*job)
        assert b.status == 0, job
        assert (b.factorcount, b.primecount, b.checksum, b.factors) == want, job
        assert (r.factorcount, r.primecount, r.checksum, r.factors) == want, job


def long_n_mix():
    # n-ranges long enough to be cut over several lanes per prime (subranges of >= 1024 n)
    return [
        (10 ** 8, 10 ** 8 + 200000, 1000, 9000),        # FP64, factors present
        (2 ** 40, 2 ** 40 + 400000, 50000, 60000),      # FP64, setup-heavy subranges
        (2 ** 60, 2 ** 60 + 300000, 20000, 26000),      # hybrid below 2^64/3
        (2 ** 62 + 2 ** 61, 2 ** 62 + 2 ** 61 + 200000, 4000, 9000),  # hybrid between 2^64/3 and 2^63
        (2 ** 63 + 1000, 2 ** 63 + 100000, 3000, 9000), # hybrid above 2^63
        (127, 20000, 127, 30000),                       # p <= n: subranges that start at n >= p
        (2 ** 64 - 100000, 2 ** 64 - 1, 500, 4700),     # Montgomery
        (10 ** 9, 10 ** 9 + 100000, 5000, 5000),        # empty n-range
    ]


@pytest.mark.parametrize("lanes,unroll", [(2, 4), (3, 2), (4, 8), (8, 4)])
def test_batch_lanes_per_prime_match_plain(gpu, lanes, unroll):
    """A job cut into n-subranges on several lanes per prime (each subrange
    from its own setup program, only the last one folding the checksum) gives
    the plain sieve's factors, candidate count and checksum."""
    jobs = long_n_mix()
    print("lanes", lanes, "unroll", unroll, flush=True)
    res, _ = gpu.batch(jobs, lanes_per_prime=lanes, unroll=unroll)
    for job, o in zip(jobs, res):
        assert o.status == 0, job
        assert outcome(o) == plain(gpu, job), job


def test_batch_invalid_job_is_isolated(gpu):
    jobs = [(10 ** 8, 10 ** 8 + 100000, 1000, 2000), (1000, 500, 100, 200), (10 ** 8, 10 ** 8 + 100000, 1000, 2000)]
    res, _ = gpu.batch(jobs)
    assert res[1].status == gpu.FS_E_ARG
    assert res[0].status == res[2].status == 0 and outcome(res[0]) == outcome(res[2]) == plain(gpu, jobs[0])


def test_batch_autotune_returns_valid_shape(gpu):
    jobs = small_mix()[:8]
    o = gpu.batch_autotune(jobs)
    assert o["tile_candidates"] in (256, 512) and o["persistent_blocks"] > 0
    assert o["unroll"] in (2, 4, 8) and o["lanes_per_prime"] in (1, 2, 4) and o["lpt_order"] in (0, 1)
    res, _ = gpu.batch(jobs, **o)
    assert all(r.status == 0 for r in res)
    base, _ = gpu.batch(jobs)
    assert [outcome(r) for r in res] == [outcome(r) for r in base]


def drain(srv, ids, limit=20000):
    """Step the server until every id has a result; {id: JobOutcome}."""
    got = {}
    for _ in range(limit):
        srv.step()
        for i in ids:
            if i not in got:
                r = srv.poll(i)
                if r is not None:
                    got[i] = r
        if len(got) == len(ids):
            break
    return got


@pytest.mark.parametrize("budget,chunk,graph", [(0, 700, 1), (64 << 20, 5000, 1), (0, 1500, 0)])
def test_server_results_equal_alone(gpu, budget, chunk, graph):
    """Workunits arriving while others run are admitted as slots/memory free
    up; each result equals the same workunit run alone, with the quanta's
    launches replayed from a graph or enqueued one by one."""
    jobs = small_mix()
    srv = gpu.Server(mem_budget=budget, n_chunk=chunk, use_graph=graph)
    try:
        ids = [srv.submit(j) for j in jobs[:10]]
        srv.step()
        srv.step()
        ids += [srv.submit(j) for j in jobs[10:]]
        got = {}
        for _ in range(20000):
            srv.step()
            for i in ids:
                if i not in got:
                    r = srv.poll(i)
                    if r is not None:
                        got[i] = r
            if len(got) == len(ids):
                break
        assert len(got) == len(ids)
        st = srv.stats()
        assert st["active"] == 0 and st["queued"] == 0
        for i, job in zip(ids, jobs):
            assert got[i].status == 0
            assert outcome(got[i]) == plain(gpu, job), job
    finally:
        srv.close()


def test_server_small_workunits_match_oracle(gpu):
    """Seeded random workunits over every modulus size, served together in
    quanta of 500 n: every result against the oracle."""
    jobs = random_jobs(11, 12)
    srv = gpu.Server(n_chunk=500)
    try:
        ids = [srv.submit(j) for j in jobs]
        got = drain(srv, ids)
        assert len(got) == len(ids)
    finally:
        srv.close()
    for i, job in zip(ids, jobs):
        want = ol.sieve(*job, threads=8)
        assert got[i].status == 0, job
        assert (got[i].factorcount, got[i].primecount, got[i].checksum) == (want["factorcount"], want["primecount"],
                                                                          want["checksum"]), job
        assert got[i].factors == want["factors"], job


def test_server_submit_does_not_wait_for_a_step(gpu):
    """Submissions go to a queue with its own lock: while another thread runs
    steps whose quanta take tens of milliseconds, every submit returns within
    a few milliseconds, and all the work still comes out right."""
    import threading
    import time
    heavy = [(2 ** 62 + k * 2 ** 26, 2 ** 62 + k * 2 ** 26 + 2 ** 24, 500000, 900000) for k in range(4)]
    light = small_mix()[:12]
    srv = gpu.Server(n_chunk=100000)
    ids, times, stop = [], [], threading.Event()
    try:
        ids += [srv.submit(j) for j in heavy]
        step_times = []

        def stepper():
            while not stop.is_set():
                t = time.perf_counter()
                srv.step()
                step_times.append(time.perf_counter() - t)

        th = threading.Thread(target=stepper)
        th.start()
        time.sleep(0.05)
        for j in light:
            t = time.perf_counter()
            ids.append(srv.submit(j))
            times.append(time.perf_counter() - t)
            time.sleep(0.005)
        stop.set()
        th.join()
        got = drain(srv, ids)
    finally:
        srv.close()
    print(f"submit max {max(times) * 1e3:.2f} ms; step max {max(step_times) * 1e3:.1f} ms", flush=True)
    assert max(step_times) > 0.02 and max(times) < 0.005
    assert len(got) == len(ids)
    for i, job in zip(ids, heavy + light):
        assert got[i].status == 0
        assert outcome(got[i]) == plain(gpu, job), job


@pytest.mark.parametrize("flag", ["factor", "candidate"])
def test_server_overflow_restarts_workunits(gpu, flag):
    """A factor-buffer or candidate-region overflow (forced from the host for
    one harvested quantum) restarts that quantum's workunits from scratch;
    their results still equal the workunits run alone."""
    ctr = gpu.CTR_FACTOROVF if flag == "factor" else gpu.CTR_PRIMEOVF
    jobs = small_mix()[:10]
    srv = gpu.Server(n_chunk=1000)
    try:
        ids = [srv.submit(j) for j in jobs]
        srv.step()                      # first quantum enqueued
        gpu.check(gpu.lib().fs_inject_fault(ctr))
        srv.step()                      # its harvest sees the overflow
        got = drain(srv, ids)
    finally:
        gpu.lib().fs_inject_fault(0)
        srv.close()
    assert len(got) == len(ids)
    for i, job in zip(ids, jobs):
        assert got[i].status == 0, job
        assert outcome(got[i]) == plain(gpu, job), job


def test_server_gives_up_after_repeated_overflows(gpu):
    """A workunit whose every quantum overflows is restarted at most three
    times and then fails with FS_E_OVERFLOW; the server stays usable."""
    job = (10 ** 9, 10 ** 9 + 100000, 1000, 4000)
    srv = gpu.Server(n_chunk=1000)
    try:
        wid = srv.submit(job)
        res = None
        for _ in range(200):
            gpu.check(gpu.lib().fs_inject_fault(gpu.CTR_FACTOROVF))
            srv.step()
            res = srv.poll(wid)
            if res is not None:
                break
        gpu.lib().fs_inject_fault(0)
        assert res is not None and res.status == gpu.FS_E_OVERFLOW
        wid2 = srv.submit(job)
        got = drain(srv, [wid2])
        assert got[wid2].status == 0 and outcome(got[wid2]) == plain(gpu, job)
    finally:
        gpu.lib().fs_inject_fault(0)
        srv.close()
