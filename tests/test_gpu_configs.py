"""GPU parity at the benchmark configurations themselves.

- The 64-job mixed batch (seed 20261203, sigma 1.0) at full size, with the
  autotuned launch shape: every job's factors, candidate count and checksum
  equal the plain single-job sieve of the same job.
- The same 64 jobs with every p-range cut to its first 1/64: batched and plain
  against the oracle's fixture (tests/golden/c4_sub.json).
- Poisson serving of the 64 full-size jobs at 70 % of the measured batched
  capacity: every workunit equals the same workunit run alone.
"""
import json
import os
import time

import pytest

from factsieve_amd import workloads

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def outcome(r):
    return (r.factorcount, r.primecount, r.checksum, r.factors)


@pytest.fixture(scope="module")
def c4_jobs():
    return workloads.mixed_jobs(64)


@pytest.fixture(scope="module")
def c4_plain(gpu, c4_jobs):
    out = []
    for job in c4_jobs:
        r = gpu.sieve(*job, verify_threads=16)
        out.append(outcome(r))
    return out


def test_c4_full_batched_autotuned_matches_plain(gpu, c4_jobs, c4_plain):
    opts = gpu.batch_autotune(c4_jobs)
    print("autotuned", opts, flush=True)
    res, sec = gpu.batch(c4_jobs, **opts)
    print(f"batch of 64 full-size jobs: {sec:.3f} s", flush=True)
    assert len(res) == 64
    for job, o, want in zip(c4_jobs, res, c4_plain):
        assert o.status == 0, job
        assert outcome(o) == want, job
    # the default shape too
    res2, _ = gpu.batch(c4_jobs)
    assert [outcome(o) for o in res2] == c4_plain


def test_c4_sub_matches_oracle_fixture(gpu):
    with open(os.path.join(HERE, "golden", "c4_sub.json")) as f:
        fx = json.load(f)
    jobs = [tuple(j["range"]) for j in fx["jobs"]]
    res, _ = gpu.batch(jobs)
    for job, o, want in zip(jobs, res, fx["jobs"]):
        expect = (want["factorcount"], want["primecount"], int(want["checksum"], 16),
                  [tuple(x) for x in want["factors"]])
        assert o.status == 0, job
        assert outcome(o) == expect, job
        assert outcome(gpu.sieve(*job, verify_threads=16)) == expect, job


def test_c5_poisson_serving_full_size_equals_alone(gpu, c4_jobs, c4_plain):
    _, t_batch = gpu.batch(c4_jobs)
    rate = 0.7 * len(c4_jobs) / t_batch  # workunits per second at 70 % of batched capacity
    arrivals = workloads.poisson_arrivals(rate, len(c4_jobs))
    srv = gpu.Server()  # the default quantum, as the benchmark serves
    ids, got = {}, {}
    try:
        t0 = time.perf_counter()
        nxt = 0
        deadline = t0 + 120.0
        while len(got) < len(c4_jobs):
            assert time.perf_counter() < deadline, "serving did not finish in 120 s"
            now = time.perf_counter() - t0
            while nxt < len(c4_jobs) and arrivals[nxt] <= now:
                ids[srv.submit(c4_jobs[nxt])] = nxt
                nxt += 1
            st = srv.stats()
            if st["active"] == 0 and st["queued"] == 0 and nxt < len(c4_jobs):
                time.sleep(max(0.0, arrivals[nxt] - (time.perf_counter() - t0)))
                continue
            srv.step()
            for wid, k in ids.items():
                if k not in got:
                    r = srv.poll(wid)
                    if r is not None:
                        got[k] = r
        st = srv.stats()
        assert st["active"] == 0 and st["queued"] == 0
    finally:
        srv.close()
    for k, job in enumerate(c4_jobs):
        assert got[k].status == 0, job
        assert outcome(got[k]) == c4_plain[k], job
