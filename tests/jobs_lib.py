"""Seeded job generators shared by the GPU tests."""
import random


def random_jobs(seed, count):
    """`count` jobs (pmin, pmax, nmin, nmax): pmin uniform in [2^k, 2^(k+1)) for
    k uniform in [17, 63], p-ranges of 1e3-2e5 integers, n from below 2e4
    over up to 3,000 values; small enough for the oracle to finish in a second."""
    rng = random.Random(seed)
    jobs = []
    for _ in range(count):
        k = rng.randrange(17, 64)
        lo = rng.randrange(2 ** k, min(2 ** (k + 1), 2 ** 64 - 300000))
        n0 = rng.randrange(127, 20000)
        n1 = n0 + rng.randrange(1, 3000)
        if lo < n1:
            lo = n1 + 1
        jobs.append((lo, lo + rng.randrange(1000, 200000), n0, n1))
    return jobs


def wide_jobs(seed, count, max_nrange=4000, max_prange=60000):
    """`count` jobs with nmin drawn from three bands (below 2e3, 2e3-6e4, 6e4-2e5,
    so setup programs of every length), pmin uniform in [2^k, 2^(k+1)) for k
    uniform in [20, 63], up to `max_nrange` n-values and `max_prange` integers."""
    rng = random.Random(seed)
    jobs = []
    for _ in range(count):
        k = rng.randrange(20, 64)
        lo = rng.randrange(2 ** k, min(2 ** (k + 1), 2 ** 64 - 300000))
        n0 = rng.choice([rng.randrange(10, 2000), rng.randrange(2000, 60000), rng.randrange(60000, 200000)])
        n1 = n0 + rng.randrange(1, max_nrange)
        if lo < n1:
            lo = n1 + 1
        jobs.append((lo, lo + rng.randrange(500, max_prange), n0, n1))
    return jobs
