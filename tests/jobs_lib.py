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
