"""Synthetic workloads of the batched and serving benchmarks (seeded).

mixed_jobs: the 64-job mix -- n_start uniform in [1e3, 5e5], n-range length
lognormal (median 2e4, sigma 1.0), p_start = 2^k with k uniform in [32, 63],
p-range length lognormal (median 2^24 integers, sigma 1.0).
poisson_arrivals: arrival times of a Poisson process of the given rate.
"""
import math

import numpy as np

DEFAULT_SEED = 20261203


def mixed_jobs(n=64, seed=DEFAULT_SEED, sigma=1.0):
    rng = np.random.default_rng(seed)
    jobs = []
    for _ in range(n):
        nmin = int(rng.integers(1000, 500000, endpoint=True))
        nlen = max(1, int(round(rng.lognormal(math.log(2e4), sigma))))
        nlen = min(nlen, 2 ** 31 - 1 - nmin)
        k = int(rng.integers(32, 63, endpoint=True))
        plen = max(1, int(round(rng.lognormal(math.log(2 ** 24), sigma))))
        pmin = 2 ** k
        pmax = min(pmin + plen, 2 ** 64 - 1)
        jobs.append((pmin, pmax, nmin, nmin + nlen))
    return jobs


def poisson_arrivals(rate, n, seed=DEFAULT_SEED + 1):
    rng = np.random.default_rng(seed)
    return list(np.cumsum(rng.exponential(1.0 / rate, size=n)))
