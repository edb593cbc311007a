"""bench.py's operation-count model (the VALU operations per step it prices
the iterate kernels' roofline with) against the kernels' own loop choices, and
its one-GPU contract.  The loop predicates are reached through the library's
host execution entry points, which accept a step exactly when the kernels'
per-wave vote would take that loop for a launch ending at n1."""
import math
import random
import subprocess
import sys

import bench


def c_fp64_loop(fs, p, n1):
    """(paired, exact) as fs_cand_iter_fp64 chooses them for one modulus p."""
    try:
        _, _, exact = fs.hostemu_fp_pair(p, 1, n1 - 1)
        return True, exact
    except fs.FactSieveError:
        pass
    try:
        fs.hostemu_fp_exact(p, 1, n1)
        return False, True
    except fs.FactSieveError:
        return False, False


def test_fp64_ops_model_matches_kernel_loops(fs):
    rng = random.Random(53)
    ps = [7, 101, 10 ** 8 + 7, 10 ** 12 + 39, 2 ** 40 + 15, 2 ** 46 + 15, 2 ** 53 - 111]
    ps += [rng.randrange(2 ** 10, 2 ** 53) | 1 for _ in range(60)]
    want = {(True, True): bench.OPS_PER_STEP_FP64_PAIR_EXACT, (True, False): bench.OPS_PER_STEP_FP64_PAIR,
            (False, True): bench.OPS_PER_STEP_FP64_EXACT, (False, False): bench.OPS_PER_STEP[1]}
    for p in ps:
        edges = {31 * p // 64, math.isqrt(2 ** 53 // p), 2 ** 53 // p, 2 ** 22}
        n1s = {n + d for n in edges for d in (-1, 0, 1)} | {rng.randrange(2, 2 ** 23)}
        for n1 in sorted(n1s):
            if not 2 <= n1 < 2 ** 31 or n1 >= p:
                continue
            assert bench.fp64_ops_at(p, p, n1) == want[c_fp64_loop(fs, p, n1)], (p, n1)


def test_hybrid_hi_model_matches_kernel_loop(fs):
    rng = random.Random(59)
    ps = [2 ** 53 + 5, (524300 << 34) + 7, 2 ** 58 + 27, 2 ** 62 + 135, 2 ** 63 - 25, 2 ** 63 + 29,
          2 ** 64 - 2 ** 58 - 1, 2 ** 64 - 2 ** 58 + 1, 2 ** 64 - 2 ** 46 - 3]
    ps += [rng.randrange(2 ** 53, 2 ** 64 - 2 ** 46) | 1 for _ in range(60)]
    for p in ps:
        top = p >> 34 if p < 2 ** 63 else p >> 39
        for n1 in {1, 1000, top - 1, top, top + 1, rng.randrange(1, 2 ** 31)}:
            if not 1 <= n1 < 2 ** 32:
                continue
            try:
                fs.hostemu_mulmod_hi(p, 5, n1)
                c_hi = True
            except fs.FactSieveError:
                c_hi = False
            assert bench.hybrid_hi(p, p + 1, n1) == c_hi, (p, n1)
    # a p-range straddling 2^63 takes the smaller limit of the two regimes
    assert bench.hybrid_hi(2 ** 63 - 1000, 2 ** 63 + 1000, 2 ** 24)
    assert not bench.hybrid_hi(2 ** 63 - 1000, 2 ** 63 + 1000, 2 ** 24 + 1)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, bench.__file__, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_hybrid_pair_model_matches_kernel_loop(fs):
    """bench.hybrid_pair against the predicate of the kernels' paired hybrid
    loop, reached through the host execution of one paired step (a + 1 = n1)."""
    rng = random.Random(71)
    ps = [2 ** 45 - 1, 2 ** 45 + 1, 2 ** 53 + 5, 2 ** 63 + 29, 2 ** 64 - 2 ** 58 - 1, 2 ** 64 - 2 ** 58 + 1,
          2 ** 64 - 2 ** 46 - 3]
    ps += [rng.randrange(2 ** 40, 2 ** 64 - 2 ** 46) | 1 for _ in range(40)]
    for p in ps:
        for n1 in (2, 1000, 2 ** 22 - 1, 2 ** 22, 2 ** 22 + 1, rng.randrange(2, 2 ** 23)):
            try:
                fs.hostemu_sh_pair(p, 5, n1 - 1)
                c_pair = True
            except fs.FactSieveError:
                c_pair = False
            assert bench.hybrid_pair(p, p + 1, n1) == c_pair, (p, n1)
