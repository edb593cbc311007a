"""fs_iterate_noalloc, the iterate entry point whose multiplier and pair tables
live in the library's static device array instead of a table it allocates,
against fs_iterate on the same candidates and start residues: end residues,
factor records and flags bit-identical, on two calls in a row per path, the
second over fewer candidates and a shorter n-range (the array then still holds
the first call's longer tables)."""
import random

import pytest

pytestmark = pytest.mark.gpu

CASES = [
    # path name, (pmin, pmax, nmin, nmax) of the first call, then of the second (smaller) call
    ("fp64", (10 ** 12, 10 ** 12 + 200000, 100000, 140000), (10 ** 8, 10 ** 8 + 30000, 5000, 8001)),
    ("hybrid", (2 ** 62, 2 ** 62 + 150000, 500000, 520000), (2 ** 62 + 2 ** 40, 2 ** 62 + 2 ** 40 + 20000, 3000, 3999)),
    ("mont", (2 ** 64 - 300000, 2 ** 64 - 1, 2000, 6000), (2 ** 64 - 40000, 2 ** 64 - 1, 700, 1500)),
]


def starts_for(ps, n0, n1, seed):
    """Start residues: every fifth puts +-1 at a chosen n within the first 300
    steps (so the factor lists are not empty), the rest are random."""
    rng = random.Random(seed)
    out = []
    for i, p in enumerate(ps):
        if i % 5 == 0:
            k, sign = rng.randrange(0, min(300, n1 - n0)), rng.choice((1, -1))
            prod = 1
            for j in range(1, k + 1):
                prod = prod * (n0 + j) % p
            out.append(sign * pow(prod, -1, p) % p)
        else:
            out.append(rng.randrange(p))
    return out


def run(gpu, seg, n, n0, n1, path, starts, noalloc):
    seg.clearresult()
    seg.write_u64("res", starts)
    seg.write_u32("nnext", [n0] * n)
    (seg.iterate_noalloc if noalloc else seg.iterate)(n0, n1, path)
    ctr = seg.counters()
    return seg.read_u64("res", n), sorted(seg.factors()), ctr[gpu.CTR_VALIDATE], ctr[gpu.CTR_FACTOROVF]


@pytest.mark.parametrize("name,first,second", CASES)
def test_iterate_noalloc_matches_iterate(gpu, name, first, second):
    path = {"fp64": gpu.PATH_FP64, "hybrid": gpu.PATH_HYBRID, "mont": gpu.PATH_MONT}[name]
    assert gpu.select_path(first[1]) == path or path == gpu.PATH_MONT
    ncand = []
    for call, (lo, hi, n0, n1) in enumerate((first, second)):
        print(f"noalloc {name} call {call}: p=[{lo},{hi}) n=[{n0},{n1})", flush=True)
        seg = gpu.Segment(cap=16384, factor_cap=1 << 16)
        try:
            seg.clearresult()
            seg.clearn()
            seg.getsegprimes(lo, hi)
            n = seg.counters()[gpu.CTR_COUNT]
            assert 0 < n <= 16384
            ncand.append(n)
            starts = starts_for(seg.read_u64("p", n), n0, n1, seed=call)
            want = run(gpu, seg, n, n0, n1, path, starts, noalloc=False)
            got = run(gpu, seg, n, n0, n1, path, starts, noalloc=True)
            assert got[2:] == (0, 0) and want[2:] == (0, 0)
            assert len(want[1]) > 0
            assert got[0] == want[0]
            assert got[1] == want[1]
        finally:
            seg.close()
    assert ncand[1] < ncand[0] and second[3] - second[2] < first[3] - first[2]


def test_iterate_noalloc_rejects_too_long_range(gpu):
    """More than FS_ITERATE_NOALLOC_MAX_N n-values in one call is an argument
    error, raised before anything is launched."""
    seg = gpu.Segment(cap=4096)
    try:
        seg.clearresult()
        seg.clearn()
        seg.getsegprimes(10 ** 9, 10 ** 9 + 20000)
        with pytest.raises(gpu.FactSieveError) as e:
            seg.iterate_noalloc(1000, 1000 + (1 << 20) + 1, gpu.PATH_FP64)
        assert e.value.code == gpu.FS_E_ARG
        n = seg.counters()[gpu.CTR_COUNT]
        seg.write_u64("res", [0] * n)
        seg.write_u32("nnext", [1000] * n)
        seg.iterate_noalloc(1000, 1000 + (1 << 20), gpu.PATH_FP64)   # exactly the maximum is accepted
        seg.sync()
        assert seg.counters()[gpu.CTR_VALIDATE] == 0
        assert seg.read_u64("res", n) == [0] * n
        assert seg.read_u32("nnext", n) == [1000 + (1 << 20)] * n
    finally:
        seg.close()
