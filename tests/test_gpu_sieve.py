"""GPU parity: the HIP kernels and the sieve driver against the reference's
self-test vectors, the oracle's golden fixtures and full-size invariants.
Every call goes through the C ABI of libfactsieve.so."""
import json
import math
import os
import subprocess

import pytest

import oracle_lib as ol
from jobs_lib import random_jobs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
KATS = json.load(open(os.path.join(GOLDEN, "kats.json")))["cases"]
TINY = json.load(open(os.path.join(GOLDEN, "tiny.json")))["cases"]
C1 = json.load(open(os.path.join(GOLDEN, "c1.json")))


def paths_for(fs, pmax):
    if pmax <= 2 ** 53:
        return [fs.PATH_FP64, fs.PATH_HYBRID, fs.PATH_MONT]
    if pmax <= 2 ** 64 - 2 ** 46 + 1:
        return [fs.PATH_HYBRID, fs.PATH_MONT]
    return [fs.PATH_MONT]


def oracle_iterate(p, n0, n1, start, cap=64):
    """The oracle's iterate from a canonical residue: (end residue, factors)."""
    import ctypes
    O = ol.oracle()
    M = ol.mont(p)
    s6 = ol.u64(O.ora_mmul(start, M.r2, p, M.q))
    s7 = ol.u64(O.ora_mmul(n0 % p, M.r2, p, M.q))
    buf = (ol.OFactor * cap)()
    nf = O.ora_iterate(ctypes.byref(M), n0, n1, ctypes.byref(s6), ctypes.byref(s7), buf, cap)
    return O.ora_mmul(s6.value, 1, p, M.q), [(buf[i].p, buf[i].nc) for i in range(min(nf, cap))]


@pytest.mark.parametrize("k", range(5))
def test_reference_self_test_case(gpu, k):
    pmin, pmax, nmin, nmax, nf, pc, ck = KATS[k]
    print(f"KAT {k + 1}: p=[{pmin},{pmax}) n=[{nmin},{nmax})", flush=True)
    r = gpu.sieve(pmin, pmax, nmin, nmax, verify_threads=16)
    assert (r.factorcount, r.primecount, f"{r.checksum:016X}") == (nf, pc, ck)


@pytest.mark.parametrize("path", ["fp64", "hybrid", "mont"])
@pytest.mark.parametrize("k", range(5))
def test_reference_self_test_case_every_path(gpu, k, path):
    """The reference's five self-test vectors with each arithmetic formulation
    forced (all five ranges lie below 2^53, so every path covers them): the
    FP64, hybrid and Montgomery residues give the reference's counts and
    checksums alike."""
    pmin, pmax, nmin, nmax, nf, pc, ck = KATS[k]
    forced = {"fp64": gpu.PATH_FP64, "hybrid": gpu.PATH_HYBRID, "mont": gpu.PATH_MONT}[path]
    print(f"KAT {k + 1} on the {path} path", flush=True)
    r = gpu.sieve(pmin, pmax, nmin, nmax, verify_threads=16, path=forced)
    assert r.path == forced
    assert (r.factorcount, r.primecount, f"{r.checksum:016X}") == (nf, pc, ck)


def test_c1_golden_factor_list(gpu):
    r = gpu.sieve(*C1["range"])
    assert r.factorcount == 25 and r.primecount == 108529 and f"{r.checksum:016X}" == C1["checksum"]
    assert [list(f) for f in r.factors] == C1["factors"]


@pytest.mark.parametrize("ci", range(len(TINY)))
def test_tiny_cases_every_path(gpu, ci):
    case = TINY[ci]
    pmin, pmax, nmin, nmax = case["range"]
    for path in paths_for(gpu, pmax):
        print(f"tiny {ci} path {path}", flush=True)
        r = gpu.sieve(pmin, pmax, nmin, nmax, path=path)
        assert r.primecount == case["primecount"]
        assert f"{r.checksum:016X}" == case["checksum"]
        assert r.raw_factors == len(case["raw"])
        assert [list(f) for f in r.factors] == case["factors"]


@pytest.mark.parametrize("ci", range(len(TINY)))
def test_kernels_per_candidate(gpu, ci):
    """Each kernel on its own against the oracle's per-candidate residues."""
    case = TINY[ci]
    pmin, pmax, nmin, nmax = case["range"]
    cands = case["cands"]
    O = ol.oracle()
    for path in paths_for(gpu, pmax):
        print(f"kernels {ci} path {path}", flush=True)
        seg = gpu.Segment(cap=len(cands) + 4096)
        try:
            seg.clearresult()
            seg.clearn()
            seg.getsegprimes(pmin, pmax)
            n = seg.counters()[gpu.CTR_COUNT]
            assert n == len(cands)
            ps = seg.read_u64("p", n)
            assert sorted(ps) == [c[0] for c in cands]
            qs, r2s = seg.read_u64("q", n), seg.read_u64("r2", n)
            for p, q, r2 in zip(ps, qs, r2s):
                assert (p * q) % 2 ** 64 == 1 and r2 == 2 ** 128 % p
            seg.load_program(nmin, path)
            seg.setup(nmin, path)
            res = seg.read_u64("res", n)
            for p, r in zip(ps, res):
                assert r == O.ora_factorial_mod(nmin, p), (path, p)
            seg.verify(nmin)
            assert seg.counters()[gpu.CTR_POWERTABLE] == 0
            mid = (nmin + nmax) // 2
            seg.iterate(nmin, mid, path)
            seg.iterate(mid, nmax, path)
            res = seg.read_u64("res", n)
            want = {c[0]: c[1] * pow(2, -64, c[0]) % c[0] for c in cands}
            for p, r in zip(ps, res):
                assert r == want[p], (path, p)
            assert sorted(seg.factors()) == sorted(tuple(f) for f in case["raw"])
            seg.check(nmax)
            ctr, sums = seg.counters(), seg.sums()
            assert ctr[gpu.CTR_VALIDATE] == 0 and ctr[gpu.CTR_MAXCOUNT] == n
            assert sums[0] == n and f"{sums[1]:016X}" == case["device_checksum"]
        finally:
            seg.close()


@pytest.mark.parametrize("lo,hi", [(2 ** 40, 2 ** 40 + 60000), (2 ** 62, 2 ** 62 + 60000),
                                   (2 ** 64 - 2 ** 46 - 60000, 2 ** 64 - 2 ** 46 + 1), (2 ** 64 - 60000, 2 ** 64 - 1)])
def test_iterate_at_largest_n(gpu, lo, hi):
    """The iterate kernel of every path at the top of the n range (n up to
    2^31 - 1, the largest the factor records hold), from random start residues
    and from residues that put +-1 at chosen n, against the oracle."""
    import random
    n0, n1 = 2 ** 31 - 2900, 2 ** 31 - 1
    rng = random.Random(lo)
    for path in paths_for(gpu, hi):
        print(f"largest n: p=[{lo},{hi}) path {path}", flush=True)
        seg = gpu.Segment(cap=8192)
        try:
            seg.clearresult()
            seg.clearn()
            seg.getsegprimes(lo, hi)
            n = seg.counters()[gpu.CTR_COUNT]
            assert 0 < n <= 8192
            ps = seg.read_u64("p", n)
            starts, planted = [], {}
            for i, p in enumerate(ps):
                k, sign = rng.randrange(0, n1 - n0), rng.choice((1, -1))
                prod = 1
                if i % 7 == 0:      # +-1 at n0 + k: start = +-((n0+1)...(n0+k))^-1
                    for j in range(1, k + 1):
                        prod = prod * (n0 + j) % p
                if i % 7 == 0 and math.gcd(prod, p) == 1:
                    starts.append(sign * pow(prod, -1, p) % p)
                    planted[p] = (p, -(n0 + k) if sign == 1 else n0 + k)
                else:
                    starts.append(rng.randrange(p))
            seg.write_u64("res", starts)
            seg.write_u32("nnext", [n0] * n)
            seg.iterate(n0, n1, path)
            res = seg.read_u64("res", n)
            facs = sorted(seg.factors())
            assert seg.counters()[gpu.CTR_VALIDATE] == 0
            want_f = []
            for p, s0, r in zip(ps, starts, res):
                want_r, f = oracle_iterate(p, n0, n1, s0)
                assert r == want_r, (path, p)
                want_f += f
            assert facs == sorted(want_f)
            assert set(planted.values()) <= set(facs)
        finally:
            seg.close()


def test_validation_flags_iterate_out_of_order(gpu):
    """Launching iterate on a chunk that does not continue the residue's n is
    caught by the residue-position check (check.cl:66-78 validation)."""
    seg = gpu.Segment(cap=8192)
    try:
        seg.clearresult()
        seg.clearn()
        seg.getsegprimes(10 ** 8, 10 ** 8 + 20000)
        seg.load_program(1000, gpu.PATH_FP64)
        seg.setup(1000, gpu.PATH_FP64)
        seg.iterate(1001, 1100, gpu.PATH_FP64)   # skips n = 1000
        seg.check(1100)
        assert seg.counters()[gpu.CTR_VALIDATE] == 1
    finally:
        seg.close()


def full_size_fixture(name):
    """(factorcount, primecount, checksum, factors) of a full-size range as the
    CPU oracle computed it (tests/golden/make_golden.py c2 c3)."""
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", name + ".json")) as f:
        fx = json.load(f)
    return fx["factorcount"], fx["primecount"], int(fx["checksum"], 16), [tuple(x) for x in fx["factors"]]


def test_paths_agree_at_full_size_c2(gpu):
    """Full C2 (3.6e12 p*n steps): the FP64 (paired steps), hybrid and
    Montgomery (the reference's own formulation) paths are independent
    arithmetic; their checksums and factor lists must agree with each other and
    with the CPU oracle's fixture, and every factor is re-verified by the
    sieve's own verifier."""
    args = (10 ** 12, 10 ** 12 + 10 ** 9, 10 ** 5, 2 * 10 ** 5)
    a = gpu.sieve(*args, verify_threads=16, path=gpu.PATH_FP64)
    b = gpu.sieve(*args, verify_threads=16, path=gpu.PATH_HYBRID)
    c = gpu.sieve(*args, verify_threads=16, path=gpu.PATH_MONT)
    assert a.primecount == b.primecount == c.primecount == 36191004
    assert (a.checksum, a.factors) == (b.checksum, b.factors) == (c.checksum, c.factors)
    want = full_size_fixture("c2")
    assert (a.factorcount, a.primecount, a.checksum, a.factors) == want


def test_paths_agree_at_full_size_c3(gpu):
    """Full C3 (p in [2^62, 2^62 + 2^30), n in [5e5, 5.1e5)): the hybrid path
    (high-word quotient, two-stage REDC setup) against the Montgomery path and
    the CPU oracle's fixture."""
    args = (2 ** 62, 2 ** 62 + 2 ** 30, 500000, 510000)
    a = gpu.sieve(*args, verify_threads=16, path=gpu.PATH_HYBRID)
    b = gpu.sieve(*args, verify_threads=16, path=gpu.PATH_MONT)
    assert a.primecount == b.primecount == 24983780
    assert (a.checksum, a.factors) == (b.checksum, b.factors)
    want = full_size_fixture("c3")
    assert (a.factorcount, a.primecount, a.checksum, a.factors) == want


def test_c3_subrange_hybrid_vs_montgomery(gpu):
    """C3's modulus regime (p ~ 2^62, nmin = 5e5) on a 1/64 p-subrange: the
    hybrid path against the reference's Montgomery formulation."""
    lo = 2 ** 62
    args = (lo, lo + 2 ** 24, 500000, 510000)
    a = gpu.sieve(*args, path=gpu.PATH_HYBRID)
    b = gpu.sieve(*args, path=gpu.PATH_MONT)
    assert a.primecount == b.primecount == ol.oracle().ora_candidates(lo, lo + 2 ** 24, None, 0)
    assert (a.checksum, a.factors) == (b.checksum, b.factors)


def test_segmentation_and_chunking_invariance(gpu):
    g = C1["range"]
    r = gpu.sieve(*g, seg_candidates=5000, n_chunk=777)
    assert r.segments > 10 and r.iterate_launches > 100
    assert (r.factorcount, r.primecount, f"{r.checksum:016X}") == (25, 108529, C1["checksum"])
    assert [list(f) for f in r.factors] == C1["factors"]


@pytest.mark.parametrize("n_chunk", [777, 1000, 4096])
def test_paired_steps_chunked_match_oracle(gpu, n_chunk):
    """Paired FP64 steps over n-chunks of odd and even length (launches whose
    first multiplier has either parity, each reading its pairs at its own
    offset of the stream's pair table), against the oracle."""
    args = (4 * 10 ** 6, 5 * 10 ** 6, 11, 997)
    r = gpu.sieve(*args, n_chunk=n_chunk)
    o = ol.sieve(*args, threads=8)
    assert r.iterate_launches > 1 or n_chunk > 986
    assert (r.factorcount, r.primecount, r.checksum) == (o["factorcount"], o["primecount"], o["checksum"])
    assert r.factors == o["factors"]


@pytest.mark.parametrize("args", [
    (10 ** 8, 10 ** 8, 1000, 2000),          # empty p-range
    (10 ** 8, 10 ** 8 + 100000, 5000, 5000),  # empty n-range
    (114, 126, 114, 200),                     # no candidates
    (2 ** 64 - 200000, 2 ** 64 - 1, 1000, 1500),
    (2 ** 52 - 100000, 2 ** 52 + 100000, 3000, 3500),
    (2 ** 53 - 100000, 2 ** 53, 3000, 3500),                              # top of the FP64 range
    (2 ** 53 - 100000, 2 ** 53 + 100000, 3000, 3500),
    (2 ** 63 - 100000, 2 ** 63 + 100000, 3000, 3500),
    (2 ** 64 - 2 ** 46 - 100000, 2 ** 64 - 2 ** 46 + 1, 3000, 3500),   # top of the hybrid range
    (2 ** 64 - 2 ** 46 - 50000, 2 ** 64 - 2 ** 46 + 50000, 3000, 3500),  # straddles it: Montgomery
    (2 ** 40 - 100000, 2 ** 40 + 100000, 8000, 8193),   # FP64 paired steps (128 n1^2 <= 63 p)
    (51539509248 - 100000, 51539509248 + 100000, 174270, 174763),  # single steps: p * n1 crosses 2^53
    (4 * 10 ** 6, 6 * 10 ** 6, 10, 1000),                # paired steps with the exact product, ~30 factors
    (1600000000 - 100000, 1600000000 + 100000, 19500, 20000),  # 4 n1^2 = 1.6e9 inside: all paired
    (812698413 - 100000, 812698413 + 100000, 19500, 20000),  # 128 n1^2 = 63 p inside: paired and single waves
    (360287970 - 100000, 360287970 + 100000, 4000, 5000),  # p * n1^2 crosses 2^53: exact and inexact pairs
    ((16385 << 39) - 100000, (16385 << 39) + 100000, 16000, 16385),   # hybrid just above 2^53: high-word loop
    ((524300 << 34) - 100000, (524300 << 34) + 100000, 523800, 524300),  # p >> 34 crosses n1
    (2 ** 63 - 3000, 2 ** 63 + 3000, 2 ** 24 - 40, 2 ** 24 + 40),  # high-word limit drops from p >> 34 to p >> 39
    (2 ** 64 - 2 ** 58 - 100000, 2 ** 64 - 2 ** 58 + 100000, 3000, 3500),  # top of the high-word quotient
    ((2 ** 64 - 1) // 3 - 100000, (2 ** 64 - 1) // 3 + 100000, 3000, 3500),  # setup: two-stage REDC / REDC64
    (2 ** 50 - 100000, 2 ** 50 + 100000, 5000, 5200),   # FP64 setup: squares in FP64 below 2^50, Montgomery above
])
def test_edge_ranges_match_oracle(gpu, args):
    print("edge", args, flush=True)
    r = gpu.sieve(*args)
    o = ol.sieve(*args, threads=8)
    assert (r.factorcount, r.primecount, r.checksum) == (o["factorcount"], o["primecount"], o["checksum"])
    assert r.factors == o["factors"]


@pytest.mark.parametrize("args", [
    (10 ** 8, 10 ** 8 + 2 * 10 ** 5, 1000, 10000),               # FP64: loop cut at 9,487 inside the last lane
    (360287970 - 100000, 360287970 + 100000, 1000, 9000),         # FP64: exact/inexact pairs cut inside a lane
    (4 * 10 ** 6, 5 * 10 ** 6, 11, 9000),                         # FP64 exact pairs, many factors
    (2 ** 62, 2 ** 62 + 2 * 10 ** 5, 5000, 15000),                 # hybrid pairs
    ((16385 << 39) - 100000, (16385 << 39) + 100000, 2000, 12000),  # hybrid just above 2^53
    (2 ** 64 - 2 ** 58 - 100000, 2 ** 64 - 2 ** 58 + 100000, 3000, 13000),  # top of the high-word quotient
])
def test_lanes_per_prime_match_oracle(gpu, args):
    """One-segment jobs that underfill the GPU run as several lanes per prime
    (one setup and one iterate launch over all lanes, each lane from its own
    setup program, loop cuts taken inside the kernel): the factor list and
    checksum equal the single-lane run's and the oracle's."""
    print("lanes", args, flush=True)
    r = gpu.sieve(*args)
    one = gpu.sieve(*args, n_chunk=1 << 22)  # an explicit n-chunk keeps one lane
    o = ol.sieve(*args, threads=8)
    assert r.lanes > 1 and one.lanes <= 1
    assert (r.factorcount, r.primecount, r.checksum) == (o["factorcount"], o["primecount"], o["checksum"])
    assert r.factors == o["factors"] == one.factors and one.checksum == r.checksum


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_jobs_match_oracle(gpu, seed):
    """Seeded random jobs over every modulus size from 2^17 to 2^64 (every
    arithmetic path and iterate loop, and the launch cuts between loops)
    against the oracle."""
    for args in random_jobs(seed, 12):
        print("random", args, flush=True)
        r = gpu.sieve(*args)
        o = ol.sieve(*args, threads=8)
        assert (r.factorcount, r.primecount, r.checksum) == (o["factorcount"], o["primecount"], o["checksum"]), args
        assert r.factors == o["factors"]


@pytest.mark.parametrize("flag,code", [(3, -6), (5, -5), (6, -4)])
def test_fault_injection_is_reported(gpu, flag, code):
    gpu.check(gpu.lib().fs_inject_fault(flag))
    with pytest.raises(gpu.FactSieveError) as e:
        gpu.sieve(10 ** 8, 10 ** 8 + 200000, 1000, 2000)
    assert e.value.code == code
    gpu.lib().fs_inject_fault(0)


def test_candidate_overflow_recovers(gpu):
    """A candidate-buffer overflow flag makes the driver grow the buffer and
    redo the segment; the result is unchanged."""
    gpu.check(gpu.lib().fs_inject_fault(gpu.CTR_PRIMEOVF))
    r = gpu.sieve(*C1["range"])
    assert (r.factorcount, r.primecount, f"{r.checksum:016X}") == (25, 108529, C1["checksum"])


def test_gpu_verifier_matches_cpu(gpu):
    """fs_verify_factors_gpu against the CPU verifier: the factor lists of the
    reference self-test ranges 1 and 3 (1,071 factors with n up to 2e6; 42,770
    with p < 1e5), each alone and with planted wrong factors at several
    positions, and the edge cases the device leaves to the host (n < 2, n >= p,
    even p, p < 3) beside large p."""
    import random
    lists = []
    for k in (0, 2):
        pmin, pmax, nmin, nmax = KATS[k][:4]
        lists.append(gpu.sieve(pmin, pmax, nmin, nmax, verify_threads=16).factors)
    assert len(lists[0]) == 1071 and len(lists[1]) == 42770
    rng = random.Random(11)
    for fl in lists:
        assert gpu.verify_factors_gpu(fl) == -1 == gpu.verify_factors(fl, threads=16)
        for _ in range(4):
            bad = list(fl)
            i = rng.randrange(len(bad))
            p, nc = bad[i]
            bad[i] = (p, -nc) if rng.random() < 0.5 else (p, nc + (1 if nc > 0 else -1))
            want = gpu.verify_factors(bad, threads=16)
            assert want >= 0 and gpu.verify_factors_gpu(bad) == want, (i, bad[i])
    edge = [(1000003, 2000000), (1000003, -5), (100, 7), (2, 1), (4294967311, 70000), (2 ** 61 - 1, -100000),
            (2 ** 64 - 59, 3000), (2 ** 64 - 59, -1), (100003, 100002), (100003, -100001)]
    for c in edge:
        assert gpu.verify_factors_gpu([c]) == gpu.verify_factors([c], threads=2), c
    assert gpu.verify_factors_gpu([]) == -1
    # Wilson's theorem at the largest n: (p-1)! = -1 and (p-2)! = +1 (mod p) for
    # p = 2^31 - 1, 524,288 ranges of 4,096 multipliers each
    w = 2 ** 31 - 1
    assert gpu.verify_factors_gpu([(w, w - 1), (w, -(w - 2))]) == -1
    assert gpu.verify_factors_gpu([(w, w - 1), (w, w - 2)]) == 1


@pytest.mark.parametrize("mode", ["cpu", "gpu", "auto"])
def test_sieve_verify_device_modes(gpu, mode):
    """Reference self-test range 1 (1,071 factors, 1.6e9 verification
    multiplications: auto verifies on the GPU) with each verifier: the same
    factors, counts and checksum."""
    pmin, pmax, nmin, nmax, nf, pc, ck = KATS[0]
    m = {"cpu": gpu.VERIFY_CPU, "gpu": gpu.VERIFY_GPU, "auto": gpu.VERIFY_AUTO}[mode]
    r = gpu.sieve(pmin, pmax, nmin, nmax, verify_threads=16, verify_device=m)
    assert (r.factorcount, r.primecount, f"{r.checksum:016X}") == (nf, pc, ck)


def test_invalid_ranges_rejected(gpu):
    for args in [(10 ** 8, 10 ** 7, 1000, 2000), (1000, 2000, 1500, 1000), (200, 300, 500, 600)]:
        with pytest.raises(gpu.FactSieveError) as e:
            gpu.sieve(*args)
        assert e.value.code == gpu.FS_E_ARG


def test_cli_writes_reference_factors_file(gpu, tmp_path):
    exe = gpu.EXE_PATH
    r = subprocess.run([exe, "-n", "1000", "-N", "1e4", "-p", "1e8", "-P", "102e6", "-v", "4"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "factors 25, prime count 108529, checksum 000009F5FD6D6CFD" in r.stdout
    lines = open(tmp_path / "factors.txt").read().splitlines()
    want = [gpu.format_factor(p, nc) for p, nc in C1["factors"]] + [C1["checksum"]]
    assert lines == want
    # a finished workunit resumes straight to "Workunit complete."
    r2 = subprocess.run([exe, "-n", "1000", "-N", "1e4", "-p", "1e8", "-P", "102e6"], cwd=tmp_path,
                        capture_output=True, text=True, timeout=120)
    assert r2.returncode == 0 and "Workunit complete." in r2.stdout


def test_cli_self_test_passes(gpu, tmp_path):
    """`factsieve -s` is the reference's run_test (cl_sieve.cpp:1273-1391): the five
    self-test ranges against their expected factor counts, candidate counts and
    checksums; exit status 0 only when all pass."""
    print("cli -s", flush=True)
    r = subprocess.run([gpu.EXE_PATH, "-s", "-v", "16"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "All test cases completed successfully!" in r.stdout
    assert r.stderr.count("passed.") == 5 and "failed" not in r.stderr


def test_checkpoint_resume_matches_uninterrupted(gpu, tmp_path):
    g = C1["range"]
    opts = dict(write_files=1, resume=1, workdir=str(tmp_path), checkpoint_secs=0, seg_candidates=8000)
    part = gpu.sieve(*g, stop_after_segments=4, **opts)
    assert part.segments == 4
    assert os.path.exists(tmp_path / "stateA.ckp")
    full = gpu.sieve(*g, **opts)
    assert full.resumed == 1
    assert (full.factorcount, full.primecount, f"{full.checksum:016X}") == (25, 108529, C1["checksum"])
    lines = open(tmp_path / "factors.txt").read().splitlines()
    assert lines == [gpu.format_factor(p, nc) for p, nc in C1["factors"]] + [C1["checksum"]]


def test_resume_skips_corrupt_checkpoint(gpu, tmp_path):
    """The newest checkpoint is damaged: the run resumes from the other file
    (recomputing the segments in between) and still ends with the reference's
    count and checksum; every golden factor is in factors.txt."""
    g = C1["range"]
    opts = dict(write_files=1, resume=1, workdir=str(tmp_path), checkpoint_secs=0, seg_candidates=8000)
    gpu.sieve(*g, stop_after_segments=6, **opts)
    newest, which = gpu.read_checkpoint(tmp_path, *g)
    name = tmp_path / ("stateB.ckp" if which else "stateA.ckp")
    raw = bytearray(open(name, "rb").read())
    raw[56] ^= 0xFF  # state_sum
    open(name, "wb").write(raw)
    older, which2 = gpu.read_checkpoint(tmp_path, *g)
    assert which2 != which and older.p < newest.p
    full = gpu.sieve(*g, **opts)
    assert full.resumed == 1
    assert (full.factorcount, full.primecount, f"{full.checksum:016X}") == (25, 108529, C1["checksum"])
    lines = open(tmp_path / "factors.txt").read().splitlines()
    assert lines[-1] == C1["checksum"]
    assert set(lines[:-1]) == {gpu.format_factor(p, nc) for p, nc in C1["factors"]}


def test_missing_factor_lines_detected(gpu, tmp_path):
    """factors.txt lost lines between checkpoint and the end: the line-count
    check of finalizeResults fails with FS_E_IO (cl_sieve.cpp:742-746)."""
    g = C1["range"]
    opts = dict(write_files=1, resume=1, workdir=str(tmp_path), checkpoint_secs=0, seg_candidates=8000)
    part = gpu.sieve(*g, stop_after_segments=10, **opts)
    st, _ = gpu.read_checkpoint(tmp_path, *g)
    assert st.factorcount > 0
    open(tmp_path / "factors.txt", "w").close()
    with pytest.raises(gpu.FactSieveError) as e:
        gpu.sieve(*g, **opts)
    assert e.value.code == gpu.FS_E_IO and "Missing factors" in str(e.value)


def test_repeated_sieves_keep_device_memory_flat(gpu):
    """Each plain sieve's stream-bound multiplier table (here 2^21 n per
    launch, about 50 MB) is released with its stream: free device memory
    does not drift over repeated calls."""
    job = (10 ** 9, 10 ** 9 + 20000, 10 ** 6, 10 ** 6 + 2 ** 21)
    first = gpu.sieve(*job)
    free0 = gpu.device_info()["free_mem"]
    for _ in range(12):
        r = gpu.sieve(*job)
        assert (r.checksum, r.factors) == (first.checksum, first.factors)
    drift = free0 - gpu.device_info()["free_mem"]
    print(f"free memory drift over 12 calls: {drift / 2 ** 20:.1f} MiB", flush=True)
    assert drift < 64 * 2 ** 20


def test_concurrent_sieves_from_host_threads(gpu):
    """fs_sieve called from four host threads at once (each call leases its own
    stream and buffers; verification threads are shared): every result equals
    the same job run alone."""
    import threading
    jobs = [tuple(C1["range"]), (10 ** 8, 10 ** 8 + 300000, 1000, 3000), (2 ** 40, 2 ** 40 + 200000, 20000, 24000),
            (2 ** 62, 2 ** 62 + 100000, 5000, 9000), (4 * 10 ** 6, 5 * 10 ** 6, 11, 997),
            ((16385 << 39) - 100000, (16385 << 39) + 100000, 2000, 12000)]
    want = {j: gpu.sieve(*j, verify_threads=4) for j in jobs}
    got, errors = [], []

    def worker(t):
        try:
            for k in range(6):
                j = jobs[(t + k) % len(jobs)]
                got.append((j, gpu.sieve(*j, verify_threads=4)))
        except Exception as e:  # reported below, on the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=120)
    assert not errors, errors
    assert len(got) == 24
    for j, r in got:
        w = want[j]
        assert (r.factorcount, r.primecount, r.checksum, r.factors) == \
            (w.factorcount, w.primecount, w.checksum, w.factors), j


def reference_order(factors):
    """True when a factor list is in the reference's order: p ascending, then
    |n| ascending (factorcompare, cl_sieve.cpp:421-436), with no duplicates."""
    keys = [(p, abs(nc)) for p, nc in factors]
    return all(a < b for a, b in zip(keys, keys[1:]))


@pytest.mark.parametrize("lo,n0,note", [
    (10 ** 12, 100000, "FP64 inexact pairs, a(a+1) < p"),
    (10 ** 8, 5000, "FP64 exact pairs (p n1^2 <= 2^53)"),
    (2 ** 32 + 1, 300000, "FP64 inexact pairs, a(a+1) far above p"),
    (2 ** 62, 500000, "hybrid pairs"),
    (2 ** 53 + 2 ** 20, 3000, "hybrid pairs just above 2^53"),
    (2 ** 64 - 2 ** 58 - 10 ** 6, 2 ** 22 - 64, "hybrid pairs at the top of their range"),
])
def test_paired_loop_planted_every_position(gpu, lo, n0, note):
    """The paired iterate kernels (16 steps per block, one product by
    (n+1)(n+2) per two steps; FP64: ballots of |v| == 1 and |v'| == n+2;
    hybrid: a filter on the quotient estimates) from start residues that put
    +-1 at every position of a block, both signs, in three consecutive blocks
    and the tail; end residues and factor lists against the oracle."""
    n1 = n0 + 3 * 16 + 7
    path = gpu.PATH_FP64 if lo + 60000 <= 2 ** 53 else gpu.PATH_HYBRID
    assert 64 * n1 <= 31 * lo and (path == gpu.PATH_FP64 or n1 <= 2 ** 22)
    seg = gpu.Segment(cap=8192)
    try:
        print(f"planted pairs: p from {lo}, n=[{n0},{n1}) ({note})", flush=True)
        seg.clearresult()
        seg.clearn()
        seg.getsegprimes(lo, lo + 60000)
        n = seg.counters()[gpu.CTR_COUNT]
        assert 200 <= n <= 8192
        ps = seg.read_u64("p", n)
        starts, planted = [], set()
        for i, p in enumerate(ps):
            k, sign = i % (n1 - n0), (1 if (i // (n1 - n0)) % 2 == 0 else -1)
            prod = 1
            for j in range(1, k + 1):
                prod = prod * (n0 + j) % p
            if i < 4 * (n1 - n0) and math.gcd(prod, p) == 1:
                starts.append(sign * pow(prod, -1, p) % p)
                planted.add((p, -(n0 + k) if sign == 1 else n0 + k))
            else:
                starts.append((i * 7919 + 12345) % p)
        covered = {(abs(nc) - n0, nc < 0) for _, nc in planted}
        assert len(covered) == 2 * (n1 - n0)      # every position, both signs
        seg.write_u64("res", starts)
        seg.write_u32("nnext", [n0] * n)
        seg.iterate(n0, n1, path)
        res = seg.read_u64("res", n)
        facs = seg.factors()
        assert seg.counters()[gpu.CTR_VALIDATE] == 0
        want_f = []
        for p, s0, r in zip(ps, starts, res):
            want_r, f = oracle_iterate(p, n0, n1, s0)
            assert r == want_r, p
            want_f += f
        assert sorted(facs) == sorted(want_f)
        assert planted <= set(facs)
    finally:
        seg.close()


@pytest.mark.parametrize("fault", ["residue", "power_table"])
def test_verify_kernels_detect_wrong_setup(gpu, fault):
    """The verify kernels (verifyslow/pow/reduce/result) compute nmin! mod p of
    the segment's first candidate by the direct product and by the reference
    power table and compare both with the setup kernel's residue
    (verifyresult.cl -> cl_sieve.cpp:903-907).  A corrupted residue, or a
    power table with one exponent off, raises the power-table flag; the
    untouched segment does not."""
    import ctypes
    L = gpu.lib()
    nmin = 5000
    seg = gpu.Segment(cap=8192)
    pb, pe = ctypes.c_void_p(), ctypes.c_void_p()
    try:
        seg.clearresult()
        seg.clearn()
        seg.getsegprimes(10 ** 9, 10 ** 9 + 20000)
        seg.load_program(nmin, gpu.PATH_FP64)
        seg.setup(nmin, gpu.PATH_FP64)
        seg.sync()
        seg.verify(nmin)
        assert seg.counters()[gpu.CTR_POWERTABLE] == 0
        b, e, _ = gpu.power_table(nmin)
        if fault == "residue":
            r0 = seg.read_u64("res", 1)[0]
            p0 = seg.read_u64("p", 1)[0]
            seg.write_u64("res", [(r0 + 1) % p0])
        else:
            e = list(e)
            e[len(e) // 2] += 1
        gpu.check(L.fs_malloc(ctypes.byref(pb), len(b) * 8))
        gpu.check(L.fs_malloc(ctypes.byref(pe), len(e) * 4))
        ab, ae = (gpu.u64 * len(b))(*b), (gpu.u32 * len(e))(*e)
        gpu.check(L.fs_memcpy_htod(pb, ab, len(b) * 8, seg.stream, 1))
        gpu.check(L.fs_memcpy_htod(pe, ae, len(e) * 4, seg.stream, 1))
        gpu.check(L.fs_verifyslow(ctypes.byref(seg.c), seg.ctr, seg.ver, nmin, seg.stream))
        gpu.check(L.fs_verifypow(ctypes.byref(seg.c), seg.ctr, pb, pe, len(b), seg.ver, seg.stream))
        gpu.check(L.fs_verifyreduce(ctypes.byref(seg.c), seg.ctr, seg.ver, seg.stream))
        gpu.check(L.fs_verifyresult(ctypes.byref(seg.c), seg.ctr, seg.ver, seg.stream))
        seg.sync()
        assert seg.counters()[gpu.CTR_POWERTABLE] == 1
    finally:
        L.fs_free(pb)
        L.fs_free(pe)
        seg.close()


def test_factor_lists_in_reference_order(gpu):
    """Kept factors come back sorted as the reference sorts them: p ascending,
    then |n| (cl_sieve.cpp:421-436); the golden lists are in that order too."""
    for case in TINY + [C1]:
        assert reference_order([tuple(f) for f in case["factors"]])
    r = gpu.sieve(127, 20000, 127, 30000)   # p <= nmax: many factors per p
    assert r.factorcount > 1000 and reference_order(r.factors)
    res, _ = gpu.batch([(127, 20000, 127, 30000), (10 ** 8, 10 ** 8 + 300000, 1000, 3000)])
    assert all(reference_order(o.factors) for o in res)


@pytest.mark.parametrize("standalone", [0, 1])
def test_boinc_hooks(gpu, tmp_path, standalone):
    """A run with BOINC hooks (cl_sieve.cpp:47-69, 225-232, 1119-1139,
    1249-1256): fraction done never decreases and ends at 1.0, critical
    sections nest properly and hold every checkpoint, and a run resumed from a
    checkpoint whose last trickle is over a day old sends exactly one
    trickle-up (none when standalone).  The results equal the plain run's."""
    import re
    g = C1["range"]
    calls = []

    def fraction(fd, user):
        calls.append(("fraction", fd))

    def begin(user):
        calls.append(("begin",))

    def end(user):
        calls.append(("end",))

    def done(user):
        calls.append(("checkpoint",))

    def trickle(variety, text, user):
        calls.append(("trickle", variety.decode(), text.decode()))
        return 0

    cbs = (gpu.FRACTION_FN(fraction), gpu.SECTION_FN(begin), gpu.SECTION_FN(end), gpu.SECTION_FN(done),
           gpu.TRICKLE_FN(trickle))
    hooks = gpu.BoincHooks(standalone, *cbs, 3600.0, None)
    st = gpu.WorkStatus(pmin=g[0], pmax=g[1], nmin=g[2], nmax=g[3], p=g[0], last_trickle=0)
    gpu.write_checkpoint(tmp_path, st, 0)
    opts = dict(write_files=1, resume=1, workdir=str(tmp_path), checkpoint_secs=0, seg_candidates=8000)
    r = gpu.sieve(*g, boinc=hooks, **opts)
    assert r.resumed == 1 and r.segments > 2
    assert (r.factorcount, r.primecount, f"{r.checksum:016X}") == (25, 108529, C1["checksum"])
    fr = [c[1] for c in calls if c[0] == "fraction"]
    assert fr and fr[-1] == 1.0 and all(a <= b for a, b in zip(fr, fr[1:]))
    depth = 0
    for c in calls:
        depth += {"begin": 1, "end": -1}.get(c[0], 0)
        assert 0 <= depth <= 1
        if c[0] in ("checkpoint", "trickle"):
            assert depth == 1
    assert depth == 0 and sum(1 for c in calls if c[0] == "checkpoint") >= 2
    tr = [c for c in calls if c[0] == "trickle"]
    if standalone:
        assert not tr
    else:
        assert len(tr) == 1 and tr[0][1] == "factsieve_progress"
        m = re.fullmatch(r"<trickle_up>\n   <progress>([0-9.]+)</progress>\n   <cputime>([0-9.]+)</cputime>\n"
                         r"   <runtime>([0-9.]+)</runtime>\n</trickle_up>\n", tr[0][2])
        assert m and 0.0 <= float(m.group(1)) <= 1.0 and float(m.group(3)) >= 3600.0
    last, _ = gpu.read_checkpoint(tmp_path, *g)
    assert (last.last_trickle > 0) == (not standalone)
