"""Host logic of the product library on the CPU: exports of the C ABI, the
setup programs, the CPU verifier, number parsing, and the three residue
arithmetic formulations executed on the host, all against the oracle."""
import ctypes
import json
import math
import os
import random
import struct
import re
import subprocess

import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def declared_functions():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"\b(fs_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol(fs):
    L = fs.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts(fs):
    assert ctypes.sizeof(fs.Factor) == 16
    assert ctypes.sizeof(fs.Cands) == 48


def test_no_device_gives_clean_error(fs):
    # this container has no GPU: init must fail with an error code, not crash
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    rc = fs.lib().fs_init(0)
    assert rc == fs.FS_E_NODEV and "device" in fs.last_error()


def test_select_path(fs):
    assert fs.select_path(2 ** 52) == fs.PATH_FP64
    assert fs.select_path(2 ** 53) == fs.PATH_FP64
    assert fs.select_path(2 ** 53 + 1) == fs.PATH_HYBRID
    assert fs.select_path(2 ** 63) == fs.PATH_HYBRID
    assert fs.select_path(2 ** 63 + 1) == fs.PATH_HYBRID
    # the signed hybrid step needs (1/2 + 2^-19) p < 2^63: p <= 2^64 - 2^46
    assert fs.select_path(2 ** 64 - 2 ** 46 + 1) == fs.PATH_HYBRID
    assert fs.select_path(2 ** 64 - 2 ** 46 + 2) == fs.PATH_MONT


# 524291 = 2 * 2^18 + 3 and 1048579 = 4 * 2^18 + 3 start the second and third
# 2^18-odd-number segments of the library's prime sieve
@pytest.mark.parametrize("nmin", [127, 1000, 10000, 100000, 500000, 524289, 524291, 524293, 1048579, 3000017])
def test_power_table_matches_oracle(fs, nmin):
    O = ol.oracle()
    B = ctypes.POINTER(ctypes.c_uint64)()
    E = ctypes.POINTER(ctypes.c_uint32)()
    C = ctypes.POINTER(ctypes.c_uint32)()
    O.ora_power_table.restype = ctypes.c_uint64
    m = O.ora_power_table(nmin, ctypes.byref(B), ctypes.byref(E), ctypes.byref(C))
    b, e, c = fs.power_table(nmin)
    assert len(b) == m
    assert b == list(B[:m]) and e == list(E[:m]) and c == list(C[:m])


def test_power_table_term_counts(fs):
    # merged term counts (run-length merge of equal exponents below 2^64-1)
    want = {127: 14, 1000: 46, 10000: 313, 100000: 2770, 500000: 13586, 1000000: 26032, 2000000: 49682}
    for nmin, terms in want.items():
        assert len(fs.power_table(nmin)[0]) == terms


@pytest.mark.parametrize("path", [1, 2])
@pytest.mark.parametrize("nmin", [0, 1, 2, 3, 127, 1000, 4096, 20000])
def test_setup_program_evaluates_to_factorial(fs, path, nmin):
    import struct
    n, words = fs.setup_program(nmin, path)
    assert n % 4 == 0
    bound = 2 ** 50 if path == 1 else 2 ** 60
    acc, e62, e64 = 1, 0, 0
    for i, w in enumerate(words[:n]):
        sq = w >> 63
        assert not sq or i % 4 == 0          # squares only open a 4-term block
        w &= (1 << 63) - 1
        m = int(struct.unpack("<d", struct.pack("<Q", w))[0]) if path == 1 else w
        assert 1 <= m < bound
        if sq:
            acc, e62, e64 = acc * acc, 2 * e62 + 64, 2 * e64 + 64
        acc, e62, e64 = acc * m, e62 + 62, e64 + 64
    assert acc == math.factorial(nmin)
    if path == 2:
        # exponents of the Montgomery factors: multiplies by 2^-62 or 2^-64, squares by 2^-64
        assert words[n:n + 2] == [e62, e64]


def test_hybrid_setup_multiply_at_edges(fs):
    """The hybrid setup products s*m*2^-k mod p against big integers at the
    edges of their bounds: below 2^64/3 the two-stage reduction by 2^31 (k = 62)
    at s = 2p - 1, m = 2^60 - 1 and p just under 2^64/3, where its first stage
    comes closest to 2^64; between 2^64/3 and 2^63 the lazy REDC by 2^64; above,
    the canonical form; squares by 2^64 in every range.  Lazy values stay below
    2p, canonical ones below p."""
    rng = random.Random(17)
    third = (2 ** 64 - 1) // 3
    ps = [third - 2 if (third - 2) % 2 else third - 1, 2 ** 62 + 2 ** 30 + 3, 2 ** 62 + 135, 2 ** 62 - 57,
          2 ** 53 + 5, 2 ** 32 + 15, 2 ** 32 - 5, 1000003, 127, 3,
          third + 2 if (third + 2) % 2 else third + 1, 2 ** 63 - 25, 2 ** 63 - 259,
          2 ** 64 - 2 ** 46 - 3, 2 ** 63 + 29]
    for p in ps:
        assert p % 2
        stop = 2 * p if p < 2 ** 63 else p
        ss = [stop - 1, stop - 2, p - 1, p, p + 1, 1, 0, 2 ** 32 - 1, 2 ** 32, 2 ** 31 - 1] + \
             [rng.randrange(0, stop) for _ in range(300)]
        ms = [2 ** 60 - 1, 2 ** 60 - 2, 2 ** 31 - 1, 2 ** 31, 2 ** 32 - 1, 1, 0, 2 ** 59 + 2 ** 31 - 1] + \
             [rng.randrange(0, 2 ** 60) for _ in range(300)]
        for i, s in enumerate(ss):
            if s >= stop:
                continue
            for m in ms[:8] + [ms[8 + i % 300]]:
                raw, res, k = fs.hostemu_setup_mul(p, s, m)
                assert k == (62 if p < third else 64)
                assert res == s * m * pow(2, -k, p) % p, (p, s, m)
                assert raw % p == res and raw < stop, (p, s, m, raw)
            raw, res, k = fs.hostemu_setup_mul(p, s, 0, square=True)
            assert k == 64 and res == s * s * pow(2, -64, p) % p and raw % p == res and raw < stop, (p, s)
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_setup_mul(2 ** 62 + 135, 2 * (2 ** 62 + 135), 5)   # s >= 2p
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_setup_mul(2 ** 62 + 135, 5, 2 ** 60)                # m >= 2^60
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_setup_mul(2 ** 63 + 29, 2 ** 63 + 29, 5)            # canonical form: s >= p


def test_is_prime_and_verify_match_oracle(fs):
    O = ol.oracle()
    rng = random.Random(2)
    vals = [rng.getrandbits(64) | 1 for _ in range(3000)] + ol.candidates(10 ** 12, 10 ** 12 + 5000)
    vals += json.load(open(os.path.join(GOLDEN, "spsp2_survivors_1e6.json")))["values"]
    # around the three-base bound 4,759,123,141 (itself composite, 48781 * 97561) and
    # strong pseudoprimes to several small bases (3215031751 = 151 * 751 * 28351)
    vals += [rng.getrandbits(33) | 1 for _ in range(3000)] + list(range(4759123141 - 400, 4759123141 + 400, 2))
    vals += [3215031751, 2152302898747, 3474749660383, 341550071728321]
    assert not fs.is_prime(4759123141) and not fs.is_prime(3215031751)
    for v in vals:
        assert fs.is_prime(v) == bool(O.ora_is_prime(v))
    assert [v for v in range(0, 200) if fs.is_prime(v)] == [v for v in range(2, 200)
                                                             if all(v % d for d in range(2, v))]
    for p, nc in json.load(open(os.path.join(GOLDEN, "c1.json")))["factors"]:
        n, c = abs(nc), (-1 if nc < 0 else 1)
        assert fs.verify_factor(p, n, c) and not fs.verify_factor(p, n, -c) and not fs.verify_factor(p, n + 1, c)
    for p in [2 ** 62 + 135, 2 ** 64 - 59, 1000003, 4294967311]:
        for n in [5, 6, 33, 34, 35, 36, 200]:  # 2..n holds 4k, 4k+1, 4k+2 and 4k+3 multipliers
            for c in (-1, 1):
                assert fs.verify_factor(p, n, c) == bool(O.ora_verify(p, n, c))


def test_verify_factor_list_matches_single(fs):
    """The threaded list verifier (ranges of each factor's n! on all threads)
    agrees with the one-factor verifier: real factors from the golden C1 list
    and KAT3-like small p, each with a planted wrong twin; n across several
    2^15 ranges; n >= p; even p; p above 2^32."""
    O = ol.oracle()
    good = [tuple(f) for f in json.load(open(os.path.join(GOLDEN, "c1.json")))["factors"]]
    rng = random.Random(9)
    cases = list(good)
    for p, nc in good[:10]:
        cases += [(p, -nc), (p, nc + (1 if nc > 0 else -1))]
    # (p-1)! = -1 and (p-2)! = +1 (mod p): Wilson, n up to ~2e5 (7 ranges)
    for p in [100003, 199999, 131071, 65537, 4294967311]:
        if p < 300000:
            cases += [(p, p - 1), (p, -(p - 2)), (p, -(p - 1)), (p, p - 2)]
    cases += [(1000003, 2000000), (1000003, -5), (100, 7), (4294967311, 70000), (2 ** 61 - 1, -100000)]
    rng.shuffle(cases)
    want = [fs.verify_factor(p, abs(nc), -1 if nc < 0 else 1) for p, nc in cases]
    for i, (p, nc) in enumerate(cases):
        if p < 2 ** 63 and p % 2 and abs(nc) < 300000:
            assert want[i] == bool(O.ora_verify(p, abs(nc), -1 if nc < 0 else 1)), (p, nc)
    first_bad = want.index(False) if False in want else -1
    assert fs.verify_factors(cases, threads=8) == first_bad
    ok = [c for c, w in zip(cases, want) if w]
    assert len(ok) > 30 and fs.verify_factors(ok, threads=8) == -1
    assert fs.verify_factors(ok, threads=1) == -1
    assert fs.verify_factors([], threads=4) == -1


def test_verify_factors_from_concurrent_callers(fs):
    """The verifier's worker threads are shared: callers on six host threads at
    once (each call one parallel region, entered one at a time) all get the
    answers a lone caller gets, good lists and lists with a planted bad factor."""
    import threading
    good = [tuple(f) for f in json.load(open(os.path.join(GOLDEN, "c1.json")))["factors"]]
    bad = list(good)
    bad[7] = (bad[7][0], -bad[7][1])
    lists = [(good, -1), (bad, 7), (good[:5], -1), (bad[3:12], 4)]
    errors = []

    def worker(t):
        try:
            for k in range(40):
                lst, want = lists[(t + k) % len(lists)]
                got = fs.verify_factors(lst, threads=1 + (t + k) % 8)
                if got != want:
                    errors.append((t, k, got, want))
        except Exception as e:  # reported on the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(t,)) for t in range(6)]
    for th in threads:
        th.start()
    for th in threads:
        th.join(timeout=300)
    assert not errors, errors[:5]


def test_verify_factors_null_list_is_an_error(fs):
    """A null list with n > 0 must not read as "all pass" (-1)."""
    assert fs.lib().fs_verify_factors(None, 3, 2) == fs.FS_E_VERIFY
    assert "null" in fs.last_error()


def test_parse_matches_oracle(fs):
    O = ol.oracle()
    strs = ["127", "1e8", "4K", "4M", "1k", "1g", "3b62", "0x7f", "12x", "1e99", "99999999999999999999",
            "18446744073709551615", "1b64", "abc", "", "1E3", "1B4", "2T", "1P", "1p", "5Kx"]
    for s in strs:
        for lo, hi in [(127, 2 ** 64 - 2), (128, 2 ** 64 - 1), (127, 2 ** 31 - 2)]:
            a = ol.u64(0)
            ra = O.ora_parse_uint64(ctypes.byref(a), s.encode(), lo, hi)
            rb, vb = fs.parse_uint64(s, lo, hi)
            assert (ra, a.value if ra == 0 else None) == (rb, vb if rb == 0 else None), (s, lo, hi)


EDGE_RANGES = [
    (10 ** 8, 10 ** 8 + 4000, [1, 2, 3]),
    ((1 << 52) - 3000, 1 << 52, [1, 2, 3]),
    ((1 << 52) + 1, (1 << 52) + 3000, [1, 2, 3]),
    ((1 << 53) - 3000, 1 << 53, [1, 2, 3]),
    ((1 << 53) + 1, (1 << 53) + 3000, [2, 3]),
    ((1 << 62) - 3000, 1 << 62, [2, 3]),
    ((1 << 63) - 3000, 1 << 63, [2, 3]),
    ((1 << 63) + 1, (1 << 63) + 3000, [2, 3]),
    ((1 << 64) - (1 << 46) - 3000, (1 << 64) - (1 << 46) + 1, [2, 3]),
    ((1 << 64) - 3000, (1 << 64) - 1, [3]),
]


@pytest.mark.parametrize("lo,hi,paths", EDGE_RANGES)
def test_host_arithmetic_matches_oracle(fs, lo, hi, paths):
    """The kernels' inline arithmetic, executed on the host, against the
    oracle's Montgomery residues, at the edges of each formulation's range and
    for n up to 2^31 - 1."""
    O = ol.oracle()
    rng = random.Random(lo)
    ps = ol.candidates(lo, hi)
    assert ps
    for p in rng.sample(ps, min(6, len(ps))):
        for nmin, span in [(127, 2000), (5000, 1500), (2 ** 31 - 1500, 1499)]:
            want = O.ora_factorial_mod(nmin, p) if nmin < 10 ** 6 else None
            if want is None:
                # start from a random residue instead of nmin!
                want = rng.randrange(p)
                start = want
            for path in paths:
                if nmin < 10 ** 6:
                    start = fs.hostemu_setup(path, p, nmin)
                    assert start == want, (path, p, nmin)
                if path == fs.PATH_HYBRID and start and nmin + span >= 2 * p:
                    # outside the hybrid step's multiplier bound m < 2p, which
                    # the sieve keeps (its residues are 0 from n = p on)
                    with pytest.raises(fs.FactSieveError):
                        fs.hostemu_iterate(path, p, nmin, nmin + span, start)
                    continue
                r, facs, nf = fs.hostemu_iterate(path, p, nmin, nmin + span, start)
                M = ol.mont(p)
                s6 = ol.u64(O.ora_mmul(start, M.r2, p, M.q))
                s7 = ol.u64(O.ora_mmul(nmin % p, M.r2, p, M.q))
                buf = (ol.OFactor * 64)()
                onf = O.ora_iterate(ctypes.byref(M), nmin, nmin + span, ctypes.byref(s6), ctypes.byref(s7), buf, 64)
                assert r == O.ora_mmul(s6.value, 1, p, M.q), (path, p, nmin)
                assert facs == [(buf[i].p, buf[i].nc) for i in range(onf)]


def test_host_arithmetic_planted_factor(fs):
    """Start residues at +1 and -1: every formulation must report the factor."""
    for p, paths in [(1000003, [1, 2, 3]), (2 ** 53 - 111, [1, 2, 3]), (2 ** 61 - 1, [2, 3]), (2 ** 63 + 29, [2, 3]),
                     (2 ** 64 - 59, [3])]:
        for start, sign in [(1, -1), (p - 1, 1)]:
            for path in paths:
                _, facs, _ = fs.hostemu_iterate(path, p, 500, 502, start)
                assert facs[0] == (p, sign * 500), (p, path)


def test_residue_step_at_error_bound_edges(fs):
    """Single steps at the extremes the error analysis allows: FP64 with
    p just below 2^53, |s| just below 0.8p and m just below 2^50; hybrid with
    p at 2^64 - 2^46, |s| up to min(0.51p, 2^63), m up to min(2^31, 2p) - 1.  Results are exact and the raw
    residues stay inside the stated bounds (|s'| < 0.8p, |s'| <= (1/2 + 2^-19)p)."""
    rng = random.Random(11)
    fp_ps = [2 ** 53 - 111, 2 ** 53 - 1, 2 ** 52 + 1, 2 ** 40 + 15, 1000003, 127]
    for p in fp_ps:
        smax = int(0.8 * p) - 2
        ms = [2 ** 50 - 1, 2 ** 50 - 3, 2 ** 49 + 1, 2 ** 31 - 1, 3] + [rng.randrange(1, 2 ** 50) for _ in range(200)]
        ss = [smax, -smax, 1, -1, 0] + [rng.randrange(-smax, smax + 1) for _ in range(200)]
        for s, m in zip(ss * 2, ms + ms[::-1]):
            raw, res = fs.hostemu_mulmod(fs.PATH_FP64, p, s, m)
            assert res == (s * m) % p, (p, s, m)
            assert raw % p == res and abs(raw) < 0.8 * p, (p, s, m, raw)
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_mulmod(fs.PATH_FP64, 2 ** 53 - 111, 0, 2 ** 50)
    hy_ps = [2 ** 64 - 2 ** 46 - 3, 2 ** 63 + 29, 2 ** 62 + 135, 2 ** 53 + 5, 1000003, 127]
    for p in hy_ps:
        smax = min(int(0.51 * p) - 2 ** 12, 2 ** 63 - 1) if p > 2 ** 40 else int(0.51 * p)
        mtop = min(2 ** 31, 2 * p)  # the step needs m < 2p (iterate: n + 1 <= p while s != 0)
        ms = [mtop - 1, mtop // 2, 3] + [rng.randrange(1, mtop) for _ in range(200)]
        ss = [smax, -smax, 1, -1] + [rng.randrange(-smax, smax + 1) for _ in range(200)]
        for s, m in zip(ss, ms):
            raw, res = fs.hostemu_mulmod(fs.PATH_HYBRID, p, s, m)
            assert res == (s * m) % p, (p, s, m)
            assert raw % p == res and abs(raw) <= (0.5 + 2 ** -19) * p, (p, s, m, raw)


def test_fp64_square_at_error_bound_edges(fs):
    """The FP64 setup's square (fs_fp_mulmod with the residue as its own
    multiplier, used while p < 2^50): |s| up to 0.8p at p just below 2^50 keeps
    the multiplier below 2^50, the result exact and |s'| < 0.74p."""
    rng = random.Random(12)
    for p in [2 ** 50 - 1, 2 ** 50 - 27, 2 ** 49 + 9, 2 ** 40 + 15, 1000003]:
        smax = int(0.8 * p) - 2
        ss = [smax, -smax, smax - 1, 1, -1, 2] + [rng.randrange(-smax, smax + 1) for _ in range(300)]
        for s in ss:
            raw, res = fs.hostemu_mulmod(fs.PATH_FP64, p, s, abs(s))
            assert res == (s * abs(s)) % p, (p, s)
            assert raw % p == res and abs(raw) < 0.74 * p, (p, s, raw)


def test_fp64_exact_step_at_error_bound_edges(fs):
    """The FP64 step without the product's error term (used when p * n1 <=
    2^53; the iterate loops keep |s| <= 0.51p, so |s*m| <= 0.51 * 2^53): exact,
    |s'| <= 0.51p, at p * m = 2^53 and |s| = 0.51p."""
    rng = random.Random(31)
    for p in [2 ** 26 + 15, 2 ** 32 - 5, 2 ** 40 + 15, 2 ** 45 + 59, 2 ** 51 - 129, 2 ** 52 + 1, 1000003, 127, 7]:
        mtop = min(2 ** 53 // p, 2 ** 32 - 1)
        smax = int(0.51 * p)
        while smax > 0.51 * p:
            smax -= 1
        ms = [mtop, mtop - 1 if mtop > 1 else 1, 1] + [rng.randrange(1, mtop + 1) for _ in range(100)]
        ss = [smax, -smax, 1, -1, 0] + [rng.randrange(-smax, smax + 1) for _ in range(100)]
        for s in ss:
            for m in ms[:3] + [ms[3 + (s % 100)]]:
                raw, res = fs.hostemu_fp_exact(p, s, m)
                assert res == (s * m) % p, (p, s, m)
                assert raw % p == res and abs(raw) <= 0.51 * p, (p, s, m, raw)
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_fp_exact(2 ** 40 + 15, 5, 2 ** 53 // (2 ** 40 + 15) + 1)   # p * m above 2^53


def test_fp64_pair_step_at_error_bound_edges(fs):
    """The FP64 paired step (one product by a*(a+1), used when
    64 n1 <= 31 p and n1 <= 2^22): exact, and |s'| <= 0.503p, at the
    largest pair the bounds allow (a + 1 = min(2^22, 31 p / 64), where
    a*(a+1) may exceed p), |s| up to 0.51p, moduli up to 2^53; the
    exact-product form exactly when p (a+1)^2 <= 2^53."""
    rng = random.Random(41)
    ps = [2 ** 53 - 111, 2 ** 53 - 1, 2 ** 46 + 15, 2 ** 46 - 1, 2 ** 40 + 15, 2 ** 33 + 17, 2 ** 26 + 15, 4 * 9 + 1,
          4 * 2 ** 44 + 1, 1000003, 10 ** 12 + 39]
    ps += [rng.randrange(2 ** 20, 2 ** 53) | 1 for _ in range(30)]
    for p in ps:
        atop = min(2 ** 22, 31 * p // 64) - 1   # largest a with 64 (a+1) <= 31 p, a+1 <= 2^22
        assert 64 * (atop + 1) <= 31 * p
        smax = int(0.51 * p)
        while smax > 0.51 * p:
            smax -= 1
        As = [atop, max(1, atop - 1), 1] + [rng.randrange(1, atop + 1) for _ in range(40)]
        ss = [smax, -smax, 1, -1, 0] + [rng.randrange(-smax, smax + 1) for _ in range(40)]
        for s in ss:
            for a in As[:3] + [As[3 + (s % 40)]]:
                raw, res, exact = fs.hostemu_fp_pair(p, s, a)
                M = a * (a + 1)
                assert res == (s * M) % p, (p, s, a)
                assert raw % p == res and abs(raw) <= 0.503 * p, (p, s, a, raw)
                assert exact == (p * (a + 1) ** 2 <= 2 ** 53), (p, a)
    p = 1000003
    fs.hostemu_fp_pair(p, 5, 31 * p // 64 - 1)               # 64 (a+1) <= 31 p: the largest pair
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_fp_pair(p, 5, 31 * p // 64)               # 64 (a+1) > 31 p: no paired step
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_fp_pair(2 ** 52 + 1, 5, 2 ** 22)         # a + 1 above 2^22


def oracle_iterate(p, n0, n1, start, cap=64):
    """The oracle's iterate from a canonical residue: (end residue, factors)."""
    O = ol.oracle()
    M = ol.mont(p)
    s6 = ol.u64(O.ora_mmul(start, M.r2, p, M.q))
    s7 = ol.u64(O.ora_mmul(n0 % p, M.r2, p, M.q))
    buf = (ol.OFactor * cap)()
    nf = O.ora_iterate(ctypes.byref(M), n0, n1, ctypes.byref(s6), ctypes.byref(s7), buf, cap)
    return O.ora_mmul(s6.value, 1, p, M.q), [(buf[i].p, buf[i].nc) for i in range(min(nf, cap))]


def test_fp64_pair_loop_matches_oracle(fs):
    """The FP64 iterate with paired steps against the oracle: moduli on both
    sides of 64 n1 = 31 p (paired loop just inside, single-step loop just
    outside), with and without the exact product, with a*(a+1) below and far
    above p, from random start residues and from residues that put +-1 at
    every position of an 8-step block.  The host execution also returns an
    error if one paired step ever differs from two single steps, leaves
    |s| > 0.51p, or if a block holds a +-1 residue and its filter (one compare
    per pair, |v'| <= the block's last a+1) does not fire."""
    rng = random.Random(43)
    cases = []
    for n1 in (1040, 5003, 20000, 300000, 2 ** 22):
        edge = -(-64 * n1 // 31)                                # smallest p with 64 n1 <= 31 p
        cases += [(edge | 1, n1), ((edge | 1) + 2, n1), (edge - 1 - edge % 2, n1)]   # pair, pair, single
    cases += [(2 ** 26 + 15, 1030), (2 ** 33 + 17, 20001), (10 ** 12 + 39, 200000), (2 ** 53 - 111, 2 ** 22),
              (10 ** 8 + 7, 9000), (2 ** 32 + 15, 500000), (2 ** 36 + 31, 2 ** 22)]   # a*(a+1) > p
    for p, n1 in cases:
        for n0 in (max(1, n1 - 8 * 4 - 3), max(1, n1 - 17)):
            starts = [rng.randrange(p) for _ in range(2)]
            for pos in range(8):
                prod = 1
                for j in range(1, pos + 1):
                    prod = prod * (n0 + j) % p
                if math.gcd(prod, p) == 1:
                    starts.append(pow(prod, -1, p) * (1 if pos % 2 else -1) % p)
            for start in starts:
                r, facs, nf = fs.hostemu_iterate(fs.PATH_FP64, p, n0, n1, start)
                want_r, want_f = oracle_iterate(p, n0, n1, start)
                assert (r, facs[:nf]) == (want_r, want_f), (p, n0, n1, start)


HYBRID_PMAX = 2 ** 64 - 2 ** 46
PAIR_PLIM = 2 ** 64 - 2 ** 58   # the hybrid paired step takes p below this


def test_hybrid_pair_step_at_error_bound_edges(fs):
    """The hybrid paired step (one product by a*(a+1), quotient from the
    estimate of s/p, used when n1 <= 2^22 and 2^45 <= p < 2^64 - 2^58): exact
    and |s'| <= 0.503p at the largest pair (a + 1 = 2^22), the smallest and
    largest moduli, |s| up to 0.51p and low words up to 2^32 - 1; a residue
    whose product with a is +-1 gives exactly +-(a+1), which the loop's
    compare relies on."""
    rng = random.Random(61)
    ps = [2 ** 45 + 59, 2 ** 45 + 1, 2 ** 53 + 5, 2 ** 62 + 135, 2 ** 63 - 25, 2 ** 63 + 29, PAIR_PLIM - 1,
          PAIR_PLIM - 3]
    ps += [rng.randrange(2 ** 45, PAIR_PLIM) | 1 for _ in range(40)]
    for p in ps:
        smax = min(int(0.51 * p), 2 ** 63 - 1) - 2 ** 12
        As = [2 ** 22 - 1, 2 ** 22 - 2, 1, 2] + [rng.randrange(1, 2 ** 22) for _ in range(30)]
        s0max = (((smax >> 32) - 1) << 32) | 0xFFFFFFFF
        ss = [smax, -smax, 1, -1, 0, s0max, -s0max] + [rng.randrange(-smax, smax + 1) for _ in range(30)]
        for s in ss:
            for a in As[:4] + [As[4 + (s % 30)]]:
                raw, res = fs.hostemu_sh_pair(p, s, a)
                assert res == (s * a * (a + 1)) % p, (p, s, a)
                assert raw % p == res and abs(raw) <= 0.503 * p, (p, s, a, raw)
        for a in As[:8]:
            if math.gcd(a, p) != 1:
                continue
            inv = pow(a, -1, p)
            for sign in (1, -1):
                s = sign * inv % p
                s = s - p if s > p // 2 else s
                raw, _ = fs.hostemu_sh_pair(p, s, a)
                assert raw == sign * (a + 1), (p, a, sign, raw)
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_sh_pair(2 ** 45 - 1, 5, 1000)          # modulus below 2^45
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_sh_pair(2 ** 60 + 33, 5, 2 ** 22)      # a + 1 above 2^22
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_sh_pair(PAIR_PLIM + 1, 5, 1000)        # |s'| could reach 2^63


def test_hybrid_pair_loop_matches_oracle(fs):
    """The hybrid iterate with paired steps against the oracle, from random
    residues and from residues that put +-1 at every position of an 8-step
    block, at moduli from 2^45 to the top of the hybrid range (single steps
    from 2^64 - 2^58) and n1 up to 2^22 (where the paired loop stops).  The host execution also returns an
    error if a pair differs from two single steps, a residue leaves 0.51p, or
    a block's filter misses a +-1."""
    rng = random.Random(67)
    cases = [(2 ** 45 + 59, 1040), (2 ** 53 + 5, 20001), (2 ** 58 + 27, 500000), (2 ** 62 + 135, 300000),
             (2 ** 63 + 29, 2 ** 22), (PAIR_PLIM - 1, 2 ** 22), (PAIR_PLIM + 1, 77777), (HYBRID_PMAX - 3, 7777),
             (2 ** 50 + 55, 2 ** 22 - 5)]
    cases += [(rng.randrange(2 ** 45, HYBRID_PMAX) | 1, rng.randrange(40, 2 ** 22)) for _ in range(10)]
    for p, n1 in cases:
        for n0 in (max(1, n1 - 8 * 4 - 3), max(1, n1 - 17)):
            starts = [rng.randrange(p) for _ in range(2)]
            for pos in range(8):
                prod = 1
                for j in range(1, pos + 1):
                    prod = prod * (n0 + j) % p
                if math.gcd(prod, p) == 1:
                    for sign in (1, -1):
                        starts.append(sign * pow(prod, -1, p) % p)
            for start in starts:
                r, facs, nf = fs.hostemu_iterate(fs.PATH_HYBRID, p, n0, n1, start)
                want_r, want_f = oracle_iterate(p, n0, n1, start)
                assert (r, facs[:nf]) == (want_r, want_f), (p, n0, n1, start)


def hi_mmax(p):
    """Largest multiplier the hybrid high-word step takes for p: p >> 34 below
    2^63 (|s'| <= 3/4 p + 2^-21 p), p >> 39 up to 2^64 - 2^58 (|s'| <= 0.51p)."""
    if p >= 2 ** 64 - 2 ** 58:
        return 0
    return p >> 34 if p < 2 ** 63 else p >> 39


def test_hybrid_hi_step_at_error_bound_edges(fs):
    """The hybrid step with the quotient from the high word of s only: exact,
    and |s'| stays within its bound (3/4 p + 2^-21 p below 2^63, 0.51p above),
    at the largest multiplier each regime allows, |s| up to that bound, and
    the largest dropped low word s0 = 2^32 - 1."""
    rng = random.Random(23)
    top = 2 ** 64 - 2 ** 58          # first modulus the high-word step does not take
    ps = [top - 1, top - 3, 2 ** 63 + 29, 2 ** 63 - 25, 2 ** 62 + 135, 2 ** 58 + 27, 2 ** 53 + 5, 2 ** 50 + 55,
          (2 ** 19 << 34) + 1, (2 ** 19 << 39) + 1, (524300 << 34) + 7]
    for p in ps:
        bound = 0.75 + 2 ** -21 if p < 2 ** 63 else 0.51
        mtop = min(hi_mmax(p), 2 ** 31 - 1)
        smax = min(int(bound * p) - 2 ** 12, 2 ** 63 - 1)
        ms = [mtop, mtop - 1, max(1, mtop // 2), 1] + [rng.randrange(1, mtop + 1) for _ in range(150)]
        s0max = (((smax >> 32) - 1) << 32) | 0xFFFFFFFF    # largest dropped low word, |s| <= smax
        ss = [smax, -smax, 1, -1, 0, s0max, -s0max, -(((smax >> 32) - 1) << 32)] + \
             [rng.randrange(-smax, smax + 1) for _ in range(150)]
        for s in ss:
            for m in ms[:4] + [ms[4 + (s % 150)]]:
                raw, res = fs.hostemu_mulmod_hi(p, s, m)
                assert res == (s * m) % p, (p, s, m)
                assert raw % p == res and abs(raw) <= bound * p, (p, s, m, raw)
                if abs(s) == 1:
                    assert raw == s * m, (p, s, m, raw)   # +-1 times m stays +-m (paired test)
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_mulmod_hi(2 ** 58 + 27, 5, ((2 ** 58 + 27) >> 34) + 1)   # multiplier above p / 2^34
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_mulmod_hi(2 ** 63 + 29, 5, ((2 ** 63 + 29) >> 39) + 1)   # above 2^63: p / 2^39
    with pytest.raises(fs.FactSieveError):
        fs.hostemu_mulmod_hi(top + 1, 5, 3)                                # |s'| could reach 2^63


def test_hybrid_hi_filter_sound(fs):
    """With the high-word quotient the +-1 filter is |y| <= 1.5 * 2^32/p; a
    residue of +-1 must never fail it (the host execution returns an error if
    it does, or if a residue leaves the loop's bound).  n-ranges chosen so that
    the high-word loop is selected."""
    rng = random.Random(29)
    ps = [2 ** 64 - 2 ** 58 - 1, 2 ** 63 + 29, 2 ** 63 - 25, 2 ** 62 + 135, 2 ** 58 + 27, 2 ** 53 + 5, 2 ** 45 + 59]
    ps += [rng.randrange(2 ** 45, 2 ** 64 - 2 ** 58) | 1 for _ in range(200)]
    for p in ps:
        top = min(hi_mmax(p), 2 ** 31 - 1)
        for start in (1, p - 1):
            for n0 in (rng.randrange(1, max(2, top - 4)), top - 3):
                if n0 < 1 or n0 + 3 > top:
                    continue
                _, facs, nf = fs.hostemu_iterate(fs.PATH_HYBRID, p, n0, n0 + 3, start)
                assert nf >= 1 and facs[0] == (p, (-1 if start == 1 else 1) * n0), p


@pytest.mark.parametrize("path", [1, 2])
def test_block_filter_catches_every_position(fs, path):
    """The iterate loops test one residue of each pair (|v[k+1]| <= m[k], or
    its quotient estimate) for +-1 at k and k+1.  Start residues are chosen so
    that +-1 falls at each position 0..7 of an 8-step block, over moduli that
    select each loop variant (FP64 with and without the exact product; hybrid
    with and without the high-word quotient); the host execution returns an
    error if a block's filter ever misses a +-1 residue."""
    rng = random.Random(37 + path)
    if path == 1:
        ps = [1000003, 2 ** 26 + 15, 2 ** 40 + 15, 2 ** 50 + 55, 2 ** 53 - 111]
        ps += [rng.randrange(10 ** 6, 2 ** 53) | 1 for _ in range(40)]
    else:
        ps = [2 ** 53 + 5, 2 ** 58 + 27, 2 ** 62 + 135, 2 ** 63 + 29, 2 ** 64 - 2 ** 58 - 1, 2 ** 64 - 2 ** 46 - 3]
        ps += [rng.randrange(2 ** 53, 2 ** 64 - 2 ** 46) | 1 for _ in range(40)]
    # path 2: n0 = 700000 and 2^24 + 100 keep the full quotient below
    # p = 700000 * 2^34 and 2^58 (the high-word loop needs n1 <= p >> 34)
    n0s = (1000, 700000, 2 ** 24 + 100) if path == 2 else (1000,)
    for p in ps:
        for n0 in n0s + (rng.randrange(2000, 400000),):
            n1 = n0 + 8 * 3 + rng.randrange(0, 8)
            if n1 >= p:
                continue
            for k in range(8):
                for sign in (1, -1):
                    blk = 8 * rng.randrange(0, 3)     # which block of the range
                    prod = 1
                    for j in range(1, blk + k + 1):
                        prod = prod * (n0 + j) % p
                    if math.gcd(prod, p) != 1:
                        continue
                    start = sign * pow(prod, -1, p) % p   # residue at n0 such that n0+blk+k gives +-1
                    _, facs, nf = fs.hostemu_iterate(path, p, n0, n1, start)
                    want = (p, -(n0 + blk + k) if sign == 1 else n0 + blk + k)
                    assert want in facs[:nf], (p, n0, k, sign)


def test_hybrid_quotient_filter_sound(fs):
    """The hybrid iterate tests s == +-1 only where the quotient estimate
    passes |y| <= 1.5/p; the host execution returns an error if a residue of
    +-1 ever fails that filter.  Start residues +-1 (and +-1 times a few
    multipliers) over odd moduli from 3 to the top of the hybrid range, with
    every multiplier below 2p as in the sieve (the residue is 0 from n = p)."""
    rng = random.Random(5)
    ps = [3, 5, 7, 127, 65537, 2 ** 31 - 1, 2 ** 32 + 15, 2 ** 52 - 1, 2 ** 52 + 1, 2 ** 63 - 25, 2 ** 63 + 29,
          2 ** 64 - 2 ** 46 - 1, 2 ** 64 - 2 ** 46 - 3]
    ps += [rng.randrange(3, 2 ** rng.randrange(3, 65)) | 1 for _ in range(300)]
    for p in ps:
        if p > 2 ** 64 - 2 ** 46:
            continue
        for start in (1, p - 1):
            n0 = rng.randrange(1, min(2 ** 31 - 10, 2 * p - 4))
            _, facs, nf = fs.hostemu_iterate(fs.PATH_HYBRID, p, n0, n0 + 3, start)
            assert nf >= 1 and facs[0] == (p, (-1 if start == 1 else 1) * n0), p


@pytest.mark.parametrize("lo,hi", [(127, 100000), (10 ** 8, 10 ** 8 + 200000), (2 ** 52 - 100000, 2 ** 52 + 100000),
                                   (2 ** 62, 2 ** 62 + 100000), (2 ** 64 - 300000, 2 ** 64 - 1), (1000, 1001),
                                   (127, 128), (2 ** 64 - 2, 2 ** 64 - 1), (2 ** 34 - 100000, 2 ** 34 + 100000),
                                   (2 ** 47 - 100000, 2 ** 47 + 100000), (2 ** 63 - 100000, 2 ** 63 + 100000),
                                   (2 ** 50 - 200000, 2 ** 50 + 100000),
                                   (10 ** 5, 10 ** 6)])
def test_candidate_masks_match_oracle(fs, lo, hi):
    assert fs.hostemu_candidates(lo, hi) == ol.candidates(lo, hi)


def test_cli_rejects_bad_arguments(fs):
    exe = fs.EXE_PATH
    r = subprocess.run([exe, "-p", "12x"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "invalid argument -p 12x" in r.stdout
    r = subprocess.run([exe, "-n", "99"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "out of range argument -n 99" in r.stdout
    r = subprocess.run([exe, "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "-N #" in r.stdout and "--verify-on" in r.stdout
    r = subprocess.run([exe, "--verify-on", "fpga", "-n", "1000"], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "invalid argument --verify-on fpga" in r.stdout


def test_default_options(fs):
    o = fs.default_options()
    assert (o.path, o.verify_threads, o.verify_device, o.checkpoint_secs) == (fs.PATH_AUTO, 2, fs.VERIFY_AUTO, 60)
    assert (o.write_files, o.resume, o.quiet, o.seg_candidates, o.n_chunk) == (0, 0, 1, 0, 0)


# ---------------------------------------------------------------- checkpoints
RANGE = (10 ** 8, 10 ** 8 + 2 * 10 ** 6, 1000, 10000)


def _state(fs, p, checksum=0x1234, primecount=77, factorcount=3):
    pmin, pmax, nmin, nmax = RANGE
    return fs.WorkStatus(pmin=pmin, pmax=pmax, p=p, checksum=checksum, primecount=primecount,
                         factorcount=factorcount, last_trickle=1700000000, nmin=nmin, nmax=nmax)


def test_checkpoint_record_layout_and_sum(fs, tmp_path):
    """72-byte little-endian workStatus, state_sum = wrapping sum of the other fields."""
    assert ctypes.sizeof(fs.WorkStatus) == 72
    st = fs.write_checkpoint(tmp_path, _state(fs, 10 ** 8 + 5000), 0)
    raw = open(tmp_path / "stateA.ckp", "rb").read()
    assert len(raw) == 72
    f = struct.unpack("<8Q2I", raw)
    assert f[7] == sum(f[:7] + f[8:]) % 2 ** 64 == st.state_sum
    assert f[2] == 10 ** 8 + 5000 and f[8:] == (1000, 10000)


def test_checkpoint_selection_and_rejection(fs, tmp_path, capfd):
    """read_state rules (cl_sieve.cpp:132-222): larger p of two good files;
    a truncated, range-mismatched or checksum-broken file is skipped with the
    reference's message; none good -> fresh start."""
    d = tmp_path
    assert fs.read_checkpoint(d, *RANGE) is None  # no files
    fs.write_checkpoint(d, _state(fs, 10 ** 8 + 4000), 0)
    fs.write_checkpoint(d, _state(fs, 10 ** 8 + 9000), 1)
    st, which = fs.read_checkpoint(d, *RANGE)
    assert (st.p, which) == (10 ** 8 + 9000, 1)
    # checksum error in B -> A
    raw = bytearray(open(d / "stateB.ckp", "rb").read())
    raw[24] ^= 1  # checksum field
    open(d / "stateB.ckp", "wb").write(raw)
    st, which = fs.read_checkpoint(d, *RANGE)
    assert (st.p, which) == (10 ** 8 + 4000, 0)
    assert "Checksum error in stateB.ckp" in capfd.readouterr().err
    # A for another range -> invalid; B broken -> none
    other = _state(fs, 10 ** 8 + 4000)
    other.nmax = 20000
    fs.write_checkpoint(d, other, 0)
    assert fs.read_checkpoint(d, *RANGE) is None
    assert "Invalid checkpoint file stateA.ckp" in capfd.readouterr().err
    # truncated B -> cannot parse; A good again -> A
    fs.write_checkpoint(d, _state(fs, 10 ** 8 + 6000), 0)
    open(d / "stateB.ckp", "wb").write(b"\0" * 40)
    st, which = fs.read_checkpoint(d, *RANGE)
    assert (st.p, which, st.primecount, st.factorcount, st.checksum) == (10 ** 8 + 6000, 0, 77, 3, 0x1234)
    assert "Cannot parse stateB.ckp" in capfd.readouterr().err


def test_checkpoint_write_error_reported(fs, tmp_path):
    with pytest.raises(fs.FactSieveError):
        fs.write_checkpoint(tmp_path / "no" / "such" / "dir", _state(fs, 10 ** 8), 0)


def test_trickle_up_text_is_the_reference_message(fs):
    """The daily trickle-up record, character for character as
    cl_sieve.cpp:59-64 formats it (%lf fields)."""
    assert fs.trickle_up_text(0.25, 1234.5, 86400.125) == (
        "<trickle_up>\n"
        "   <progress>0.250000</progress>\n"
        "   <cputime>1234.500000</cputime>\n"
        "   <runtime>86400.125000</runtime>\n"
        "</trickle_up>\n")


def test_read_init_data(fs, tmp_path):
    """The stand-alone device choice from init_data.xml (README.md:42-56):
    both documented formats, a missing file, and files without a valid
    <gpu_device_num>."""
    assert fs.read_init_data(tmp_path) is None
    for gpu_type, num in (("NVIDIA", 0), ("ATI", 2), ("ATI", 63)):
        (tmp_path / "init_data.xml").write_text(
            f"<app_init_data>\n<gpu_type>{gpu_type}</gpu_type>\n<gpu_device_num>{num}</gpu_device_num>\n"
            "</app_init_data>\n")
        assert fs.read_init_data(tmp_path) == (num, gpu_type)
    (tmp_path / "init_data.xml").write_text("<app_init_data><gpu_device_num> 5 </gpu_device_num></app_init_data>")
    assert fs.read_init_data(tmp_path) == (5, "")
    for bad in ("<app_init_data></app_init_data>", "<gpu_device_num>x</gpu_device_num>",
                "<gpu_device_num>64</gpu_device_num>", "<gpu_device_num>-1</gpu_device_num>",
                "<gpu_device_num>1"):
        (tmp_path / "init_data.xml").write_text(bad)
        with pytest.raises(fs.FactSieveError):
            fs.read_init_data(tmp_path)


def test_cli_takes_the_device_from_init_data(fs, tmp_path):
    """Without -d the executable names the device it takes from
    init_data.xml, or device 0 when there is no file; with -d it reads no
    file.  (Only the device choice is checked: the run itself needs a GPU.)"""
    import subprocess
    args = [fs.EXE_PATH, "-n", "1000", "-N", "1100", "-p", "100000000", "-P", "100010000"]

    def first_lines():
        r = subprocess.run(args, cwd=tmp_path, capture_output=True, text=True, timeout=120)
        return r.stdout

    assert "init_data.xml not found, using device 0." in first_lines()
    (tmp_path / "init_data.xml").write_text(
        "<app_init_data>\n<gpu_type>ATI</gpu_type>\n<gpu_device_num>0</gpu_device_num>\n</app_init_data>\n")
    assert "init_data.xml: using GPU device 0 (ATI)." in first_lines()
    args += ["-d", "0"]
    assert "init_data.xml" not in first_lines()
