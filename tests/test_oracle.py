"""Pins the CPU oracle: reference self-test vectors, the reference's own
verifyprime.c / putil.c (oracle/_ref), closed forms and brute force.  CPU only."""
import ctypes
import json
import math
import os
import random

import pytest

import oracle_lib as ol

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def small_primes(n):
    sieve = bytearray([1]) * (n + 1)
    sieve[0:2] = b"\x00\x00"
    for i in range(2, int(n ** 0.5) + 1):
        if sieve[i]:
            sieve[i * i::i] = bytearray(len(sieve[i * i::i]))
    return [i for i in range(n + 1) if sieve[i]]


def spsp2_py(n):
    d, t = n - 1, 0
    while d % 2 == 0:
        d //= 2
        t += 1
    a = pow(2, d, n)
    if a in (1, n - 1):
        return True
    for _ in range(t - 1):
        a = a * a % n
        if a == n - 1:
            return True
    return False


SMALL = [q for q in small_primes(113) if q > 2]


# --------------------------------------------------------------- arithmetic

def test_montgomery_against_bigints():
    O = ol.oracle()
    rng = random.Random(7)
    for _ in range(3000):
        bits = rng.choice([20, 40, 52, 62, 63, 64])
        p = rng.getrandbits(bits) | 1 | (1 << (bits - 1))
        q = O.ora_invert(p)
        assert (p * q) % 2 ** 64 == 1
        a, b = rng.randrange(p), rng.randrange(p)
        assert O.ora_mmul(a, b, p, q) == a * b * pow(2, -64, p) % p
        assert O.ora_addmod(a, b, p) == (a + b) % p
        m = ol.mont(p)
        assert m.one == 2 ** 64 % p and m.nmo == p - m.one and m.r2 == 2 ** 128 % p


# --------------------------------------------------------------- candidates

def test_spsp2_matches_bigint_test():
    O = ol.oracle()
    rng = random.Random(3)
    for _ in range(4000):
        n = rng.getrandbits(rng.choice([17, 33, 50, 64])) | 1
        if n < 5:
            continue
        assert bool(O.ora_spsp2(n)) == spsp2_py(n), n


def test_known_base2_strong_pseudoprimes():
    # OEIS A001262, all terms below 1e6
    a001262 = [2047, 3277, 4033, 4681, 8321, 15841, 29341, 42799, 49141, 52633, 65281, 74665, 80581, 85489,
               88357, 90751, 104653, 130561, 196093, 220729, 233017, 252601, 253241, 256999, 271951, 280601,
               314821, 357761, 390937, 458989, 476971, 486737, 489997, 514447, 580337, 635401, 647089, 741751,
               800605, 818201, 838861, 873181, 877099, 916327, 976873, 983401]
    O = ol.oracle()
    for n in a001262:
        assert O.ora_spsp2(n)
    survivors = [n for n in a001262 if all(n % q for q in SMALL)]
    assert survivors == load("spsp2_survivors_1e6.json")["values"]
    found = [n for n in range(115, 10 ** 6, 2) if O.ora_is_candidate(n) and not O.ora_is_prime(n)]
    assert found == survivors


def test_candidates_bruteforce_small_range():
    pr = set(small_primes(200000))
    want = [n for n in range(127, 200000) if n in pr or (n % 2 and all(n % q for q in SMALL) and spsp2_py(n))]
    assert ol.candidates(127, 200000) == want


def test_fermat_numbers_d1_edge():
    """N = 2^t + 1 (odd part of N-1 is 1): the reference's masked 64-bit shift
    (getsegprimes.cl:81-84) evaluates 2^(2^64+1) instead of 2^1; for every such N
    that survives the 113-sieve both give the same verdict."""
    O = ol.oracle()
    for t in range(7, 64):
        n = 2 ** t + 1
        if not all(n % q for q in SMALL):
            continue
        a = pow(2, 2 ** 64 + 1, n)
        masked = a in (1, n - 1) or any(pow(a, 2 ** s, n) == n - 1 for s in range(1, t))
        assert masked == spsp2_py(n) == bool(O.ora_spsp2(n)), t


@pytest.mark.parametrize("case", load("kats.json")["cases"])
def test_kat_candidate_counts(case):
    pmin, pmax = case[0], case[1]
    O = ol.oracle()
    assert O.ora_candidates(pmin, pmax, None, 0) == case[5]


def test_c1_candidate_count():
    assert ol.oracle().ora_candidates(100000000, 102000000, None, 0) == 108529


def test_wheel_offset_definition():
    O = ol.oracle()
    res = [7, 11, 13, 17, 19, 23, 29, 1]
    for start in list(range(120, 400)) + [10 ** 12 + k for k in range(40)]:
        s, i = ol.u64(start), ol.i32()
        O.ora_find_wheel_offset(ctypes.byref(s), ctypes.byref(i))
        n = next(x for x in range(start, start + 30) if math.gcd(x, 30) == 1)
        assert s.value == n and res[i.value] == n % 30


# --------------------------------------------------------------- power table / setup

def test_power_table_closed_form():
    O = ol.oracle()
    for nmin in [1, 2, 127, 1000, 5000]:
        prod = 1
        for q in small_primes(nmin):
            e, cb = ol.u32(), ol.u32()
            O.ora_get_power(q, nmin, ctypes.byref(e), ctypes.byref(cb))
            legendre = sum(nmin // q ** k for k in range(1, 40) if q ** k <= nmin)
            assert e.value == legendre
            if e.value > 1:
                assert cb.value == 1 << (e.value.bit_length() - 2)
            prod *= q ** e.value
        assert prod == math.factorial(nmin)


def test_setup_residue_equals_factorial():
    O = ol.oracle()
    rng = random.Random(11)
    for nmin in [127, 1000, 4000]:
        B = ctypes.POINTER(ctypes.c_uint64)()
        E = ctypes.POINTER(ctypes.c_uint32)()
        C = ctypes.POINTER(ctypes.c_uint32)()
        O.ora_power_table.restype = ctypes.c_uint64
        m = O.ora_power_table(nmin, ctypes.byref(B), ctypes.byref(E), ctypes.byref(C))
        O.ora_setup_residue.restype = ctypes.c_uint64
        for _ in range(20):
            p = rng.choice(ol.candidates(10 ** 9, 10 ** 9 + 2000) + ol.candidates(2 ** 63, 2 ** 63 + 3000))
            M = ol.mont(p)
            s6 = O.ora_setup_residue(ctypes.byref(M), B, E, C, ctypes.c_uint64(m))
            assert s6 == math.factorial(nmin) % p * 2 ** 64 % p
            assert O.ora_factorial_mod(nmin, p) == math.factorial(nmin) % p


# --------------------------------------------------------------- iterate

def brute_factors(p, nmin, nmax):
    out, f = [], math.factorial(nmin) % p
    for n in range(nmin, nmax):
        if f == 1:
            out.append((p, -n))
        elif f == p - 1:
            out.append((p, n))
        f = f * (n + 1) % p
    return out, f


def test_iterate_bruteforce_and_wilson():
    O = ol.oracle()
    for p in [p for p in small_primes(6000) if p > 130][:120]:
        M = ol.mont(p)
        nmin, nmax = 127, p + 5
        s6 = ol.u64(math.factorial(nmin) % p * 2 ** 64 % p)
        s7 = ol.u64(nmin * 2 ** 64 % p)
        buf = (ol.OFactor * 256)()
        nf = O.ora_iterate(ctypes.byref(M), nmin, nmax, ctypes.byref(s6), ctypes.byref(s7), buf, 256)
        got = [(buf[i].p, buf[i].nc) for i in range(nf)]
        want, fend = brute_factors(p, nmin, nmax)
        assert got == want
        # Wilson: (p-1)! = -1 and (p-2)! = +1 (mod p)
        assert (p, p - 1) in got and (p, -(p - 2)) in got
        assert s6.value == fend * 2 ** 64 % p and s7.value == nmax * 2 ** 64 % p


# --------------------------------------------------------------- whole sieve vs reference vectors

def test_kat3_full():
    c = load("kats.json")["cases"][2]
    r = ol.sieve(c[0], c[1], c[2], c[3], threads=8, cap=60000)
    assert (r["factorcount"], r["primecount"], f"{r['checksum']:016X}") == (c[4], c[5], c[6])


def test_c1_golden():
    g = load("c1.json")
    r = ol.sieve(*g["range"], threads=8)
    assert r["factorcount"] == g["factorcount"] == 25
    assert r["primecount"] == g["primecount"] == 108529
    assert f"{r['checksum']:016X}" == g["checksum"] == "000009F5FD6D6CFD"
    assert [list(f) for f in r["factors"]] == g["factors"]


def test_checksum_definition_on_tiny_case():
    """checksum = sum over candidates of mont(nmax!) + mont(nmax), plus n + c per kept factor."""
    pmin, pmax, nmin, nmax = 127, 3000, 127, 5000
    r = ol.sieve(pmin, pmax, nmin, nmax, residues=True)
    total = 0
    for p, s6, s7 in r["cands"]:
        assert s6 == math.factorial(nmax) % p * 2 ** 64 % p and s7 == nmax % p * 2 ** 64 % p
        total += s6 + s7
    for p, nc in r["factors"]:
        total += abs(nc) + (-1 if nc < 0 else 1)
    assert total % 2 ** 64 == r["checksum"]


def test_finalize_rejects_planted_bad_factor():
    O = ol.oracle()
    f = (ol.OFactor * 2)()
    f[0].p, f[0].nc = 131, -(131 - 2)   # Wilson: true factor of 129!-1
    f[1].p, f[1].nc = 137, 100          # planted: 137 does not divide 100!+1
    ck, bad = ol.u64(0), ol.i64()
    assert O.ora_finalize(f, 2, ctypes.byref(ck), ctypes.byref(bad)) == -1 and bad.value == 1


# --------------------------------------------------------------- against the reference's own C

REF = ol.ref()
needs_ref = pytest.mark.skipif(REF is None, reason="oracle/_ref not built (reference checkout absent)")


@needs_ref
def test_is_prime_matches_reference_isPrime():
    O = ol.oracle()
    rng = random.Random(5)
    vals = [rng.getrandbits(64) | 1 for _ in range(5000)] + ol.candidates(10 ** 12, 10 ** 12 + 20000)
    vals += load("spsp2_survivors_1e6.json")["values"] + [3215031751, 3825123056546413051, 2 ** 64 - 59]
    for v in vals:
        assert bool(O.ora_is_prime(v)) == REF._Z7isPrimem(v), v


@needs_ref
def test_verify_matches_reference_verify():
    O = ol.oracle()
    g = load("c1.json")
    cases = [tuple(f) for f in g["factors"]]
    rng = random.Random(9)
    cases += [(p, rng.choice([-1, 1]) * rng.randrange(35, 3000)) for p in ol.candidates(10 ** 8, 10 ** 8 + 3000)]
    for p, nc in cases:
        n, c = abs(nc), (-1 if nc < 0 else 1)
        assert bool(O.ora_verify(p, n, c)) == REF._Z6verifymji(p, n, c)


@needs_ref
def test_parse_matches_reference_putil():
    O = ol.oracle()
    strs = ["127", "1e8", "1e12", "100e6", "101e6", "4K", "4M", "3G", "2T", "1P", "1k", "1m", "1g", "3t", "1p",
            "1b10", "3b62", "0x7f", "010", "12x", "1e", "1e99", "99999999999999999999", "5Kx", "18446744073709551615",
            "1b64", "2e18", "abc", "", "1E3", "1B4"]
    for s in strs:
        for lo, hi in [(127, 2 ** 64 - 2), (128, 2 ** 64 - 1), (127, 2 ** 31 - 2), (1, 128)]:
            a, b = ol.u64(0), ol.u64(0)
            ra = O.ora_parse_uint64(ctypes.byref(a), s.encode(), lo, hi)
            rb = REF.parse_uint64(ctypes.byref(b), s.encode(), lo, hi)
            assert (ra, a.value if ra == 0 else None) == (rb, b.value if rb == 0 else None), (s, lo, hi)


def test_kat1_full():
    """Reference self-test case 1 end to end on the oracle (~40 s on 8 cores)."""
    c = load("kats.json")["cases"][0]
    r = ol.sieve(c[0], c[1], c[2], c[3], threads=8, cap=4000)
    assert (r["factorcount"], r["primecount"], f"{r['checksum']:016X}") == (c[4], c[5], c[6])
