"""Writes the golden fixtures under tests/golden/ from the CPU oracle.

    python tests/golden/make_golden.py              # kats, c1, tiny, spsp2 survivors
    python tests/golden/make_golden.py c2 c3 c4_sub # the full-size ranges (tens of minutes)

kats.json is the reference's own self-test table (cl_sieve.cpp:1282-1375).
Every other file is produced by the oracle, whose functions the CPU tests pin
against those KATs, closed forms and brute force (tests/test_oracle.py).
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_lib  # noqa: E402

KATS = [
    # pmin, pmax, nmin, nmax, factors, primecount, checksum (cl_sieve.cpp:1283-1364)
    [100000000, 101000000, 1000000, 2000000, 1071, 54211, "000004F4F744CA97"],
    [1000000000000, 1000010000000, 10000, 2000000, 3, 361727, "05052A2D65F3D735"],
    [127, 100000, 127, 1000000, 42770, 9566, "0000000065C074F0"],
    [2000000000000, 2000010000000, 2000000, 6000000, 3, 352866, "09C9742A908451EB"],
    [1000000000000, 1000001000000, 100000000, 110000000, 3, 36249, "0081056DC98DA092"],
]

# Small ranges with per-candidate residues: every arithmetic regime of the
# kernels (tiny p <= nmax, 2^40, just below 2^52, 2^62, just below 2^63, the
# top of the 64-bit range).
TINY = [
    [127, 3000, 127, 5000],
    [1000003, 1006000, 1000, 4000],
    [1 << 40, (1 << 40) + 6000, 10000, 12000],
    [(1 << 52) - 6000, 1 << 52, 5000, 6500],
    [1 << 62, (1 << 62) + 4000, 200, 1200],
    [(1 << 63) - 4000, 1 << 63, 3000, 3500],
    [(1 << 64) - 5000, (1 << 64) - 1, 300, 1300],
]

C1 = [100000000, 102000000, 1000, 10000]
# Plain single job at production shape and the full-width modulus stress range.
C2 = [1000000000000, 1001000000000, 100000, 200000]
C3 = [1 << 62, (1 << 62) + (1 << 30), 500000, 510000]


def write_range(name, rng, threads=8, jobs_note=None):
    r = oracle_lib.sieve(*rng, threads=threads)
    doc = {"range": rng, "factorcount": r["factorcount"], "primecount": r["primecount"],
           "checksum": f"{r['checksum']:016X}", "device_checksum": f"{r['device_checksum']:016X}",
           "raw": r["raw"], "factors": r["factors"]}
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(doc, f, indent=1)
    print(name, doc["factorcount"], doc["primecount"], doc["checksum"], flush=True)


def write_c4_sub(threads=8):
    """The 64-job mixed batch (seed 20261203, sigma 1.0) with every p-range cut
    to its first 1/64 (at least one integer) and the full n-ranges."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from factsieve_amd import workloads
    jobs = []
    for pmin, pmax, nmin, nmax in workloads.mixed_jobs(64):
        job = [pmin, pmin + max(1, (pmax - pmin) // 64), nmin, nmax]
        r = oracle_lib.sieve(*job, threads=threads)
        jobs.append({"range": job, "factorcount": r["factorcount"], "primecount": r["primecount"],
                     "checksum": f"{r['checksum']:016X}", "factors": r["factors"]})
        print("c4_sub", len(jobs), job, r["primecount"], flush=True)
    with open(os.path.join(HERE, "c4_sub.json"), "w") as f:
        json.dump({"seed": workloads.DEFAULT_SEED, "sigma": 1.0, "p_fraction": "1/64", "jobs": jobs}, f, indent=1)


def main(targets):
    if "c2" in targets:
        write_range("c2", C2)
    if "c3" in targets:
        write_range("c3", C3)
    if "c4_sub" in targets:
        write_c4_sub()
    if "base" not in targets:
        return
    with open(os.path.join(HERE, "kats.json"), "w") as f:
        json.dump({"source": "cl_sieve.cpp:1282-1375 (reference self-test)",
                   "fields": ["pmin", "pmax", "nmin", "nmax", "factors", "primecount", "checksum"],
                   "cases": KATS}, f, indent=1)

    r = oracle_lib.sieve(*C1, threads=8)
    with open(os.path.join(HERE, "c1.json"), "w") as f:
        json.dump({"range": C1, "factorcount": r["factorcount"], "primecount": r["primecount"],
                   "checksum": f"{r['checksum']:016X}", "device_checksum": f"{r['device_checksum']:016X}",
                   "factors": r["factors"]}, f, indent=1)

    cases = []
    for pmin, pmax, nmin, nmax in TINY:
        r = oracle_lib.sieve(pmin, pmax, nmin, nmax, threads=8, residues=True)
        cases.append({"range": [pmin, pmax, nmin, nmax], "factorcount": r["factorcount"],
                      "primecount": r["primecount"], "checksum": f"{r['checksum']:016X}",
                      "device_checksum": f"{r['device_checksum']:016X}", "raw": r["raw"],
                      "factors": r["factors"], "cands": r["cands"]})
    with open(os.path.join(HERE, "tiny.json"), "w") as f:
        json.dump({"cases": cases}, f)

    O = oracle_lib.oracle()
    surv = [n for n in range(115, 10**6, 2) if O.ora_sieve_survivor(n) and O.ora_spsp2(n)
            and not O.ora_is_prime(n)]
    with open(os.path.join(HERE, "spsp2_survivors_1e6.json"), "w") as f:
        json.dump({"desc": "base-2 strong pseudoprimes < 1e6 with no prime factor <= 113", "values": surv}, f)


if __name__ == "__main__":
    main(sys.argv[1:] or ["base"])
