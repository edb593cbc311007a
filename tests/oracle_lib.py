"""ctypes access to the CPU oracle (oracle/liboracle.so) and, where it was
built, to the reference's own verifyprime.c / putil.c (oracle/_ref/libfsref.so).

Test infrastructure only: the product never imports this module."""
import ctypes
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfsref.so")

u64, u32, i32, i64 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int64


class OFactor(ctypes.Structure):
    _fields_ = [("p", u64), ("nc", i32), ("pad", i32)]


class OMont(ctypes.Structure):
    _fields_ = [("p", u64), ("q", u64), ("one", u64), ("two", u64), ("nmo", u64), ("r2", u64)]


_O = None


def oracle():
    global _O
    if _O is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle/liboracle.so missing: run __graft_entry__.build()")
        O = ctypes.CDLL(ORACLE_SO)
        sig = {
            "ora_invert": (u64, [u64]),
            "ora_mmul": (u64, [u64, u64, u64, u64]),
            "ora_addmod": (u64, [u64, u64, u64]),
            "ora_mont_init": (None, [u64, ctypes.POINTER(OMont)]),
            "ora_to_mont": (u64, [u64, ctypes.POINTER(OMont)]),
            "ora_sieve_survivor": (ctypes.c_int, [u64]),
            "ora_spsp2": (ctypes.c_int, [u64]),
            "ora_is_candidate": (ctypes.c_int, [u64]),
            "ora_candidates": (u64, [u64, u64, ctypes.c_void_p, u64]),
            "ora_find_wheel_offset": (None, [ctypes.POINTER(u64), ctypes.POINTER(i32)]),
            "ora_get_power": (None, [u32, u32, ctypes.POINTER(u32), ctypes.POINTER(u32)]),
            "ora_factorial_mod": (u64, [u32, u64]),
            "ora_is_prime": (ctypes.c_int, [u64]),
            "ora_verify": (ctypes.c_int, [u64, u32, i32]),
            "ora_iterate": (u32, [ctypes.POINTER(OMont), u32, u32, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                  ctypes.POINTER(OFactor), u32]),
            "ora_finalize": (i64, [ctypes.POINTER(OFactor), u64, ctypes.POINTER(u64), ctypes.POINTER(i64)]),
            "ora_parse_uint64": (ctypes.c_int, [ctypes.POINTER(u64), ctypes.c_char_p, u64, u64]),
            "ora_sieve_device": (ctypes.c_int, [u64, u64, u32, u32, ctypes.c_int, ctypes.POINTER(OFactor), u64,
                                                ctypes.POINTER(u64), ctypes.POINTER(u64), ctypes.POINTER(u64),
                                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, u64]),
        }
        for name, (res, args) in sig.items():
            f = getattr(O, name)
            f.restype, f.argtypes = res, args
        _O = O
    return _O


def ref():
    """The reference's verifyprime.c + putil.c, or None where not built."""
    if not os.path.exists(REF_SO):
        return None
    R = ctypes.CDLL(REF_SO)
    R._Z7isPrimem.restype, R._Z7isPrimem.argtypes = ctypes.c_bool, [u64]
    R._Z6verifymji.restype, R._Z6verifymji.argtypes = ctypes.c_bool, [u64, u32, i32]
    R.parse_uint64.restype, R.parse_uint64.argtypes = ctypes.c_int, [ctypes.POINTER(u64), ctypes.c_char_p, u64, u64]
    return R


def mont(p):
    m = OMont()
    oracle().ora_mont_init(p, ctypes.byref(m))
    return m


def candidates(lo, hi):
    O = oracle()
    n = O.ora_candidates(lo, hi, None, 0)
    buf = (u64 * max(n, 1))()
    O.ora_candidates(lo, hi, buf, n)
    return list(buf[:n])


def sieve(pmin, pmax, nmin, nmax, threads=0, cap=1 << 17, residues=False):
    """Oracle run: dict with raw (device) factors, kept factors, checksums, counts
    and, optionally, per-candidate (p, s6, s7) sorted by p."""
    O = oracle()
    raw = (OFactor * cap)()
    nraw, dck, pc = u64(), u64(), u64()
    ncand = O.ora_candidates(pmin, pmax, None, 0) if residues else 0
    cp = (u64 * max(ncand, 1))()
    c6 = (u64 * max(ncand, 1))()
    c7 = (u64 * max(ncand, 1))()
    rc = O.ora_sieve_device(pmin, pmax, nmin, nmax, threads, raw, cap, ctypes.byref(nraw), ctypes.byref(dck),
                            ctypes.byref(pc), cp if residues else None, c6 if residues else None,
                            c7 if residues else None, ncand)
    if rc != 0:
        raise RuntimeError("oracle factor buffer too small")
    raw_list = [(raw[i].p, raw[i].nc) for i in range(nraw.value)]
    ck = u64(dck.value)
    bad = i64()
    kept = O.ora_finalize(raw, nraw.value, ctypes.byref(ck), ctypes.byref(bad))
    if kept < 0:
        raise RuntimeError(f"oracle factor verification failed at {bad.value}")
    out = {"raw": raw_list, "factors": [(raw[i].p, raw[i].nc) for i in range(kept)], "factorcount": kept,
           "primecount": pc.value, "device_checksum": dck.value, "checksum": ck.value}
    if residues:
        out["cands"] = sorted(zip(list(cp), list(c6), list(c7)))
    return out
