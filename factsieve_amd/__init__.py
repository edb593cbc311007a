"""factsieve_amd -- MI355X-native FactSieve (factors p of n!+-1) behind a C ABI.

The product is the shared library ``libfactsieve.so`` (HIP kernels for gfx950 and
the C++ host driver) and the ``factsieve`` executable; this module is a thin
ctypes binding over ``include/factsieve.h`` for tests, benchmarks and Python
callers.  Loading fails loudly when the library has not been built: there is no
CPU fallback for any entry point that runs the sieve.

Reference interface mirrored here (mfl0p/FactSieve):
    cl_sieve(hw, st, sd)  cl_sieve.cpp:918-1270  -> sieve(pmin, pmax, nmin, nmax, ...)
    run_test(hw, st, sd)  cl_sieve.cpp:1273-1391 -> self_test()
    isPrime / verify      verifyprime.c:96-164   -> is_prime / verify_factor
    parse_uint64          putil.c:33-115         -> parse_uint64
    the ten kernels       kernels/<name>.cl      -> Segment.<name>() launches
"""
import ctypes
import threading
import os
from collections import namedtuple

_HERE = os.path.dirname(os.path.abspath(__file__))
# FACTSIEVE_LIB names another build of the same C ABI (the host-sanitizer
# build of factsieve_amd.build.build_sanitized, for instance).
LIB_PATH = os.environ.get("FACTSIEVE_LIB") or os.path.join(_HERE, "libfactsieve.so")
EXE_PATH = os.path.join(os.path.dirname(LIB_PATH), "factsieve")  # built beside the library

FS_OK = 0
FS_E_ARG, FS_E_HIP, FS_E_NOMEM, FS_E_OVERFLOW, FS_E_VALIDATE = -1, -2, -3, -4, -5
FS_E_POWERTABLE, FS_E_VERIFY, FS_E_IO, FS_E_NODEV, FS_E_BUSY = -6, -7, -8, -9, -10
PATH_AUTO, PATH_FP64, PATH_HYBRID, PATH_MONT = 0, 1, 2, 3
VERIFY_AUTO, VERIFY_CPU, VERIFY_GPU = 0, 1, 2
CTR_COUNT, CTR_MAXCOUNT, CTR_NFACTORS, CTR_POWERTABLE = 0, 1, 2, 3
CTR_PRIMEOVF, CTR_VALIDATE, CTR_FACTOROVF, CTR_WORDS = 4, 5, 6, 8
SUM_WORDS = 2

u64, u32, i32, i64 = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32, ctypes.c_int64
vp = ctypes.c_void_p


class Factor(ctypes.Structure):
    _fields_ = [("p", u64), ("nc", i32), ("pad", i32)]


class Cands(ctypes.Structure):
    _fields_ = [("p", vp), ("q", vp), ("r2", vp), ("res", vp), ("nnext", vp), ("cap", u32),
                ("grid_hint", u32)]


FRACTION_FN = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_void_p)
SECTION_FN = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
TRICKLE_FN = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p)


class BoincHooks(ctypes.Structure):
    """fs_boinc_hooks: the BOINC API calls of a run (cl_sieve.cpp:47-69, 1119-1139)."""
    _fields_ = [("standalone", ctypes.c_int), ("fraction_done", FRACTION_FN),
                ("begin_critical_section", SECTION_FN), ("end_critical_section", SECTION_FN),
                ("checkpoint_completed", SECTION_FN), ("send_trickle_up", TRICKLE_FN),
                ("starting_elapsed_time", ctypes.c_double), ("user", ctypes.c_void_p)]


class Options(ctypes.Structure):
    _fields_ = [("path", ctypes.c_int), ("verify_threads", ctypes.c_int), ("write_files", ctypes.c_int),
                ("resume", ctypes.c_int), ("quiet", ctypes.c_int), ("test_mode", ctypes.c_int),
                ("seg_candidates", u32), ("n_chunk", u32),
                ("checkpoint_secs", u32), ("workdir", ctypes.c_char_p), ("stop_after_segments", u64),
                ("verify_device", ctypes.c_int), ("boinc", ctypes.POINTER(BoincHooks))]


class WorkStatus(ctypes.Structure):
    """fs_work_status: the 72-byte checkpoint record (workStatus)."""
    _fields_ = [("pmin", u64), ("pmax", u64), ("p", u64), ("checksum", u64), ("primecount", u64),
                ("factorcount", u64), ("last_trickle", u64), ("state_sum", u64), ("nmin", u32), ("nmax", u32)]


class Result(ctypes.Structure):
    _fields_ = [("factorcount", u64), ("primecount", u64), ("checksum", u64), ("raw_factors", u64),
                ("steps", u64), ("setup_mulmods", u64), ("seconds", ctypes.c_double),
                ("kernel_ms_iterate", ctypes.c_double), ("kernel_ms_setup", ctypes.c_double),
                ("kernel_ms_primegen", ctypes.c_double), ("iterate_launches", u32), ("path", ctypes.c_int),
                ("segments", u32), ("resumed", u32), ("lanes", u32)]


class Job(ctypes.Structure):
    _fields_ = [("pmin", u64), ("pmax", u64), ("nmin", u32), ("nmax", u32)]


class JobResult(ctypes.Structure):
    _fields_ = [("factorcount", u64), ("primecount", u64), ("checksum", u64), ("raw_factors", u64),
                ("steps", u64), ("status", ctypes.c_int), ("path", ctypes.c_int)]


class BatchOptions(ctypes.Structure):
    _fields_ = [("tile_candidates", u32), ("persistent_blocks", u32), ("use_graph", u32), ("lpt_order", u32),
                ("verify_threads", ctypes.c_int), ("lanes_per_prime", u32), ("n_chunk", u32), ("unroll", u32)]


class ServerOptions(ctypes.Structure):
    _fields_ = [("mem_budget", u64), ("max_slots", u32), ("n_chunk", u32), ("verify_threads", ctypes.c_int),
                ("use_graph", u32)]


class FactSieveError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"factsieve error {code}: {msg}")
        self.code = code


# (name, restype, argtypes) of every C-ABI entry point used from Python
_SIGS = [
    ("fs_init", ctypes.c_int, [ctypes.c_int]),
    ("fs_valu_peak", ctypes.c_int, [ctypes.c_int, u32, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]),
    ("fs_trickle_up_text", ctypes.c_int, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_char_p,
                                          ctypes.c_size_t]),
    ("fs_read_init_data", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p,
                                         ctypes.c_size_t]),
    ("fs_device_info", ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(u32),
                                      ctypes.POINTER(u64), ctypes.POINTER(u64)]),
    ("fs_last_error", ctypes.c_char_p, []),
    ("fs_inject_fault", ctypes.c_int, [ctypes.c_int]),
    ("fs_malloc", ctypes.c_int, [ctypes.POINTER(vp), ctypes.c_size_t]),
    ("fs_free", ctypes.c_int, [vp]),
    ("fs_memcpy_htod", ctypes.c_int, [vp, vp, ctypes.c_size_t, vp, ctypes.c_int]),
    ("fs_memcpy_dtoh", ctypes.c_int, [vp, vp, ctypes.c_size_t, vp, ctypes.c_int]),
    ("fs_memset", ctypes.c_int, [vp, ctypes.c_int, ctypes.c_size_t, vp]),
    ("fs_memcpy_dtod", ctypes.c_int, [vp, vp, ctypes.c_size_t, vp]),
    ("fs_cands_alloc", ctypes.c_int, [ctypes.POINTER(Cands), u32]),
    ("fs_cands_free", ctypes.c_int, [ctypes.POINTER(Cands)]),
    ("fs_stream_create", ctypes.c_int, [ctypes.POINTER(vp)]),
    ("fs_stream_destroy", ctypes.c_int, [vp]),
    ("fs_stream_sync", ctypes.c_int, [vp]),
    ("fs_event_create", ctypes.c_int, [ctypes.POINTER(vp)]),
    ("fs_event_destroy", ctypes.c_int, [vp]),
    ("fs_event_record", ctypes.c_int, [vp, vp]),
    ("fs_event_query", ctypes.c_int, [vp]),
    ("fs_event_sync", ctypes.c_int, [vp]),
    ("fs_event_elapsed_ms", ctypes.c_int, [vp, vp, ctypes.POINTER(ctypes.c_float)]),
    ("fs_graph_begin", ctypes.c_int, [vp]),
    ("fs_graph_end", ctypes.c_int, [vp, ctypes.POINTER(vp)]),
    ("fs_graph_launch", ctypes.c_int, [vp, vp]),
    ("fs_graph_destroy", ctypes.c_int, [vp]),
    ("fs_clearn", ctypes.c_int, [vp, vp]),
    ("fs_clearresult", ctypes.c_int, [vp, vp, vp]),
    ("fs_getsegprimes", ctypes.c_int, [u64, u64, ctypes.POINTER(Cands), vp, vp]),
    ("fs_setup", ctypes.c_int, [ctypes.POINTER(Cands), vp, vp, u32, u32, u32, u32, ctypes.c_int, vp]),
    ("fs_iterate", ctypes.c_int, [ctypes.POINTER(Cands), vp, vp, u32, u32, u32, ctypes.c_int, vp]),
    ("fs_iterate_noalloc", ctypes.c_int, [ctypes.POINTER(Cands), vp, vp, u32, u32, u32, ctypes.c_int, vp]),
    ("fs_check", ctypes.c_int, [ctypes.POINTER(Cands), vp, vp, u32, vp]),
    ("fs_verify_bytes", ctypes.c_size_t, []),
    ("fs_verifyslow", ctypes.c_int, [ctypes.POINTER(Cands), vp, vp, u32, vp]),
    ("fs_verifypow", ctypes.c_int, [ctypes.POINTER(Cands), vp, vp, vp, u32, vp, vp]),
    ("fs_verifyreduce", ctypes.c_int, [ctypes.POINTER(Cands), vp, vp, vp]),
    ("fs_verifyresult", ctypes.c_int, [ctypes.POINTER(Cands), vp, vp, vp]),
    ("fs_power_table", i64, [u32, vp, vp, vp]),
    ("fs_setup_program", i64, [u32, ctypes.c_int, vp]),
    ("fs_select_path", ctypes.c_int, [u64]),
    ("fs_is_prime", ctypes.c_int, [u64]),
    ("fs_verify_factor", ctypes.c_int, [u64, u32, i32]),
    ("fs_verify_factors", i64, [ctypes.POINTER(Factor), u64, ctypes.c_int]),
    ("fs_verify_factors_gpu", i64, [ctypes.POINTER(Factor), u64, vp]),
    ("fs_read_checkpoint", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(WorkStatus), ctypes.POINTER(ctypes.c_int)]),
    ("fs_write_checkpoint", ctypes.c_int, [ctypes.c_char_p, ctypes.POINTER(WorkStatus), ctypes.c_int]),
    ("fs_parse_uint64", ctypes.c_int, [ctypes.POINTER(u64), ctypes.c_char_p, u64, u64]),
    ("fs_hostemu_setup", ctypes.c_int, [ctypes.c_int, u64, u32, ctypes.POINTER(u64)]),
    ("fs_hostemu_iterate", ctypes.c_int, [ctypes.c_int, u64, u32, u32, ctypes.POINTER(u64),
                                          ctypes.POINTER(Factor), u32, ctypes.POINTER(u32)]),
    ("fs_hostemu_candidates", i64, [u64, u64, vp, u64]),
    ("fs_hostemu_mulmod", ctypes.c_int, [ctypes.c_int, u64, i64, u64, ctypes.POINTER(i64), ctypes.POINTER(u64)]),
    ("fs_hostemu_fp_exact", ctypes.c_int, [u64, i64, u32, ctypes.POINTER(i64), ctypes.POINTER(u64)]),
    ("fs_hostemu_mulmod_hi", ctypes.c_int, [u64, i64, u32, ctypes.POINTER(i64), ctypes.POINTER(u64)]),
    ("fs_hostemu_fp_pair", ctypes.c_int, [u64, i64, u32, ctypes.POINTER(i64), ctypes.POINTER(u64),
                                          ctypes.POINTER(ctypes.c_int)]),
    ("fs_hostemu_sh_pair", ctypes.c_int, [u64, i64, u32, ctypes.POINTER(i64), ctypes.POINTER(u64)]),
    ("fs_hostemu_setup_mul", ctypes.c_int, [u64, u64, u64, ctypes.c_int, ctypes.POINTER(u64), ctypes.POINTER(u64),
                                            ctypes.POINTER(u32)]),
    ("fs_default_options", None, [ctypes.POINTER(Options)]),
    ("fs_sieve", ctypes.c_int, [u64, u64, u32, u32, ctypes.POINTER(Options), ctypes.POINTER(Result),
                                ctypes.POINTER(Factor), u64, ctypes.POINTER(u64)]),
    ("fs_self_test", ctypes.c_int, [ctypes.POINTER(Options), ctypes.POINTER(ctypes.c_int)]),
    ("fs_default_batch_options", None, [ctypes.POINTER(BatchOptions)]),
    ("fs_batch", ctypes.c_int, [ctypes.POINTER(Job), u32, ctypes.POINTER(BatchOptions), ctypes.POINTER(JobResult),
                                ctypes.POINTER(Factor), u64, ctypes.POINTER(u64), ctypes.POINTER(ctypes.c_double)]),
    ("fs_batch_autotune", ctypes.c_int, [ctypes.POINTER(Job), u32, ctypes.POINTER(BatchOptions)]),
    ("fs_default_server_options", None, [ctypes.POINTER(ServerOptions)]),
    ("fs_server_create", ctypes.c_int, [ctypes.POINTER(ServerOptions), ctypes.POINTER(vp)]),
    ("fs_server_submit", i64, [vp, ctypes.POINTER(Job)]),
    ("fs_server_step", ctypes.c_int, [vp]),
    ("fs_server_poll", ctypes.c_int, [vp, i64, ctypes.POINTER(JobResult), ctypes.POINTER(Factor), u64,
                                      ctypes.POINTER(u64)]),
    ("fs_server_stats", ctypes.c_int, [vp, ctypes.POINTER(u32), ctypes.POINTER(u32), ctypes.POINTER(u32),
                                       ctypes.POINTER(u64)]),
    ("fs_server_destroy", ctypes.c_int, [vp]),
]

_lib = None


def lib():
    """The loaded C library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with python -m factsieve_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    m = lib().fs_last_error()
    return m.decode() if m else ""


def check(rc):
    if rc < 0:
        raise FactSieveError(rc, last_error())
    return rc


def init(device=0):
    return check(lib().fs_init(device))


def device_info():
    name = ctypes.create_string_buffer(256)
    cus, total, free = u32(), u64(), u64()
    check(lib().fs_device_info(name, 256, ctypes.byref(cus), ctypes.byref(total), ctypes.byref(free)))
    return {"name": name.value.decode(), "compute_units": cus.value, "total_mem": total.value,
            "free_mem": free.value}


# ------------------------------------------------------------------ host utilities

def is_prime(p):
    return bool(lib().fs_is_prime(p))


def verify_factor(p, n, c):
    return bool(lib().fs_verify_factor(p, n, c))


def verify_factors(factors, threads=16):
    """CPU verification of [(p, nc), ...] (nc < 0: n!-1); index of the first
    factor that fails, or -1."""
    arr = (Factor * max(len(factors), 1))()
    for i, (p, nc) in enumerate(factors):
        arr[i].p, arr[i].nc, arr[i].pad = p, nc, 0
    r = lib().fs_verify_factors(arr, len(factors), threads)
    if r < -1:
        raise FactSieveError(r, last_error())
    return r


def verify_factors_gpu(factors, stream=None):
    """The same verification on the GPU (fs_verify_factors_gpu): index of the
    first factor that fails, or -1."""
    arr = (Factor * max(len(factors), 1))()
    for i, (p, nc) in enumerate(factors):
        arr[i].p, arr[i].nc, arr[i].pad = p, nc, 0
    r = lib().fs_verify_factors_gpu(arr, len(factors), stream)
    if r < -1:
        raise FactSieveError(r, last_error())
    return r


def read_checkpoint(directory, pmin, pmax, nmin, nmax):
    """read_state: (WorkStatus, 0 for stateA / 1 for stateB) or None."""
    st = WorkStatus(pmin=pmin, pmax=pmax, nmin=nmin, nmax=nmax)
    which = ctypes.c_int(-1)
    r = check(lib().fs_read_checkpoint(str(directory).encode(), ctypes.byref(st), ctypes.byref(which)))
    return (st, which.value) if r == 1 else None


def write_checkpoint(directory, st, which):
    """write_state into stateA.ckp (which=0) or stateB.ckp (which=1); sets st.state_sum."""
    check(lib().fs_write_checkpoint(str(directory).encode(), ctypes.byref(st), which))
    return st


VALU_PEAK_KINDS = {0: "v_mad_u64_u32", 1: "v_mul_lo_u32", 2: "v_mul_hi_u32", 3: "v_fma_f64"}


def valu_peak(kind, iters=1 << 16):
    """(lane-operations per second, launch ms) of one VALU instruction kind
    (VALU_PEAK_KINDS) at 8 waves per SIMD."""
    rate, ms = ctypes.c_double(), ctypes.c_double()
    check(lib().fs_valu_peak(kind, iters, ctypes.byref(rate), ctypes.byref(ms)))
    return rate.value, ms.value


def trickle_up_text(progress, cputime, runtime):
    """The daily trickle-up message (cl_sieve.cpp:59-64)."""
    buf = ctypes.create_string_buffer(512)
    check(lib().fs_trickle_up_text(progress, cputime, runtime, buf, 512))
    return buf.value.decode()


def read_init_data(directory):
    """(gpu_device_num, gpu_type) from init_data.xml in `directory`, or None
    when there is no such file; FactSieveError when it has no valid
    <gpu_device_num>."""
    dev = ctypes.c_int(-1)
    typ = ctypes.create_string_buffer(64)
    r = check(lib().fs_read_init_data(str(directory).encode(), ctypes.byref(dev), typ, 64))
    return (dev.value, typ.value.decode()) if r == 1 else None


def parse_uint64(s, lo, hi):
    """(status, value): status 0 ok, -1 cannot parse, -2 out of range (putil.c)."""
    out = u64(0)
    st = lib().fs_parse_uint64(ctypes.byref(out), s.encode(), lo, hi)
    return st, out.value


def select_path(pmax):
    return lib().fs_select_path(pmax)


def power_table(nmin):
    n = check(lib().fs_power_table(nmin, None, None, None))
    b, e, c = (u64 * max(n, 1))(), (u32 * max(n, 1))(), (u32 * max(n, 1))()
    lib().fs_power_table(nmin, b, e, c)
    return list(b[:n]), list(e[:n]), list(c[:n])


def setup_program(nmin, path):
    """(terms, words): the device words of the setup program -- FP64: terms
    doubles; HYBRID: terms multipliers then the two power-of-two exponent words;
    MONT: the power table as bases[terms] then (exp | curbit << 32)[terms]."""
    n = check(lib().fs_setup_program(nmin, path, None))
    words = (u64 * (max(2 * n, 1) + 2))()
    lib().fs_setup_program(nmin, path, words)
    nw = {PATH_MONT: 2 * n, PATH_HYBRID: n + 2}.get(path, n)
    return n, list(words[:nw])


def setup_program_terms(nmin, path):
    """Number of setup program terms for nmin on a path (no words copied)."""
    return check(lib().fs_setup_program(nmin, path, None))


def hostemu_setup(path, p, nmin):
    r = u64()
    check(lib().fs_hostemu_setup(path, p, nmin, ctypes.byref(r)))
    return r.value


def hostemu_mulmod(path, p, s, m):
    """One residue step s*m mod p of a path on the host: (raw signed value, canonical)."""
    raw, res = i64(), u64()
    check(lib().fs_hostemu_mulmod(path, p, s, m, ctypes.byref(raw), ctypes.byref(res)))
    return raw.value, res.value


def hostemu_fp_exact(p, s, m):
    """One FP64 residue step without the product's error term (p * m <= 2^52)
    on the host: (raw signed value, canonical)."""
    raw, res = i64(), u64()
    check(lib().fs_hostemu_fp_exact(p, s, m, ctypes.byref(raw), ctypes.byref(res)))
    return raw.value, res.value


def hostemu_mulmod_hi(p, s, m):
    """One hybrid residue step with the high-word quotient (p >= m * 2^34 below
    2^63, p >= m * 2^39 above) on the host: (raw signed value, canonical)."""
    raw, res = i64(), u64()
    check(lib().fs_hostemu_mulmod_hi(p, s, m, ctypes.byref(raw), ctypes.byref(res)))
    return raw.value, res.value


def hostemu_fp_pair(p, s, a):
    """One FP64 paired step s*a*(a+1) mod p (64 (a+1) <= 31 p) on the host:
    (raw signed value, canonical, whether the exact-product form ran)."""
    raw, res, ex = i64(), u64(), ctypes.c_int()
    check(lib().fs_hostemu_fp_pair(p, s, a, ctypes.byref(raw), ctypes.byref(res), ctypes.byref(ex)))
    return raw.value, res.value, bool(ex.value)


def hostemu_sh_pair(p, s, a):
    """One hybrid paired step s*a*(a+1) mod p (a + 1 <= 2^22, 2^45 <= p <
    2^64 - 2^58) on the host: (raw signed value, canonical)."""
    raw, res = i64(), u64()
    check(lib().fs_hostemu_sh_pair(p, s, a, ctypes.byref(raw), ctypes.byref(res)))
    return raw.value, res.value


def hostemu_setup_mul(p, s, m, square=False):
    """One multiply (or square) of the hybrid setup program, s*m*2^-k mod p, on
    the host: (the value the setup loop keeps, canonical, k)."""
    raw, res, k = u64(), u64(), u32()
    check(lib().fs_hostemu_setup_mul(p, s, m, int(square), ctypes.byref(raw), ctypes.byref(res), ctypes.byref(k)))
    return raw.value, res.value, k.value


def hostemu_iterate(path, p, n0, n1, res, cap=64):
    r = u64(res)
    f = (Factor * cap)()
    nf = u32()
    check(lib().fs_hostemu_iterate(path, p, n0, n1, ctypes.byref(r), f, cap, ctypes.byref(nf)))
    return r.value, [(f[i].p, f[i].nc) for i in range(min(nf.value, cap))], nf.value


def hostemu_candidates(low, high):
    n = check(lib().fs_hostemu_candidates(low, high, None, 0))
    out = (u64 * max(n, 1))()
    lib().fs_hostemu_candidates(low, high, out, n)
    return list(out[:n])


# ------------------------------------------------------------------ whole sieve

SieveResult = namedtuple("SieveResult", "factors factorcount primecount checksum raw_factors steps seconds "
                                        "kernel_ms_iterate kernel_ms_setup kernel_ms_primegen path segments "
                                        "iterate_launches setup_mulmods resumed lanes")


def default_options(**kw):
    o = Options()
    lib().fs_default_options(ctypes.byref(o))
    for k, v in kw.items():
        if k == "workdir" and v is not None:
            v = v.encode() if isinstance(v, str) else v
        if k == "boinc" and isinstance(v, BoincHooks):
            v = ctypes.pointer(v)  # the caller keeps the hooks alive for the call
        setattr(o, k, v)
    return o


_tls = threading.local()


def _factor_buffer(cap):
    """Per-thread reusable factor array (results are copied out right after a
    call), so repeated calls do not pay for a fresh zeroed host buffer."""
    buf = getattr(_tls, "fac", None)
    if buf is None or len(buf) < cap:
        buf = (Factor * cap)()
        _tls.fac = buf
    return buf


def sieve(pmin, pmax, nmin, nmax, factor_cap=1 << 20, **opts):
    """Sieve p in [pmin, pmax), n in [nmin, nmax) on the GPU (cl_sieve)."""
    o = default_options(**opts)
    r = Result()
    f = _factor_buffer(factor_cap)
    nf = u64()
    check(lib().fs_sieve(pmin, pmax, nmin, nmax, ctypes.byref(o), ctypes.byref(r), f, factor_cap,
                         ctypes.byref(nf)))
    facs = [(f[i].p, f[i].nc) for i in range(min(nf.value, factor_cap))]
    return SieveResult(facs, r.factorcount, r.primecount, r.checksum, r.raw_factors, r.steps, r.seconds,
                       r.kernel_ms_iterate, r.kernel_ms_setup, r.kernel_ms_primegen, r.path, r.segments,
                       r.iterate_launches, r.setup_mulmods, r.resumed, r.lanes)


cl_sieve = sieve


def self_test(**opts):
    o = default_options(**opts)
    mask = ctypes.c_int()
    good = lib().fs_self_test(ctypes.byref(o), ctypes.byref(mask))
    return good, mask.value


run_test = self_test


def format_factor(p, nc):
    """factors.txt line of one factor (cl_sieve.cpp:539)."""
    return f"{p} | {abs(nc)}!{'-1' if nc < 0 else '+1'}"


# ------------------------------------------------------------------ low-level segment

class Segment:
    """Device buffers of one segment and launches of the individual kernels on
    one stream, mirroring the reference's per-kernel sclEnqueueKernel calls."""

    def __init__(self, cap, factor_cap=1 << 16):
        L = lib()
        self.c = Cands()
        check(L.fs_cands_alloc(ctypes.byref(self.c), cap))
        self.ctr, self.sum, self.fac, self.ver, self.prog = vp(), vp(), vp(), vp(), vp()
        self.stream = vp()
        check(L.fs_stream_create(ctypes.byref(self.stream)))
        check(L.fs_malloc(ctypes.byref(self.ctr), CTR_WORDS * 4))
        check(L.fs_malloc(ctypes.byref(self.sum), SUM_WORDS * 8))
        check(L.fs_malloc(ctypes.byref(self.fac), factor_cap * 16))
        check(L.fs_malloc(ctypes.byref(self.ver), L.fs_verify_bytes()))
        self.factor_cap = factor_cap
        self.prog_terms = 0
        self.pt = None
        check(L.fs_memset(self.ctr, 0, CTR_WORDS * 4, self.stream))
        check(L.fs_memset(self.sum, 0, SUM_WORDS * 8, self.stream))
        self.sync()

    def close(self):
        L = lib()
        L.fs_cands_free(ctypes.byref(self.c))
        for b in (self.ctr, self.sum, self.fac, self.ver, self.prog):
            L.fs_free(b)
        if self.pt:
            for b in self.pt[:2]:
                L.fs_free(b)
        L.fs_stream_destroy(self.stream)

    def sync(self):
        check(lib().fs_stream_sync(self.stream))

    def clearn(self):
        check(lib().fs_clearn(self.ctr, self.stream))

    def clearresult(self):
        check(lib().fs_clearresult(self.ctr, self.sum, self.stream))

    def getsegprimes(self, low, high):
        check(lib().fs_getsegprimes(low, high, ctypes.byref(self.c), self.ctr, self.stream))

    def load_program(self, nmin, path):
        L = lib()
        n, words = setup_program(nmin, path)
        arr = (u64 * max(len(words), 1))(*words)
        if self.prog:
            L.fs_free(self.prog)
            self.prog = vp()
        check(L.fs_malloc(ctypes.byref(self.prog), max(len(words), 1) * 8))
        check(L.fs_memcpy_htod(self.prog, arr, len(words) * 8, self.stream, 1))
        self.prog_terms = n
        return n

    def setup(self, nmin, path, start=0, end=None):
        end = self.prog_terms if end is None else end
        check(lib().fs_setup(ctypes.byref(self.c), self.ctr, self.prog, self.prog_terms, start, end, nmin, path,
                             self.stream))

    def iterate(self, n0, n1, path):
        check(lib().fs_iterate(ctypes.byref(self.c), self.ctr, self.fac, self.factor_cap, n0, n1, path,
                               self.stream))

    def iterate_noalloc(self, n0, n1, path):
        """fs_iterate with the library's static tables (no device allocation)."""
        check(lib().fs_iterate_noalloc(ctypes.byref(self.c), self.ctr, self.fac, self.factor_cap, n0, n1, path,
                                       self.stream))

    def check(self, nmax):
        check(lib().fs_check(ctypes.byref(self.c), self.ctr, self.sum, nmax, self.stream))

    def verify(self, nmin):
        L = lib()
        b, e, _ = power_table(nmin)
        pb, pe = vp(), vp()
        check(L.fs_malloc(ctypes.byref(pb), max(len(b), 1) * 8))
        check(L.fs_malloc(ctypes.byref(pe), max(len(e), 1) * 4))
        ab, ae = (u64 * max(len(b), 1))(*b), (u32 * max(len(e), 1))(*e)
        check(L.fs_memcpy_htod(pb, ab, len(b) * 8, self.stream, 1))
        check(L.fs_memcpy_htod(pe, ae, len(e) * 4, self.stream, 1))
        check(L.fs_verifyslow(ctypes.byref(self.c), self.ctr, self.ver, nmin, self.stream))
        check(L.fs_verifypow(ctypes.byref(self.c), self.ctr, pb, pe, len(b), self.ver, self.stream))
        check(L.fs_verifyreduce(ctypes.byref(self.c), self.ctr, self.ver, self.stream))
        check(L.fs_verifyresult(ctypes.byref(self.c), self.ctr, self.ver, self.stream))
        self.sync()
        L.fs_free(pb)
        L.fs_free(pe)

    def counters(self):
        h = (u32 * CTR_WORDS)()
        check(lib().fs_memcpy_dtoh(h, self.ctr, CTR_WORDS * 4, self.stream, 1))
        return list(h)

    def sums(self):
        h = (u64 * SUM_WORDS)()
        check(lib().fs_memcpy_dtoh(h, self.sum, SUM_WORDS * 8, self.stream, 1))
        return list(h)

    def read_u64(self, field, n):
        h = (u64 * max(n, 1))()
        check(lib().fs_memcpy_dtoh(h, getattr(self.c, field), n * 8, self.stream, 1))
        return list(h[:n])

    def read_u32(self, field, n):
        h = (u32 * max(n, 1))()
        check(lib().fs_memcpy_dtoh(h, getattr(self.c, field), n * 4, self.stream, 1))
        return list(h[:n])

    def write_u64(self, field, values):
        h = (u64 * max(len(values), 1))(*values)
        check(lib().fs_memcpy_htod(getattr(self.c, field), h, len(values) * 8, self.stream, 1))

    def write_u32(self, field, values):
        h = (u32 * max(len(values), 1))(*values)
        check(lib().fs_memcpy_htod(getattr(self.c, field), h, len(values) * 4, self.stream, 1))

    def factors(self):
        n = min(self.counters()[CTR_NFACTORS], self.factor_cap)
        h = (Factor * max(n, 1))()
        check(lib().fs_memcpy_dtoh(h, self.fac, n * 16, self.stream, 1))
        return [(h[i].p, h[i].nc) for i in range(n)]


# ------------------------------------------------------------------ batched form and serving

JobOutcome = namedtuple("JobOutcome", "factorcount primecount checksum raw_factors steps status path factors")


def default_batch_options(**kw):
    o = BatchOptions()
    lib().fs_default_batch_options(ctypes.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def _jobs_array(jobs):
    arr = (Job * max(len(jobs), 1))()
    for i, (pmin, pmax, nmin, nmax) in enumerate(jobs):
        arr[i].pmin, arr[i].pmax, arr[i].nmin, arr[i].nmax = pmin, pmax, nmin, nmax
    return arr


def batch(jobs, factor_cap=1 << 20, **opts):
    """Run many (pmin, pmax, nmin, nmax) jobs in one batched persistent launch.
    Returns (list of JobOutcome, wall seconds)."""
    o = default_batch_options(**opts)
    arr = _jobs_array(jobs)
    out = (JobResult * max(len(jobs), 1))()
    fac = _factor_buffer(factor_cap)
    per = (u64 * max(len(jobs), 1))()
    sec = ctypes.c_double()
    check(lib().fs_batch(arr, len(jobs), ctypes.byref(o), out, fac, factor_cap, per, ctypes.byref(sec)))
    res, k = [], 0
    for i in range(len(jobs)):
        r = out[i]
        fl = [(fac[k + t].p, fac[k + t].nc) for t in range(per[i]) if k + t < factor_cap]
        k += per[i]
        res.append(JobOutcome(r.factorcount, r.primecount, r.checksum, r.raw_factors, r.steps, r.status, r.path, fl))
    return res, sec.value


def batch_autotune(jobs, **opts):
    o = default_batch_options(**opts)
    arr = _jobs_array(jobs)
    check(lib().fs_batch_autotune(arr, len(jobs), ctypes.byref(o)))
    return {f: getattr(o, f) for f, _ in BatchOptions._fields_}


class Server:
    """Continuous-batching serving loop over the C ABI (fs_server_*)."""

    def __init__(self, mem_budget=0, max_slots=256, n_chunk=49152, verify_threads=16, use_graph=1):
        o = ServerOptions()
        lib().fs_default_server_options(ctypes.byref(o))
        o.mem_budget, o.max_slots, o.n_chunk, o.verify_threads = mem_budget, max_slots, n_chunk, verify_threads
        o.use_graph = use_graph
        self.h = vp()
        check(lib().fs_server_create(ctypes.byref(o), ctypes.byref(self.h)))

    def submit(self, job):
        j = _jobs_array([job])
        return check(lib().fs_server_submit(self.h, ctypes.byref(j[0])))

    def step(self):
        return check(lib().fs_server_step(self.h))

    def poll(self, wid, factor_cap=1 << 16):
        r = JobResult()
        fac = _factor_buffer(factor_cap)
        nf = u64()
        done = check(lib().fs_server_poll(self.h, wid, ctypes.byref(r), fac, factor_cap, ctypes.byref(nf)))
        if not done:
            return None
        fl = [(fac[i].p, fac[i].nc) for i in range(min(nf.value, factor_cap))]
        return JobOutcome(r.factorcount, r.primecount, r.checksum, r.raw_factors, r.steps, r.status, r.path, fl)

    def stats(self):
        q, a, d, b = u32(), u32(), u32(), u64()
        check(lib().fs_server_stats(self.h, ctypes.byref(q), ctypes.byref(a), ctypes.byref(d), ctypes.byref(b)))
        return {"queued": q.value, "active": a.value, "done": d.value, "bytes_in_use": b.value}

    def close(self):
        if self.h:
            lib().fs_server_destroy(self.h)
            self.h = vp()
