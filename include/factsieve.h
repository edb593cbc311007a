/*
 * factsieve.h -- C ABI of the MI355X-native FactSieve GPU sieve (n!+-1 factors).
 *
 * This is the drop-in boundary that replaces the reference's OpenCL layer:
 *   - simpleCL.h:67-139 (sclGetCLSoftware / sclSetKernelArg / sclEnqueueKernel /
 *     sclRead / sclWrite / release / sclPrintErrorFlags) and the raw cl* calls in
 *     main.cpp:235-303 and cl_sieve.cpp:250-342, 654-1063;
 *   - the ten OpenCL kernels (the kernels/<name>.cl files), which become entry points that
 *     enqueue precompiled gfx950 code objects on a HIP stream;
 *   - the host driver cl_sieve() / run_test() (cl_sieve.cpp:918-1391), kept with
 *     the same inputs, outputs, factors.txt format and BOINC quorum checksum.
 *
 * Conventions: plain C types, device pointers are opaque `void*`-compatible
 * pointers returned by fs_malloc, streams are `fs_stream` handles (0 = the
 * library's default stream).  Every entry point returns FS_OK (0) or a negative
 * FS_E* code and records a message for fs_last_error(); none of them exits the
 * process (the reference prints and calls exit(EXIT_FAILURE), simpleCL.c:347 --
 * the CLI executable restores that behaviour on top of these codes).
 */
#ifndef FACTSIEVE_H
#define FACTSIEVE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define FS_OK 0
#define FS_E_ARG -1          /* invalid argument / range (setupSearch, cl_sieve.cpp:557-589) */
#define FS_E_HIP -2          /* HIP runtime failure (sclPrintErrorFlags, simpleCL.c:30-347) */
#define FS_E_NOMEM -3        /* allocation failure (clCreateBuffer failure, cl_sieve.cpp:929-1068) */
#define FS_E_OVERFLOW -4     /* candidate or factor buffer overflow (cl_sieve.cpp:451-476) */
#define FS_E_VALIDATE -5     /* device validation flag (check.cl:66-78 -> cl_sieve.cpp:463-467) */
#define FS_E_POWERTABLE -6   /* nmin! cross-check failed (verifyresult.cl -> cl_sieve.cpp:903-907) */
#define FS_E_VERIFY -7       /* CPU re-verification of a factor failed (cl_sieve.cpp:502-506) */
#define FS_E_IO -8           /* factors.txt / checkpoint I/O (cl_sieve.cpp:523-527, 742-765) */
#define FS_E_NODEV -9        /* no usable gfx950 device */
#define FS_E_BUSY -10        /* serving: no slot or memory free for the request */

/* Device-side counter block, same index meaning as the reference's
 * g_primecount[6] (clearresult.cl:15-23), plus two bounded-append flags. */
#define FS_CTR_COUNT 0       /* candidates in the current segment */
#define FS_CTR_MAXCOUNT 1    /* largest segment count since the last drain */
#define FS_CTR_NFACTORS 2    /* factors appended since the last drain */
#define FS_CTR_POWERTABLE 3  /* power-table / setup cross-check failed */
#define FS_CTR_PRIMEOVF 4    /* candidate buffer overflow in getsegprimes */
#define FS_CTR_VALIDATE 5    /* residue validation failed in check */
#define FS_CTR_FACTOROVF 6   /* factor buffer overflow in iterate */
#define FS_CTR_WORDS 8

/* Device-side sum block: [0] candidates since last drain, [1] checksum fold of
 * mont(nmax!) + mont(nmax) over those candidates (g_sum, check.cl:88-104). */
#define FS_SUM_WORDS 2

/* Arithmetic formulation of the residue kernels.  All produce bit-identical
 * residues; the driver picks by pmax unless a run forces one. */
#define FS_PATH_AUTO 0
#define FS_PATH_FP64 1       /* p < 2^53: exact FP64 error-free products + rounded quotient */
#define FS_PATH_HYBRID 2     /* p <= 2^64 - 2^46: exact 64x32 integer product, FP64 quotient */
#define FS_PATH_MONT 3       /* any odd p < 2^64: 64-bit Montgomery REDC (reference formulation) */

/* 16-byte factor record, identical layout to `factor` (cl_sieve.h:4-7,
 * iterate.cl:13-16): nc = -n for p | n!-1, +n for p | n!+1. */
typedef struct {
  uint64_t p;
  int32_t nc;
  int32_t pad;
} fs_factor;

/* Candidate set of one segment, structure-of-arrays in HBM (the reference packs
 * a 64-byte ulong8 record per candidate, getsegprimes.cl:227-228).  All arrays
 * have `cap` elements; the live count is counter FS_CTR_COUNT. */
typedef struct {
  uint64_t* p;      /* candidate modulus (base-2 strong probable prime) */
  uint64_t* q;      /* p^-1 mod 2^64 (invert, getsegprimes.cl:73-77) */
  uint64_t* r2;     /* 2^128 mod p */
  uint64_t* res;    /* canonical residue n! mod p in [0, p) between launches */
  uint32_t* nnext;  /* next n the residue belongs to (validated by check) */
  uint32_t cap;
  uint32_t grid_hint; /* 0, or the live count when the host knows it: kernels then
                         launch only enough workgroups for min(grid_hint, cap) */
} fs_cands;

typedef void* fs_stream;
typedef void* fs_event;
typedef void* fs_graph;

/* ------------------------------------------------ device / error handling */
/* Open device `index` (HIP ordinal; main.cpp:230-275 picks platform 0 / GPU 0).
 * Idempotent.  Replaces clGetPlatformIDs/clGetDeviceIDs/clCreateContext/
 * clCreateCommandQueue (main.cpp:235-270). */
int fs_init(int device_index);
/* Name, compute units, total and free bytes (clGetDeviceInfo, main.cpp:283-308). */
int fs_device_info(char* name, size_t name_len, uint32_t* compute_units, uint64_t* total_mem,
                   uint64_t* free_mem);
/* Message of the last failing call on this thread (sclPrintErrorFlags). */
const char* fs_last_error(void);
/* Fault injection for tests: the next check/verify launch raises the given
 * counter flag (FS_CTR_POWERTABLE / _VALIDATE / _PRIMEOVF / _FACTOROVF); 0 clears. */
int fs_inject_fault(int counter_index);

/* -------------------------------------------------- memory (simpleCL.c I/O) */
int fs_malloc(void** dptr, size_t bytes);                 /* clCreateBuffer */
int fs_free(void* dptr);                                  /* sclReleaseMemObject */
int fs_memcpy_htod(void* dst, const void* src, size_t bytes, fs_stream s, int blocking); /* sclWrite[NB] */
int fs_memcpy_dtoh(void* dst, const void* src, size_t bytes, fs_stream s, int blocking); /* sclRead[NB] */
int fs_memset(void* dptr, int value, size_t bytes, fs_stream s); /* no OpenCL call: the reference clears
                                                   device words with its clearn/clearresult kernels */
int fs_memcpy_dtod(void* dst, const void* src, size_t bytes, fs_stream s); /* device copy (also the
                                                   HBM-counter calibration kernel of the benchmark) */
int fs_cands_alloc(fs_cands* c, uint32_t cap);            /* clCreateBuffer(d_primes) (cl_sieve.cpp:1057) */
int fs_cands_free(fs_cands* c);                           /* clReleaseMemObject (cl_sieve.cpp:79-96) */

/* -------------------------------------------- streams, events, graphs */
int fs_stream_create(fs_stream* s);                       /* clCreateCommandQueue (main.cpp:266) */
int fs_stream_destroy(fs_stream s);                       /* sclReleaseClHard (simpleCL.c) */
int fs_stream_sync(fs_stream s);                          /* sclFinish / sleepCPU (cl_sieve.cpp:289-354) */
int fs_event_create(fs_event* e);
int fs_event_destroy(fs_event e);
int fs_event_record(fs_event e, fs_stream s);             /* clEnqueueMarker */
int fs_event_query(fs_event e);                           /* 1 done, 0 pending (clGetEventInfo) */
int fs_event_sync(fs_event e);                            /* waitOnEvent (cl_sieve.cpp:237-285) */
int fs_event_elapsed_ms(fs_event start, fs_event stop, float* ms); /* ProfilesclEnqueueKernel */
/* Graphs have no counterpart in the reference: the batched form captures its
 * segment loop (k_bclear ... k_bsieve) once per shape and replays it. */
int fs_graph_begin(fs_stream s);                          /* start capturing launches on s */
int fs_graph_end(fs_stream s, fs_graph* g);               /* instantiate the captured graph */
int fs_graph_launch(fs_graph g, fs_stream s);
int fs_graph_destroy(fs_graph g);

/* ------------------------------------ kernels (reference kernels/<name>.cl) */
/* clearn.cl:10-19: counter[FS_CTR_COUNT] = 0. */
int fs_clearn(uint32_t* d_ctr, fs_stream s);
/* clearresult.cl:10-25: zero the sum block and counters 1..7. */
int fs_clearresult(uint32_t* d_ctr, uint64_t* d_sum, fs_stream s);
/* getsegprimes.cl:170-239: append every base-2 strong probable prime in
 * [low, high) that has no prime factor <= 113 to `c` (with q and r2);
 * bounded append, overflow raises FS_CTR_PRIMEOVF.  low, high: any values with
 * low <= high (no wheel alignment needed; findWheelOffset is internal). */
int fs_getsegprimes(uint64_t low, uint64_t high, const fs_cands* c, uint32_t* d_ctr, fs_stream s);
/* setup.cl:49-103: res = (res * prod of program[start..end)) per the setup
 * program of `path` (built by fs_setup_program, prog_terms = its term count,
 * see fs_setup_program); start == 0 resets res to 1 (FP64/HYBRID) or starts the
 * power table at the prime 2 (MONT).  Sets nnext = nmin.  d_prog is the device
 * copy of the program words. */
int fs_setup(const fs_cands* c, const uint32_t* d_ctr, const void* d_prog, uint32_t prog_terms,
             uint32_t start, uint32_t end, uint32_t nmin, int path, fs_stream s);
/* iterate.cl:54-80: for n in [startN, endN): record (p, -n) if n! = 1 and
 * (p, +n) if n! = -1 (mod p), then res <- res * (n+1).  Bounded factor append. */
int fs_iterate(const fs_cands* c, uint32_t* d_ctr, fs_factor* d_factor, uint32_t factor_cap,
               uint32_t startN, uint32_t endN, int path, fs_stream s);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): fs_iterate with its multiplier and pair tables in a
 * static device array of the library, refilled by a small kernel on stream s
 * at every call, instead of a table that fs_iterate allocates per stream.  The
 * array holds tables for at most FS_ITERATE_NOALLOC_MAX_N n-values per call
 * (about 24 MiB, one copy on every GPU the library runs on, held until the
 * process ends); endN - startN above that returns FS_E_ARG and launches
 * nothing.  Same arguments, results and errors as fs_iterate otherwise.  Every
 * call on one GPU rewrites the same array and reads it while its kernel runs:
 * calls that overlap on one GPU (other streams or threads) race. */
#define FS_ITERATE_NOALLOC_MAX_N (1u << 20)
int fs_iterate_noalloc(const fs_cands* c, uint32_t* d_ctr, fs_factor* d_factor, uint32_t factor_cap,
                       uint32_t startN, uint32_t endN, int path, fs_stream s);
/* check.cl:50-109: fold mont(nmax!) + mont(nmax) into d_sum[1], count into
 * d_sum[0], track the max count, validate every residue reached nmax. */
int fs_check(const fs_cands* c, uint32_t* d_ctr, uint64_t* d_sum, uint32_t nmax, fs_stream s);
/* verifyslow.cl / verifypow.cl / verifyreduce.cl / verifyresult.cl: for the
 * first candidate of the segment, compute nmin! two ways (direct product and
 * the reference power table) and compare with the setup kernel's residue;
 * mismatch raises FS_CTR_POWERTABLE.  d_verify: scratch of fs_verify_bytes(). */
size_t fs_verify_bytes(void);
int fs_verifyslow(const fs_cands* c, const uint32_t* d_ctr, void* d_verify, uint32_t nmin, fs_stream s);
int fs_verifypow(const fs_cands* c, const uint32_t* d_ctr, const uint64_t* d_bases,
                 const uint32_t* d_exps, uint32_t count, void* d_verify, fs_stream s);
int fs_verifyreduce(const fs_cands* c, const uint32_t* d_ctr, void* d_verify, fs_stream s);
int fs_verifyresult(const fs_cands* c, uint32_t* d_ctr, void* d_verify, fs_stream s);

/* --------------------------------------------------- setup programs (host) */
/* Reference power table (getPower + merge, cl_sieve.cpp:773-838): primes <= nmin,
 * Legendre exponents, runs of equal exponent merged below 2^64-1.  Pass NULL
 * arrays to get the term count only. */
int64_t fs_power_table(uint32_t nmin, uint64_t* bases, uint32_t* exps, uint32_t* curbits);
/* Setup program for `path` in words of 8 bytes; returns the term count T
 * (pass NULL words to get it).  FP64 and HYBRID: T a multiple of 4, in blocks of
 * four multipliers (FP64: doubles < 2^50 as bits; HYBRID: integers < 2^60;
 * padding multipliers are 1).  Bit 63 set on a block's first word (FS_PROG_SQUARE
 * in the kernels) squares the running product before that block; evaluating
 * from 1 yields nmin!.  HYBRID has two more words, the total power of 2^-1 its
 * Montgomery products leave: [T] with multiplies by 2^-62 (p < 2^64/3), [T+1]
 * with multiplies by 2^-64 (other p); squares are by 2^-64 in both.  MONT: the
 * reference power table, 2*T words: bases[T] then (exp | curbit << 32)[T]. */
int64_t fs_setup_program(uint32_t nmin, int path, uint64_t* words);
/* Path the driver uses for p < pmax. */
int fs_select_path(uint64_t pmax);

/* ------------------------------------------------ CPU verifier (verifyprime.c) */
int fs_is_prime(uint64_t p);                              /* isPrime, verifyprime.c:96-132 */
int fs_verify_factor(uint64_t p, uint32_t n, int32_t c);  /* verify, verifyprime.c:137-164 */
/* The OpenMP verify loop over a factor list (cl_sieve.cpp:491-515) on `threads`
 * CPU threads; the sign of nc gives c (negative: n!-1).  Returns the index of
 * the first factor that fails, -1 if all pass, or a negative FS_E_* below -1. */
int64_t fs_verify_factors(const fs_factor* f, uint64_t n, int threads);
/* The same verification on the GPU (verify, verifyprime.c:137-164, in place of
 * the CPU loop of cl_sieve.cpp:491-515): each factor's product 2..n mod p is cut
 * into ranges of 4096 multipliers, one per thread, with integer Montgomery
 * products (independent of the sieve's FP64 and hybrid arithmetic).  Same
 * return contract as fs_verify_factors; runs on stream s and waits for it. */
int64_t fs_verify_factors_gpu(const fs_factor* f, uint64_t n, fs_stream s);
int fs_parse_uint64(uint64_t* result, const char* str, uint64_t lo, uint64_t hi); /* putil.c:33-115 */

/* ----------------------------- host execution of the kernel arithmetic */
/* The per-candidate loops of the setup / iterate / getsegprimes kernels run on
 * the CPU with the same inline arithmetic, for testing the three formulations
 * without a GPU.  No reference counterpart; the sieve driver never calls them. */
int fs_hostemu_setup(int path, uint64_t p, uint32_t nmin, uint64_t* res);
/* Residue *res at startN advanced to endN.  The HYBRID path needs its multipliers
 * below 2p, as in the sieve (a residue is 0 from n = p on): endN < 2p unless
 * *res == 0 (FS_E_ARG otherwise). */
int fs_hostemu_iterate(int path, uint64_t p, uint32_t startN, uint32_t endN, uint64_t* res,
                       fs_factor* factors, uint32_t cap, uint32_t* nfactors);
int64_t fs_hostemu_candidates(uint64_t low, uint64_t high, uint64_t* out, uint64_t cap);
/* One multiply s*m of a path's residue step at the edges of its error bounds:
 * FP64 needs |s| < 0.8p, m < 2^50, p < 2^53; HYBRID |s| <= 0.51p (int64), m < min(2^31, 2p),
 * p <= 2^64 - 2^46 (FS_E_ARG otherwise).  *raw = the step's own signed
 * representation of the result, *res = its canonical value in [0, p). */
int fs_hostemu_mulmod(int path, uint64_t p, int64_t s, uint64_t m, int64_t* raw, uint64_t* res);
/* The FP64 step without the product's error term, used when p * n1 <= 2^53
 * (s*m exact): |s| <= 0.51p, 1 <= m, p * m <= 2^53, odd p >= 7 (FS_E_ARG
 * otherwise); |raw| <= 0.51p. */
int fs_hostemu_fp_exact(uint64_t p, int64_t s, uint32_t m, int64_t* raw, uint64_t* res);
/* The HYBRID step with the quotient from the high word of s, used when every
 * multiplier m <= n1 of a launch has p >= m * 2^34 and p < 2^63, or
 * p >= m * 2^39 and p < 2^64 - 2^58: |s| <= 0.7500005p, 1 <= m, odd p (FS_E_ARG
 * otherwise); |raw| <= 0.7500005p, and <= 0.51p in the second case. */
int fs_hostemu_mulmod_hi(uint64_t p, int64_t s, uint32_t m, int64_t* raw, uint64_t* res);
/* The FP64 paired step, one product by a*(a+1) for the multipliers a and a+1,
 * used when 64 * n1 <= 31 * p and n1 <= 2^22 (for every pair a+1 <= n1 of a launch;
 * without the error term when also p * n1^2 <= 2^53): |s| <= 0.51p, 1 <= a,
 * 64 (a+1) <= 31 p, a + 1 <= 2^22, odd p < 2^53 (FS_E_ARG otherwise); *exact
 * tells which form ran; |raw| <= 0.503p. */
int fs_hostemu_fp_pair(uint64_t p, int64_t s, uint32_t a, int64_t* raw, uint64_t* res, int* exact);
/* The HYBRID paired step, one product by a*(a+1) with the quotient from the
 * estimate of s/p, used when every pair a+1 <= n1 of a launch has n1 <= 2^22
 * and 2^45 <= p < 2^64 - 2^58: |s| <= 0.51p, 1 <= a, a + 1 <= 2^22, odd p in
 * that range (FS_E_ARG otherwise); |raw| <= 0.503p. */
int fs_hostemu_sh_pair(uint64_t p, int64_t s, uint32_t a, int64_t* raw, uint64_t* res);
/* One multiply (square = 0) or square (square = 1, m ignored) of the HYBRID
 * setup program, s*m*2^-k mod p: below 2^64/3 a multiply reduces by k = 62 and
 * a square by k = 64, both lazily (s < 2p, result < 2p); below 2^63 both by
 * k = 64, lazily; above, k = 64 in canonical form (s < p).  m < 2^60, odd
 * 3 <= p <= 2^64 - 2^46 (FS_E_ARG otherwise).  *raw = the value as the setup
 * loop keeps it, *res = its canonical value, *shift = k. */
int fs_hostemu_setup_mul(uint64_t p, uint64_t s, uint64_t m, int square, uint64_t* raw, uint64_t* res,
                         uint32_t* shift);

/* ------------------------------------------------ BOINC client glue (shim) */
/* The calls the reference makes into the BOINC API during a run, as hooks a
 * BOINC wrapper points at boinc_* (BOINC itself is not linked here); a NULL
 * hook is skipped, and a NULL fs_options.boinc is a standalone run:
 *   fraction_done           boinc_fraction_done (cl_sieve.cpp:1119-1124): about
 *                           every 2 s of the segment loop, and 1.0 at the end
 *                           (cl_sieve.cpp:1251)
 *   begin/end_critical_...  boinc_begin/end_critical_section: around each wait
 *                           for the GPU and each drain + checkpoint
 *                           (cl_sieve.cpp:248-280, 1131-1135, 1249-1256)
 *   checkpoint_completed    boinc_checkpoint_completed after a checkpoint write
 *                           (cl_sieve.cpp:225-232)
 *   send_trickle_up         boinc_send_trickle_up (cl_sieve.cpp:47-69): at a
 *                           checkpoint, once a day (workStatus.last_trickle),
 *                           unless standalone; variety "factsieve_progress",
 *                           text as fs_trickle_up_text writes it with the last
 *                           fraction done, the process CPU time and the run's
 *                           elapsed time plus starting_elapsed_time. */
typedef struct {
  int standalone;                 /* 1: boinc_is_standalone(), no trickle-up */
  void (*fraction_done)(double fraction, void* user);
  void (*begin_critical_section)(void* user);
  void (*end_critical_section)(void* user);
  void (*checkpoint_completed)(void* user);
  int (*send_trickle_up)(const char* variety, const char* text, void* user);
  double starting_elapsed_time;   /* APP_INIT_DATA.starting_elapsed_time, seconds */
  void* user;
} fs_boinc_hooks;
/* The trickle-up message of cl_sieve.cpp:59-64 into buf; returns its length, or
 * FS_E_ARG when it does not fit. */
int fs_trickle_up_text(double progress, double cputime, double runtime, char* buf, size_t len);
/* init_data.xml in `dir` ("" = current directory), the stand-alone device choice
 * (README.md:42-56, boinc_get_opencl_ids in main.cpp:230-247): 1 and
 * *gpu_device_num (and the <gpu_type> text into gpu_type, when given) when the
 * file has a <gpu_device_num>; 0 when there is no such file; FS_E_ARG when it
 * has no valid <gpu_device_num> (0..63). */
int fs_read_init_data(const char* dir, int* gpu_device_num, char* gpu_type, size_t gpu_type_len);

/* ------------------------------------------------------- whole sieve */
typedef struct {
  int path;                 /* FS_PATH_* (AUTO = by pmax) */
  int verify_threads;       /* -v: OpenMP threads for CPU verification (default 2) */
  int write_files;          /* 1: factors.txt + stateA/B.ckp in `workdir` (reference behaviour) */
  int resume;               /* 1: resume from checkpoint files when present (non-test runs) */
  int quiet;                /* 1: no stdout progress lines */
  int test_mode;            /* 1: self-test semantics (truncate factors.txt, no resume) */
  uint32_t seg_candidates;  /* target candidates per segment (0 = autotuned) */
  uint32_t n_chunk;         /* n values per iterate launch (0 = whole range) */
  uint32_t checkpoint_secs; /* drain + checkpoint period (reference: 60) */
  const char* workdir;      /* directory for factors.txt / checkpoints (NULL = cwd) */
  uint64_t stop_after_segments; /* test hook: stop (as if killed) after k segments (0 = off) */
  int verify_device;        /* FS_VERIFY_AUTO (default): factors of a drain are re-verified on the
                               GPU when they need more than 2^24 multiplications, else on the CPU;
                               FS_VERIFY_CPU: always on verify_threads CPU threads (the reference);
                               FS_VERIFY_GPU: always on the GPU (fs_verify_factors_gpu) */
  const fs_boinc_hooks* boinc; /* BOINC client hooks (NULL = standalone, no hooks) */
} fs_options;
#define FS_VERIFY_AUTO 0
#define FS_VERIFY_CPU 1
#define FS_VERIFY_GPU 2

typedef struct {
  uint64_t factorcount;     /* kept factors (prime p) */
  uint64_t primecount;      /* candidates sieved */
  uint64_t checksum;        /* BOINC quorum checksum */
  uint64_t raw_factors;     /* factors found on the device before the isPrime filter */
  uint64_t steps;           /* p*n residue steps executed (candidates x (nmax-nmin)) */
  uint64_t setup_mulmods;   /* setup program steps x candidates */
  double seconds;           /* wall time of the device work + drains */
  double kernel_ms_iterate; /* summed iterate kernel time (HIP events) */
  double kernel_ms_setup;
  double kernel_ms_primegen;
  uint32_t iterate_launches;
  int path;                 /* path used */
  uint32_t segments;
  uint32_t resumed;         /* 1 if resumed from a checkpoint */
  uint32_t lanes;           /* lanes per prime of a one-segment job that underfills the GPU (its n-range
                               cut into subranges run side by side); 1 otherwise */
} fs_result;

void fs_default_options(fs_options* o);
/* cl_sieve (cl_sieve.cpp:918-1270): sieve p in [pmin, pmax), n in [nmin, nmax).
 * If factors/cap given, the kept factors (sorted by p, |n|) are copied there
 * (*nfactors = total kept, may exceed cap). */
int fs_sieve(uint64_t pmin, uint64_t pmax, uint32_t nmin, uint32_t nmax, const fs_options* opt,
             fs_result* res, fs_factor* factors, uint64_t cap, uint64_t* nfactors);
/* run_test (cl_sieve.cpp:1273-1391): the five reference ranges; returns the
 * number of passing cases (5 = all). */
int fs_self_test(const fs_options* opt, int* passed_mask);

/* Checkpoint record, byte-identical to workStatus (cl_sieve.h:9-12): 72 bytes,
 * state_sum = wrapping sum of the other nine fields. */
typedef struct {
  uint64_t pmin, pmax, p, checksum, primecount, factorcount, last_trickle, state_sum;
  uint32_t nmin, nmax;
} fs_work_status;
/* read_state (cl_sieve.cpp:132-222): in `dir` ("" = current directory), reads
 * stateA.ckp and stateB.ckp; a file counts when it parses, its ranges equal
 * st->pmin/pmax/nmin/nmax and its state_sum checks ("Cannot parse", "Invalid
 * checkpoint file", "Checksum error" on stderr otherwise).  With two good files
 * the larger p wins.  Returns 1 and fills *st and *which (0 = A, 1 = B) when one
 * is found, 0 when none is. */
int fs_read_checkpoint(const char* dir, fs_work_status* st, int* which);
/* write_state (cl_sieve.cpp:100-129): sets st->state_sum and writes stateA.ckp
 * (which = 0) or stateB.ckp (which = 1); FS_OK or FS_E_IO. */
int fs_write_checkpoint(const char* dir, fs_work_status* st, int which);

/* ---------------------------------------- batched persistent form (many jobs) */
typedef struct {
  uint64_t pmin, pmax;
  uint32_t nmin, nmax;
} fs_job;

typedef struct {
  uint64_t factorcount, primecount, checksum, raw_factors, steps;
  int status;               /* FS_OK or error for this job */
  int path;
} fs_job_result;

typedef struct {
  uint32_t tile_candidates;   /* candidates per tile, a multiple of 256 (default 256) */
  uint32_t persistent_blocks; /* resident workgroups of the persistent sieve kernel (0 = 8 per CU) */
  uint32_t use_graph;         /* 1: the launches of a batch run as one captured hipGraph */
  uint32_t lpt_order;         /* 1: tiles queued longest job first */
  int verify_threads;         /* CPU threads verifying factors */
  uint32_t lanes_per_prime;   /* 1..8: each job's n-range cut into this many subranges (of at least
                                 1024 n) that run on different lanes, each from its own setup (default 1) */
  uint32_t n_chunk;           /* n values per round of the persistent kernel, residues kept in HBM
                                 between rounds (0 = whole range in one round; jobs cut over several
                                 lanes run whole) */
  uint32_t unroll;            /* iterate steps per +-1 branch: 2, 4 or 8 (default and 0: 8) */
} fs_batch_options;

void fs_default_batch_options(fs_batch_options* o);
/* Many sieve jobs (no reference counterpart: the reference runs one job per
 * process, cl_sieve.cpp:918).  One candidate-generation launch for all jobs,
 * then a persistent launch (one per round when n_chunk cuts the n-ranges) whose
 * workgroups pull (job, tile of candidates, n-range) tiles from a device queue
 * and run setup + iterate + checksum fold per candidate.  Results per job are
 * bit-identical to fs_sieve on the job alone, for every option value.
 * factors: optional concatenated kept factors (job order, each sorted);
 * per_job_nf: kept count per job; seconds: wall time of the call. */
int fs_batch(const fs_job* jobs, uint32_t njobs, const fs_batch_options* opt, fs_job_result* out,
             fs_factor* factors, uint64_t cap, uint64_t* per_job_nf, double* seconds);
/* Time the launch shapes on a 1/8 p-range sample of the jobs and write the
 * fastest into *opt: one dimension at a time from *opt, in the order tile size
 * (primes per tile), persistent workgroups, queue order, unroll depth, lanes
 * per prime and n-chunk length. */
int fs_batch_autotune(const fs_job* jobs, uint32_t njobs, fs_batch_options* opt);

/* ------------------------------------------------ VALU issue-rate probe */
/* No reference counterpart: the measured peak the benchmark prices the
 * kernels against.  Lane-operations per second (and the launch's ms) of one
 * instruction kind -- 0 v_mad_u64_u32, 1 v_mul_lo_u32, 2 v_mul_hi_u32,
 * 3 v_fma_f64 -- in `iters` iterations of 8 independent chains per lane, 8 waves
 * on every SIMD, timed over one launch after a warm-up launch. */
int fs_valu_peak(int kind, uint32_t iters, double* lane_ops_per_s, double* ms);

/* ------------------------------------------------------ serving loop */
/* Continuous batching of workunits (no reference counterpart: BOINC hands one
 * workunit to one process).  Admission by device memory; residues stay in HBM
 * between quanta; each result equals fs_sieve on the workunit alone.  Quanta
 * are pipelined one deep: a step enqueues its quantum and then harvests the
 * previous one (verification and filtering on the host) while the new one
 * runs, so a workunit's result is published by the step after its last
 * quantum.  A workunit whose candidates or factors overflowed a buffer starts
 * again with a larger one (at most 3 times; then FS_E_OVERFLOW). */
typedef void* fs_server;
typedef struct {
  uint64_t mem_budget;      /* device bytes the server may hold for admitted work (0 = 50% free) */
  uint32_t max_slots;       /* concurrently resident workunits */
  uint32_t n_chunk;         /* n values per scheduling quantum (default 49,152) */
  int verify_threads;
  uint32_t use_graph;       /* 1 (default): each quantum's launches replay a captured hipGraph */
} fs_server_options;

void fs_default_server_options(fs_server_options* o);
int fs_server_create(const fs_server_options* o, fs_server* srv);
/* Submit a workunit; returns its id (>= 0) or FS_E_BUSY/FS_E_ARG.  Never waits
 * for device work: the incoming queue has its own lock, held only to append. */
int64_t fs_server_submit(fs_server srv, const fs_job* job);
/* Advance the device work by one scheduling quantum (admit queued work when
 * memory and slots free up, enqueue a quantum that advances every active
 * workunit by n_chunk n-values, harvest the previous quantum).  Returns the
 * number of workunits finished by the harvest (>= 0) or an error.  Steps of one
 * server run one at a time. */
int fs_server_step(fs_server srv);
/* Fetch a finished workunit's result (1 = done and copied, 0 = not yet). */
int fs_server_poll(fs_server srv, int64_t id, fs_job_result* out, fs_factor* factors, uint64_t cap,
                   uint64_t* nfactors);
/* queued: submitted, not yet admitted; active: admitted, result not yet
 * published; done: results published; bytes_in_use: candidate regions held. */
int fs_server_stats(fs_server srv, uint32_t* queued, uint32_t* active, uint32_t* done,
                    uint64_t* bytes_in_use);
int fs_server_destroy(fs_server srv);

#ifdef __cplusplus
}
#endif
#endif /* FACTSIEVE_H */
