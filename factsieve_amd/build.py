"""Build the gfx950 library, the command-line executable and (as test
infrastructure) the CPU oracle.

    python -m factsieve_amd.build        # or __graft_entry__.build()

Outputs, all in-tree so they travel with the repository to a GPU machine:
    factsieve_amd/libfactsieve.so   C ABI of include/factsieve.h (HIP kernels + host driver)
    factsieve_amd/factsieve         drop-in command line (-n -N -p -P -v -s)
    oracle/liboracle.so             CPU oracle (tests and the CPU baseline only)
    oracle/_ref/libfsref.so         reference verifyprime.c + putil.c (only where
                                    the reference checkout is present)
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libfactsieve.so")
EXE = os.path.join(HERE, "factsieve")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("FS_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: never emit v_pk_{add,mul,fma}_f32 (the host pass warns that
# it does not know the feature; that warning is harmless).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
SOURCES = ["fs_device.hip", "fs_batch.hip", "fs_peak.hip", "fs_host.cpp", "fs_hostemu.cpp", "fs_batch_host.cpp",
           "fs_server.cpp"]


def _run(cmd, cwd=None):
    res = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l)
    if res.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build_library(verbose=False):
    objs = []
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    for src in SOURCES:
        path = os.path.join(CSRC, src)
        if not os.path.exists(path):
            continue
        obj = os.path.join(HERE, "_build", os.path.splitext(src)[0] + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        out = _run([HIPCC] + HIP_FLAGS + lang + ["-c", path, "-o", obj])
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, LIB)
    _run([HIPCC, "-O2", "-std=c++17", "-o", EXE, os.path.join(CSRC, "fs_main.cpp"), "-L", HERE,
          "-lfactsieve", "-Wl,-rpath,$ORIGIN"])
    return LIB


# AddressSanitizer and UndefinedBehaviorSanitizer on the host code only: each
# -fsanitize= goes after -Xarch_host, so the gfx950 code is built as usual.
SAN_DIR = os.path.join(HERE, "_build", "sanitized")
SAN_HOST = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
            "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]


def sanitizer_runtime():
    """Path of the clang AddressSanitizer runtime (it must be the first library
    of a process that loads the sanitized build, through LD_PRELOAD)."""
    llvm = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "llvm")
    found = sorted(glob.glob(os.path.join(llvm, "lib", "clang", "*", "lib", "*", "libclang_rt.asan-x86_64.so")))
    if not found:
        raise RuntimeError("no libclang_rt.asan-x86_64.so under " + llvm)
    return found[-1]


def build_sanitized(jobs=8):
    """The library and the command line with the host code under ASan + UBSan,
    into factsieve_amd/_build/sanitized/ (load it with FACTSIEVE_LIB and the
    runtime of sanitizer_runtime() in LD_PRELOAD).  Returns the library path."""
    os.makedirs(SAN_DIR, exist_ok=True)
    flags = [f for f in HIP_FLAGS if f != "-O3"] + ["-O1", "-g"] + SAN_HOST

    def one(src):
        obj = os.path.join(SAN_DIR, os.path.splitext(src)[0] + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        _run([HIPCC] + flags + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(one, SOURCES))
    rt_dir = os.path.dirname(sanitizer_runtime())
    lib = os.path.join(SAN_DIR, "libfactsieve.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-shared-libsan", "-o", lib + ".tmp"] + objs
         + ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-Wl,-rpath," + rt_dir,
            "-lpthread"])
    os.replace(lib + ".tmp", lib)
    _run([HIPCC, "-O1", "-g", "-std=c++17", "-o", os.path.join(SAN_DIR, "factsieve"), os.path.join(CSRC, "fs_main.cpp"),
          "-L", SAN_DIR, "-lfactsieve", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + rt_dir])
    return lib


# ThreadSanitizer on the host code only (again after -Xarch_host): the worker
# threads of the factor verifier, the buffer pools and the caches are shared by
# every host thread that calls the library.
TSAN_DIR = os.path.join(HERE, "_build", "tsan")
CLANGXX = os.path.join(os.path.dirname(os.path.dirname(os.path.realpath(HIPCC))), "llvm", "bin", "clang++")


def build_tsan(driver_src, jobs=8):
    """The library with its host code under ThreadSanitizer, into
    factsieve_amd/_build/tsan/, and a host-only driver program (driver_src, C++
    that calls the C ABI) linked against it.  Returns the driver's path."""
    os.makedirs(TSAN_DIR, exist_ok=True)
    flags = [f for f in HIP_FLAGS if f != "-O3"] + ["-O1", "-g", "-Xarch_host", "-fsanitize=thread"]

    def one(src):
        obj = os.path.join(TSAN_DIR, os.path.splitext(src)[0] + ".o")
        lang = [] if src.endswith(".hip") else ["-x", "hip"]
        _run([HIPCC] + flags + lang + ["-c", os.path.join(CSRC, src), "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(one, SOURCES))
    rt = sorted(glob.glob(os.path.join(os.path.dirname(os.path.dirname(CLANGXX)), "lib", "clang", "*", "lib", "*",
                                       "libclang_rt.tsan-x86_64.so")))
    if not rt:
        raise RuntimeError("no libclang_rt.tsan-x86_64.so beside " + CLANGXX)
    rt_dir = os.path.dirname(rt[-1])
    lib = os.path.join(TSAN_DIR, "libfactsieve.so")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-shared-libsan", "-o", lib + ".tmp"] + objs
         + ["-Xarch_host", "-fsanitize=thread", "-Wl,-rpath," + rt_dir, "-lpthread"])
    os.replace(lib + ".tmp", lib)
    drv = os.path.join(TSAN_DIR, "driver")
    _run([CLANGXX, "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-shared-libsan", "-I", os.path.join(ROOT, "include"),
          driver_src, "-L", TSAN_DIR, "-lfactsieve", "-Wl,-rpath," + TSAN_DIR, "-Wl,-rpath," + rt_dir, "-lpthread",
          "-o", drv])
    return drv


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "all"])


def build(verbose=False):
    build_library(verbose)
    build_oracle()
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    print("built", LIB)
