"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer.

factsieve_amd.build.build_sanitized() compiles the library and the command line
with -fsanitize=address,undefined on the host side only (the gfx950 code is the
ordinary build), and this test runs the CPU host suite (tests/test_host.py:
number parsing, power tables, setup programs, the host execution of every
kernel loop, CPU verification, factor sorting and filtering, trickle-up text,
init_data.xml, the command line) against it in a child process.  Any sanitizer
report ends that process and fails the test with the report."""
import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_suite_is_clean_under_asan_and_ubsan(tmp_path):
    from factsieve_amd import build as B
    lib = B.build_sanitized()
    log = str(tmp_path / "san")
    env = dict(os.environ)
    env.update(
        FACTSIEVE_LIB=lib,
        # the runtime first, ahead of anything the environment already preloads
        LD_PRELOAD=":".join(x for x in (B.sanitizer_runtime(), os.environ.get("LD_PRELOAD", "")) if x),
        # Python's own allocations are not freed at exit: leak checking is off
        ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:log_path=" + log,
        UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:log_path=" + log,
    )
    r = subprocess.run([sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_host.py"), "-q",
                        "-m", "not gpu", "-p", "no:cacheprovider"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    reports = "".join(open(f).read()[:4000] for f in sorted(glob.glob(log + "*")))
    assert r.returncode == 0 and not reports, r.stdout[-4000:] + r.stderr[-2000:] + reports
    assert " passed" in r.stdout
