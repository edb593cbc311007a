"""The host code under ThreadSanitizer.

factsieve_amd.build.build_tsan() compiles the library with -fsanitize=thread
on the host side only (the gfx950 code is the ordinary build) and links the
host-only driver tests/tsan_verify_driver.cpp against it; the driver calls the
factor verifier from six threads at once, and builds setup programs and power
tables on two more.  The verifier's worker threads enter and leave each
parallel region through atomics, and the prime list behind the programs is a
shared cache, so a data race in either would be reported here."""
import json
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_verifier_threads_are_race_free_under_tsan(tmp_path):
    from factsieve_amd import build as B
    drv = B.build_tsan(os.path.join(ROOT, "tests", "tsan_verify_driver.cpp"))
    factors = json.load(open(os.path.join(ROOT, "tests", "golden", "c1.json")))["factors"]
    log = str(tmp_path / "tsan")
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1 exitcode=66 log_path=" + log)
    r = subprocess.run([drv], input="".join(f"{p} {nc}\n" for p, nc in factors), env=env, capture_output=True,
                       text=True, timeout=600)
    reports = "".join(open(os.path.join(tmp_path, f)).read()[:4000] for f in sorted(os.listdir(tmp_path))
                      if f.startswith("tsan"))
    assert r.returncode == 0 and not reports, r.stdout[-2000:] + r.stderr[-2000:] + reports
    assert "wrong answers: 0" in r.stdout
