"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(oracle/Makefile target `sanitized`): a driver linked with the oracle's
sources walks every label fixture, singular and regular pressure systems, a
dam break with a falling body stepped stage by stage and through run(), file
output and error paths; it must finish without a sanitizer report (which
aborts the run).  CPU only."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_build", "oracle_sanitized")


def test_oracle_clean_under_sanitizers(tmp_path):
    r = subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "sanitized"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([EXE, os.path.join(ROOT, "tests", "golden", "inputdata"), str(tmp_path)], capture_output=True,
                       text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "OK" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
