"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
tests/cpp/oracle_sanitize (built by `make -C oracle sanitize`, which
__graft_entry__.build() runs) calls every entry point of oracle/oracle.cpp and
oracle/cnn_oracle.cpp on frames from 1x1 to 64x48, all four Bayer phases,
both precisions, one and several OpenMP threads.  Any out-of-bounds access,
use after free, leak, signed overflow or invalid shift stops it with a
report.  A planted off-by-one in the box blur's clamp (row[clamp(i, 0, w)])
is reported as a heap-buffer-overflow by the same run."""
import os
import subprocess

from conftest import ROOT

EXE = os.path.join(ROOT, "tests", "cpp", "oracle_sanitize")


def test_oracle_clean_under_asan_ubsan():
    if not os.path.exists(EXE):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True, timeout=600)
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=600, env=env)
    report = r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "ERROR: AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert "cnn, auto-wb ok" in r.stdout
