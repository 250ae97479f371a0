"""The oracle's own checks again under UndefinedBehaviorSanitizer (CPU).

The reference runs its tests with Rust's overflow checks on, where an
arithmetic overflow panics.  The oracle's UBSan build
(oracle/liboxav1_oracle_ubsan.so, -fsanitize=undefined with no recovery)
aborts on a signed overflow, an out-of-range shift or a misaligned or
out-of-bounds access it can see; the known-answer, pin, self-check, edge
and election suites run against it in a child process, which must exit
cleanly."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SUITES = ["test_oracle_kat.py", "test_oracle_pins.py", "test_oracle_selfcheck.py",
          "test_oracle_lf_edges.py", "test_oracle_elect.py", "test_oracle_lr_elect.py", "test_tables.py"]


def test_oracle_suites_under_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "liboxav1_oracle_ubsan.so"],
                   check=True)
    env = dict(os.environ, OXAV1_ORACLE_SAN="ubsan")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", "-s",
                        "-m", "not gpu", "-k", "not mutant and not library_tables"]
                       + [os.path.join(ROOT, "tests", s) for s in SUITES],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-3000:]
    assert r.returncode == 0, tail
    assert "runtime error" not in r.stdout + r.stderr, tail
