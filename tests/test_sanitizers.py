"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU
only): builds the host tree builder and the KDTree3f / MCAP codec
(`make -C kd_slam2_amd/csrc sanitize-host`) and the oracle
(`make -C oracle sanitize`) with -fsanitize=address,undefined, then runs
tests/sanitize_driver.py in a child process with the ASan runtime preloaded:
the tree builder on the shared test clouds and degenerate ones, the CDR
decoder on a seeded corpus of 3000 mutated messages, the MCAP reader on
300 mutated files, and the oracle's alignments.  Any report fails the run."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _runtime(name):
    r = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_host_code_under_asan_and_ubsan():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("the g++ AddressSanitizer runtime is not installed")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "kd_slam2_amd", "csrc"), "sanitize-host"], check=True)
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    host = os.path.join(ROOT, "kd_slam2_amd", "csrc", "_build", "libkdhost_san.so")
    orc = os.path.join(ROOT, "oracle", "liboracle_san.so")
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "tests.sanitize_driver", host, orc], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=900)
    print(r.stdout[-3000:])
    assert r.returncode == 0, r.stderr[-6000:]
    assert "sanitizers: clean" in r.stdout
    for err in ("AddressSanitizer", "runtime error"):
        assert err not in r.stderr, r.stderr[-6000:]
