"""The host code the product shares with the oracle (initial partitioning
with its worker threads, the km1 rebalancer) and the oracle's pipelines,
built with -fsanitize=address,undefined and with -fsanitize=thread
(oracle/Makefile, target `sanitizers`) and run by oracle/sanitize_check.cpp
on seeded random hypergraphs.  CPU only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitizers"], capture_output=True, text=True)
    if r.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + r.stdout + r.stderr)
    return os.path.join(ORACLE, "_build")


def run(path, *args):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", TSAN_OPTIONS="halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", OMP_NUM_THREADS="8")
    r = subprocess.run([path, *args], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout[-3000:] + r.stderr[-6000:]


def test_address_and_undefined_behaviour_sanitizers(built):
    run(os.path.join(built, "san_address"))


def test_thread_sanitizer_on_the_threaded_host_code(built):
    run(os.path.join(built, "san_thread"), "threads")
