"""The threaded host ingest (graph_sort, both compressors, the edge-list reader and writer) under
ThreadSanitizer and under AddressSanitizer + UndefinedBehaviorSanitizer, built by
`make -C grass_amd sanitize` from grass_amd/csrc/ingest.cpp (host code only) and driven by
tests/san/ingest_selftest.cpp: every call with 4 threads must give the 1-thread output, with no
sanitizer report."""
import os
import subprocess

import pytest

from conftest import REPO

SAN = os.path.join(REPO, "grass_amd", "_san")


@pytest.fixture(scope="module")
def built():
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "grass_amd"), "sanitize"],
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        pytest.skip(f"sanitizer build unavailable: {r.stderr[-500:]}")
    return SAN


@pytest.mark.parametrize("kind", ["tsan", "asan"])
def test_ingest_clean_under_sanitizers(built, tmp_path, kind):
    env = dict(os.environ, TSAN_OPTIONS="halt_on_error=1", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1")
    p = subprocess.run([os.path.join(built, f"ingest_{kind}"), "15", str(tmp_path)], capture_output=True,
                       text=True, timeout=300, env=env)
    assert p.returncode == 0, p.stdout + p.stderr[-4000:]
    for bad in ("WARNING: ThreadSanitizer", "ERROR: AddressSanitizer", "runtime error"):
        assert bad not in p.stderr
    assert "ingest selftest ok" in p.stdout
