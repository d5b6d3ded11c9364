"""The oracle under AddressSanitizer + UndefinedBehaviorSanitizer (oracle/selftest.c, built by
`make -C oracle sanitize`): decode, random access, walks, k-hop, MDRW and corrupted inputs on
every golden fixture, in both codecs, with no sanitizer report."""
import os
import subprocess

import pytest

from conftest import CASES, REPO, load_case

SAN = os.path.join(REPO, "oracle", "_san", "selftest")


@pytest.fixture(scope="module")
def selftest():
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "sanitize"],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0 or not os.path.exists(SAN):
        pytest.skip(f"sanitizer build unavailable: {r.stderr[-500:]}")
    return SAN


@pytest.mark.parametrize("fmt", ["le", "gt"])
@pytest.mark.parametrize("case", CASES)
def test_oracle_clean_under_sanitizers(selftest, tmp_path, case, fmt):
    c = load_case(case)
    prefix = tmp_path / case
    (tmp_path / f"{case}.graph").write_bytes(c[f"{fmt}_graph"].tobytes())
    (tmp_path / f"{case}.offset").write_bytes(c[f"{fmt}_offset"].tobytes())
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", OMP_NUM_THREADS="2")
    p = subprocess.run([selftest, str(prefix), fmt], capture_output=True, text=True, timeout=300, env=env)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "runtime error" not in p.stderr and "ERROR: AddressSanitizer" not in p.stderr
    assert "selftest ok" in p.stdout
