"""The rw / ks / mrw command-line tools: the reference's appMain behaviour (LE/src/main.cu:5-40)
without a GPU -- usage text and exit status 0 with no option or without -g, a load failure
reported with exit status 0 -- and rejected options."""
import os
import subprocess

import pytest

from conftest import REPO

BIN = os.path.join(REPO, "grass_amd", "bin")
TOOLS = ("rw", "ks", "mrw")


def _run(tool, *args):
    exe = os.path.join(BIN, tool)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (make -C grass_amd)")
    return subprocess.run([exe, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("tool", TOOLS)
def test_usage_without_options(tool):
    p = _run(tool)
    assert p.returncode == 0
    assert "-g --graph-file" in p.stdout and "-n --nruns" in p.stdout


@pytest.mark.parametrize("tool", TOOLS)
def test_usage_without_graph(tool):
    p = _run(tool, "-n", "3")
    assert p.returncode == 0 and "usage" in p.stdout


@pytest.mark.parametrize("tool", TOOLS)
def test_missing_graph_reported(tool, tmp_path):
    p = _run(tool, "-g", str(tmp_path / "absent.cgr"))
    assert p.returncode == 0
    assert "Open graph file failed" in p.stderr


def test_bad_options():
    assert _run("rw", "--format", "xx", "-g", "a").returncode == 2
    assert _run("rw", "--bogus").returncode == 2
    assert _run("rw", "-g").returncode == 2
