"""The rw / ks / mrw tools on the GPU: each runs its App on a fixture graph written to files and
reports the sampled-edge count the oracle predicts for the same samples and Philox counters."""
import json
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO, load_case

pytestmark = pytest.mark.gpu

BIN = os.path.join(REPO, "grass_amd", "bin")


def _files(tmp_path, fmt):
    c = load_case("rmat12")
    prefix = str(tmp_path / f"rmat12_{fmt}.cgr")
    with open(prefix + ".graph", "wb") as f:
        f.write(c[f"{fmt}_graph"].tobytes())
    with open(prefix + ".offset", "wb") as f:
        f.write(c[f"{fmt}_offset"].tobytes())
    return prefix, c


def _tool(name, *args):
    print(f"[cli] {name} {' '.join(args)}", flush=True)
    p = subprocess.run([os.path.join(BIN, name), *args], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    return p


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_rw_tool(tmp_path, fmt):
    prefix, c = _files(tmp_path, fmt)
    out = json.loads(_tool("rw", "-g", prefix, "--format", fmt, "--samples", "5000", "--batch", "2048",
                           "-n", "2", "--json").stdout)
    assert out["samples"] == 5000 and out["runs"] == 2
    assert out["sampled_edges"] == 2 * 5000 * 80
    p = _tool("rw", "-g", prefix, "--format", fmt, "--samples", "100")
    for line in ("Transit Parallel: End to end time", "EndToEndTime:", "transferDataTime:", "runSamplingTime:"):
        assert line in p.stdout


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_ks_tool(tmp_path, fmt, oracle_mod):
    prefix, c = _files(tmp_path, fmt)
    n = int(c["n"])
    out = json.loads(_tool("ks", "-g", prefix, "--format", fmt, "--samples", "3000", "--batch", "1000",
                           "--seed", "7", "--json").stdout)
    og = oracle_mod.Graph(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    seeds = [np.array([i % n], dtype=np.uint32) for i in range(3000)]
    exp = og.khop(seeds, [[10, 25]] * 3000, seed=7, instances=list(range(3000)))
    assert out["sampled_edges"] == len(exp["dst"])


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_mrw_tool(tmp_path, fmt):
    prefix, c = _files(tmp_path, fmt)
    out = json.loads(_tool("mrw", "-g", prefix, "--format", fmt, "--samples", "300", "--json").stdout)
    # rmat12 has no vertex without edges, so every instance runs all 100 steps
    assert out["sampled_edges"] == 300 * 100


@pytest.mark.parametrize("directed", ["u", "d"])
def test_graph_sort_and_compressor_device_mode(tmp_path, directed):
    """`--device 0` on the reference-compatible graph_sort and compressor tools writes the same
    files as their host mode."""
    from grass_amd import ingest
    src, dst = ingest.rmat_edges(14, 16, seed=14, threads=4)
    raw = tmp_path / "raw.txt"
    with open(raw, "w") as f:
        f.write("# header line\n")
        f.write("\n".join(f"{u} {v}" for u, v in zip(src.tolist(), dst.tolist())))
    outs = {}
    for mode in ("host", "dev"):
        extra = ["--device", "0"] if mode == "dev" else []
        srt = tmp_path / f"sorted_{mode}.txt"
        _tool("graph_sort", str(raw), str(srt), directed, "1", *extra)
        outs[mode] = srt.read_bytes()
        for fmt in ("le", "gt"):
            _tool("compressor", "--format", fmt, *extra, str(srt), str(tmp_path / f"{fmt}_{mode}"))
    assert outs["host"] == outs["dev"]
    for fmt in ("le", "gt"):
        for ext in (".graph", ".offset"):
            assert (tmp_path / f"{fmt}_host{ext}").read_bytes() == (tmp_path / f"{fmt}_dev{ext}").read_bytes()
