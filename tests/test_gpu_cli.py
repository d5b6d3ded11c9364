"""The rw / ks / mrw tools on the GPU: each runs its App on a fixture graph written to files, and
the samples it writes with --out equal the oracle's for the same starts and Philox counters
(sample i: start i % n; walks keyed as walker i - s0 of instance s0 for the batch [s0, s1) that
ran it, k-hop and MultiRW as instance i; run r with seed + r)."""
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


def _u64(buf, pos):
    return int(np.frombuffer(buf, np.uint64, 1, pos)[0]), pos + 8


def _u32s(buf, pos, count):
    return np.frombuffer(buf, np.uint32, count, pos), pos + 4 * count


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_rw_tool(tmp_path, fmt, oracle_mod):
    prefix, c = _files(tmp_path, fmt)
    n = int(c["n"])
    samples, batch, runs = 5000, 2048, 2
    outf = tmp_path / "walks.bin"
    out = json.loads(_tool("rw", "-g", prefix, "--format", fmt, "--samples", str(samples), "--batch", str(batch),
                           "-n", str(runs), "--json", "--out", str(outf)).stdout)
    assert out["samples"] == samples and out["runs"] == runs
    assert out["sampled_edges"] == runs * samples * 80
    walks = np.fromfile(outf, dtype=np.uint32).reshape(runs, samples, 81)
    og = oracle_mod.Graph(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    for r in range(runs):
        for s0 in range(0, samples, batch):
            s1 = min(s0 + batch, samples)
            starts = np.arange(s0, s1, dtype=np.uint64) % n
            exp, _ = og.random_walk(s1 - s0, 80, 42 + r, instance=s0, starts=starts.astype(np.uint32))
            np.testing.assert_array_equal(walks[r, s0:s1], exp)
    p = _tool("rw", "-g", prefix, "--format", fmt, "--samples", "100")
    for line in ("Transit Parallel: End to end time", "EndToEndTime:", "transferDataTime:", "runSamplingTime:"):
        assert line in p.stdout


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_ks_tool(tmp_path, fmt, oracle_mod):
    prefix, c = _files(tmp_path, fmt)
    n = int(c["n"])
    samples = 3000
    outf = tmp_path / "khop.bin"
    out = json.loads(_tool("ks", "-g", prefix, "--format", fmt, "--samples", str(samples), "--batch", "1000",
                           "--seed", "7", "--json", "--out", str(outf)).stdout)
    og = oracle_mod.Graph(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    seeds = [np.array([i % n], dtype=np.uint32) for i in range(samples)]
    exp = og.khop(seeds, [[10, 25]] * samples, seed=7, instances=list(range(samples)))
    assert out["sampled_edges"] == len(exp["dst"])
    buf, pos = outf.read_bytes(), 0
    for i in range(samples):
        e = int(exp["edge_off"][i])
        for h in range(2):
            cnt, pos = _u64(buf, pos)
            assert cnt == int(exp["hop_count"][i, h])
            src, pos = _u32s(buf, pos, cnt)
            dst, pos = _u32s(buf, pos, cnt)
            np.testing.assert_array_equal(src, exp["src"][e:e + cnt])
            np.testing.assert_array_equal(dst, exp["dst"][e:e + cnt])
            e += cnt
        assert e == int(exp["edge_off"][i + 1])
    assert pos == len(buf)


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_mrw_tool(tmp_path, fmt, oracle_mod):
    prefix, c = _files(tmp_path, fmt)
    samples = 300
    outf = tmp_path / "mrw.bin"
    out = json.loads(_tool("mrw", "-g", prefix, "--format", fmt, "--samples", str(samples), "--batch", "128",
                           "--json", "--out", str(outf)).stdout)
    # rmat12 has no vertex without edges, so every instance runs all 100 steps
    assert out["sampled_edges"] == samples * 100
    og = oracle_mod.Graph(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    exp = og.mdrw([80] * samples, 180, 42, instances=list(range(samples)))
    buf, pos = outf.read_bytes(), 0
    for i in range(samples):
        for name, pairs in (("step", True), ("vert", False), ("ind", True)):
            cnt, pos = _u64(buf, pos)
            a, b = int(exp[f"{name}_off"][i]), int(exp[f"{name}_off"][i + 1])
            assert cnt == b - a
            if pairs:
                src, pos = _u32s(buf, pos, cnt)
                dst, pos = _u32s(buf, pos, cnt)
                key = "step" if name == "step" else "ind"
                np.testing.assert_array_equal(src, exp[f"{key}_src"][a:b])
                np.testing.assert_array_equal(dst, exp[f"{key}_dst"][a:b])
            else:
                verts, pos = _u32s(buf, pos, cnt)
                np.testing.assert_array_equal(verts, exp["verts"][a:b])
    assert pos == len(buf)


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
