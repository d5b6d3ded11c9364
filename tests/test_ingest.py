"""The product's ingest (graph_sort + LE/GT compressors, libgrass_amd.so host code) must
reproduce the reference tools byte for byte.  CPU only."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import CASES, REPO, csr_of, load_case

ingest = pytest.importorskip("grass_amd.ingest")
REF_BIN = os.path.join(REPO, "oracle", "_ref")


@pytest.mark.parametrize("case", CASES)
def test_compressors_byte_identical(case):
    c = load_case(case)
    csr = ingest.csr_from_sorted_edges(c["src"], c["dst"])
    assert csr.n == int(c["n"])
    for fmt in ("le", "gt"):
        gb, ob = csr.compress(fmt, threads=4)
        assert gb == c[f"{fmt}_graph"].tobytes(), f"{case} {fmt} .graph"
        assert ob == c[f"{fmt}_offset"].tobytes(), f"{case} {fmt} .offset"


@pytest.mark.parametrize("case", ["gs_small_u", "gs_small_d", "gs_rmat12_u"])
def test_graph_sort_matches_reference(case):
    c = load_case(case)
    csr = ingest.graph_sort(c["raw_u"].astype(np.uint32), c["raw_v"].astype(np.uint32),
                            directed=bool(c["directed"]), threads=4)
    rowptr, col = csr.arrays()
    n = int(max(c["out_u"].max(), c["out_v"].max())) + 1
    assert csr.n == n
    src = np.repeat(np.arange(n), np.diff(rowptr.astype(np.int64)))
    np.testing.assert_array_equal(src, c["out_u"])
    np.testing.assert_array_equal(col, c["out_v"])


def test_unsorted_edges_rejected():
    with pytest.raises(Exception):
        ingest.csr_from_sorted_edges(np.array([2, 1], np.uint32), np.array([0, 0], np.uint32))


def test_rmat_quadrants():
    src, dst = ingest.rmat_edges(12, 16, seed=42, threads=4)
    assert len(src) == 16 << 12 and src.max() < 4096 and dst.max() < 4096
    top_u, top_v = src >> 11, dst >> 11
    q = np.bincount(top_u * 2 + top_v, minlength=4) / len(src)
    np.testing.assert_allclose(q, [0.57, 0.19, 0.19, 0.05], atol=0.01)
    s2, d2 = ingest.rmat_edges(12, 16, seed=42, threads=1)
    np.testing.assert_array_equal(src, s2)
    np.testing.assert_array_equal(dst, d2)


def test_roundtrip_through_oracle(oracle_mod):
    src, dst = ingest.rmat_edges(15, 16, seed=7, threads=4)
    csr = ingest.graph_sort(src, dst, threads=4)
    rowptr, col = csr.arrays()
    for fmt in ("le", "gt"):
        gb, ob = csr.compress(fmt, threads=4)
        g = oracle_mod.Graph(gb, ob, fmt)
        rp, cl = g.decode_csr()
        np.testing.assert_array_equal(rp, rowptr)
        np.testing.assert_array_equal(cl, col)


@pytest.mark.skipif(not os.path.exists(os.path.join(REF_BIN, "le_compressor")),
                    reason="reference tools not built here (make -C oracle ref)")
def test_against_reference_tools_rmat14(tmp_path):
    """R-MAT s14 through the compiled reference graph_sort and compressors."""
    src, dst = ingest.rmat_edges(14, 16, seed=3, threads=4)
    raw = tmp_path / "raw.txt"
    np.savetxt(raw, np.stack([src, dst], 1), fmt="%d")
    sorted_txt = tmp_path / "sorted.txt"
    subprocess.run([os.path.join(REF_BIN, "graph_sort"), str(raw), str(sorted_txt), "u"],
                   check=True, stdout=subprocess.DEVNULL)
    csr = ingest.graph_sort(src, dst, threads=4)
    ref = np.loadtxt(sorted_txt, dtype=np.int64, ndmin=2)
    rowptr, col = csr.arrays()
    np.testing.assert_array_equal(np.repeat(np.arange(csr.n), np.diff(rowptr.astype(np.int64))),
                                  ref[:, 0])
    np.testing.assert_array_equal(col, ref[:, 1])
    for fmt in ("le", "gt"):
        prefix = tmp_path / fmt
        subprocess.run([os.path.join(REF_BIN, f"{fmt}_compressor"), str(sorted_txt), str(prefix)],
                       check=True, stdout=subprocess.DEVNULL)
        gb, ob = csr.compress(fmt, threads=4)
        assert gb == (tmp_path / f"{fmt}.graph").read_bytes()
        assert ob == (tmp_path / f"{fmt}.offset").read_bytes()


BIN = os.path.join(REPO, "grass_amd", "bin")


def _edge_text(path, rng, header=0):
    """Directed edges over sparse large ids, with self loops, duplicates and a header."""
    ids = rng.choice(10**9, size=400, replace=False)
    u, v = ids[rng.integers(0, 400, 6000)], ids[rng.integers(0, 400, 6000)]
    u[:50] = v[:50]  # self loops
    lines = [f"# header line {i}" for i in range(header)] + [f"{a} {b}" for a, b in zip(u, v)]
    lines += [f"{u[7]} {v[7]}"] * 5  # duplicates
    path.write_text("\n".join(lines) + "\n")


@pytest.mark.parametrize("mode", ["u", "d"])
def test_graph_sort_tool(tmp_path, mode):
    raw, out = tmp_path / "raw.txt", tmp_path / "out.txt"
    _edge_text(raw, np.random.default_rng(1), header=2)
    p = subprocess.run([os.path.join(BIN, "graph_sort"), str(raw), str(out), mode, "2"],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    assert "Skip 2 Lines" in p.stdout and "Graph has" in p.stdout
    if os.path.exists(os.path.join(REF_BIN, "graph_sort")):
        ref = tmp_path / "ref.txt"
        r = subprocess.run([os.path.join(REF_BIN, "graph_sort"), str(raw), str(ref), mode, "2"],
                           capture_output=True, text=True, timeout=60)
        assert r.returncode == 0
        assert out.read_bytes() == ref.read_bytes()
        assert p.stdout.splitlines()[-1] == r.stdout.splitlines()[-1]  # "Graph has ..."


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_compressor_tool(tmp_path, fmt):
    raw, srt = tmp_path / "raw.txt", tmp_path / "sorted.txt"
    _edge_text(raw, np.random.default_rng(2))
    assert subprocess.run([os.path.join(BIN, "graph_sort"), str(raw), str(srt), "u"],
                          capture_output=True, timeout=60).returncode == 0
    prefix = tmp_path / "g"
    p = subprocess.run([os.path.join(BIN, "compressor"), "--format", fmt, str(srt), str(prefix)],
                       capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stderr
    assert "graph loaded." in p.stdout and "CGR generation completed." in p.stdout
    if os.path.exists(os.path.join(REF_BIN, f"{fmt}_compressor")):
        rp = tmp_path / "r"
        assert subprocess.run([os.path.join(REF_BIN, f"{fmt}_compressor"), str(srt), str(rp)],
                              capture_output=True, timeout=60).returncode == 0
        for ext in (".graph", ".offset"):
            assert (tmp_path / f"g{ext}").read_bytes() == (tmp_path / f"r{ext}").read_bytes()


def test_ingest_tools_reject_bad_input(tmp_path):
    assert subprocess.run([os.path.join(BIN, "graph_sort"), "a", "b"], capture_output=True).returncode == 1
    unsorted = tmp_path / "uns.txt"
    unsorted.write_text("3 1\n1 2\n")
    p = subprocess.run([os.path.join(BIN, "compressor"), str(unsorted), str(tmp_path / "x")],
                       capture_output=True, text=True)
    assert p.returncode == 1 and "sorted" in p.stderr
    p = subprocess.run([os.path.join(BIN, "compressor"), str(unsorted), str(tmp_path / "x"), "head_bit"],
                       capture_output=True, text=True)
    assert p.returncode == 1
