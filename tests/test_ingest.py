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
