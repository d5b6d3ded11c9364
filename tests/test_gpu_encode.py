"""The device-side LE / GT encoder (grass_compress_device) must write the same .graph and .offset
bytes as the reference compressors: the golden fixtures the compiled reference tools wrote, and the
host compressor (pinned to those tools by test_ingest) on R-MAT and on ragged synthetic CSRs."""
import numpy as np
import pytest

from conftest import CASES, csr_of, load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    import grass_amd
    from grass_amd import ingest
    assert torch.cuda.is_available()
    grass_amd.lib()
    return torch, ingest


def _dev(torch, rowptr, col):
    rp = torch.from_numpy(rowptr.astype(np.int64)).cuda()
    cl = torch.from_numpy(np.ascontiguousarray(col, dtype=np.uint32).view(np.int32)).cuda()
    return rp, cl


def _check(torch, ingest, rowptr, col, fmt, expect, stream=None):
    rp, cl = _dev(torch, rowptr, col)
    if stream is not None:
        stream.wait_stream(torch.cuda.current_stream())
    gb, ob = ingest.compress_device(rp, cl, fmt, stream=stream)
    eg, eo = expect
    assert len(gb) == len(eg) and len(ob) == len(eo)
    bad = np.flatnonzero(np.frombuffer(gb, np.uint8) != np.frombuffer(eg, np.uint8))
    assert bad.size == 0, f"{fmt} .graph differs at bytes {bad[:4]}"
    assert ob == eo, f"{fmt} .offset differs"


@pytest.mark.parametrize("fmt", ["le", "gt"])
@pytest.mark.parametrize("case", CASES)
def test_encode_golden(env, case, fmt):
    print(f"[encode golden] case={case} fmt={fmt}", flush=True)
    torch, ingest = env
    c = load_case(case)
    rowptr, col = csr_of(c)
    _check(torch, ingest, rowptr, col, fmt, (c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes()))


def _ragged_csr(seed, n, big):
    """Rows of degree 0, 1, 2, 3, a few dozen and up to `big` (sorted, unique, self loops allowed),
    neighbour ids anywhere in [0, n) so rock, head and sea fields take every width."""
    rng = np.random.default_rng(seed)
    deg = rng.choice([0, 1, 2, 3, 5, 17, 64, 65, 129, big], size=n,
                     p=[.1, .15, .1, .1, .1, .15, .1, .1, .09, .01]).astype(np.int64)
    deg[n - 1] = 1  # a self loop on the last vertex, so an edge-list reader finds the same n
    rowptr = np.zeros(n + 1, np.uint64)
    rowptr[1:] = np.cumsum(deg)
    cols = []
    for v, d in enumerate(deg):
        if d == 0:
            continue
        if v == n - 1:
            cols.append(np.array([v], np.uint32))
            continue
        if v % 3 == 0:  # clustered around v: small rocks, small seas
            lo = max(0, v - 4 * int(d))
            pool = np.arange(lo, min(n, lo + 8 * int(d) + 8))
        else:
            pool = None
        row = rng.choice(pool if pool is not None else n, size=int(d), replace=False)
        cols.append(np.sort(row).astype(np.uint32))
    col = np.concatenate(cols) if cols else np.zeros(0, np.uint32)
    return rowptr, col


@pytest.mark.parametrize("fmt", ["le", "gt"])
@pytest.mark.parametrize("seed,n,big", [(1, 3000, 2500), (2, 20000, 9000)])
def test_encode_ragged_vs_host(env, fmt, seed, n, big):
    print(f"[encode ragged] seed={seed} n={n} fmt={fmt}", flush=True)
    torch, ingest = env
    rowptr, col = _ragged_csr(seed, n, big)
    src = np.repeat(np.arange(n, dtype=np.uint32), np.diff(rowptr.astype(np.int64)))
    csr = ingest.csr_from_sorted_edges(src, col)
    assert csr.n == n
    expect = csr.compress(fmt, threads=8)
    _check(torch, ingest, rowptr, col, fmt, expect)


@pytest.mark.parametrize("scale", [16, 20])
def test_encode_rmat_vs_host(env, scale):
    print(f"[encode rmat] scale={scale}", flush=True)
    torch, ingest = env
    src, dst = ingest.rmat_edges(scale, 16, seed=scale, threads=8)
    csr = ingest.graph_sort(src, dst, threads=8)
    rowptr, col = csr.arrays()
    stream = torch.cuda.Stream()  # the work is ordered on the caller's stream
    for fmt in ("le", "gt"):
        _check(torch, ingest, rowptr, col, fmt, csr.compress(fmt, threads=8), stream=stream)


def test_encode_decodes_back(env):
    """Encode on the GPU, load the images, decode the whole graph: the input CSR comes back."""
    torch, ingest = env
    import grass_amd
    src, dst = ingest.rmat_edges(14, 16, seed=3, threads=8)
    rowptr, col = ingest.graph_sort(src, dst, threads=8).arrays()
    rp, cl = _dev(torch, rowptr, col)
    for fmt in ("le", "gt"):
        gb, ob = ingest.compress_device(rp, cl, fmt)
        g = grass_amd.Graph.from_bytes(gb, ob, fmt)
        r2, c2 = g.decode_csr()
        np.testing.assert_array_equal(r2, rowptr)
        np.testing.assert_array_equal(c2, col)
        g.close()


def test_encode_rejects_empty(env):
    torch, ingest = env
    rp = torch.zeros(1, dtype=torch.int64, device="cuda")
    cl = torch.zeros(1, dtype=torch.int32, device="cuda")
    with pytest.raises(Exception):
        ingest.compress_device(rp, cl, "le")


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_encode_one_huge_list(env, fmt):
    """A vertex with 300,000 neighbours (thousands of 64-field passes, wide rocks and seas) among
    short lists: the same bytes as the host compressor."""
    print(f"[encode huge list] fmt={fmt}", flush=True)
    torch, ingest = env
    rng = np.random.default_rng(77)
    n = 400000
    short = np.repeat(np.arange(n, dtype=np.int64), rng.integers(0, 4, size=n))
    short = short[short != 1234]
    keys = [short * n + rng.integers(0, n, size=len(short)),
            1234 * n + rng.choice(n, size=300000, replace=False).astype(np.int64),
            np.array([(n - 1) * n + (n - 1)], np.int64)]
    keys = np.unique(np.concatenate(keys))  # sorted (row, column) pairs, no repeats
    src, col = (keys // n).astype(np.uint32), (keys % n).astype(np.uint32)
    rowptr = np.zeros(n + 1, np.uint64)
    rowptr[1:] = np.cumsum(np.bincount(src, minlength=n))
    assert int(np.diff(rowptr.astype(np.int64))[1234]) == 300000
    csr = ingest.csr_from_sorted_edges(src, col)
    assert csr.n == n
    _check(torch, ingest, rowptr, col, fmt, csr.compress(fmt, threads=8))


@pytest.mark.parametrize("cp", ["2", "4"])
def test_gt_decode_across_empty_vertices(env, monkeypatch, cp):
    """GT full decode by runs over lists separated by long stretches of vertices without entries
    (more than a pass's 64 vertices), short lists whose runs span many vertices per pass, and a
    few long ones: the CSR comes back, at checkpoint intervals 2 and 4."""
    print(f"[gt decode empty stretches] P={cp}", flush=True)
    torch, ingest = env
    import grass_amd
    monkeypatch.setenv("GRASS_GT_CHECKPOINT", cp)
    rng = np.random.default_rng(11)
    n = 6000
    deg = rng.integers(1, 12, size=n).astype(np.int64)
    deg[rng.choice(n, size=8, replace=False)] = rng.integers(300, 3000, size=8)
    for lo, hi in ((100, 300), (1000, 1070), (2000, 2065), (4000, 4500)):
        deg[lo:hi] = 0
    deg[n - 1] = 1
    rowptr = np.zeros(n + 1, np.uint64)
    rowptr[1:] = np.cumsum(deg)
    # no self loops: a GT block head stores |first - v| - 1 (graph_sort drops self loops)
    col = np.concatenate([np.sort(np.setdiff1d(rng.choice(n, size=int(d) + 1, replace=False), [v])[:int(d)])
                          .astype(np.uint32) for v, d in enumerate(deg) if d])
    assert len(col) == int(rowptr[-1])
    rp, cl = _dev(torch, rowptr, col)
    g = grass_amd.Graph.from_csr_device(rp, cl, "gt")
    assert g.info()["gt_checkpoint"] == int(cp)
    for e in (512, 2048):
        g.set_tuning(decode_edges_per_wave=e)
        r2, c2 = g.decode_csr()
        np.testing.assert_array_equal(r2, rowptr)
        np.testing.assert_array_equal(c2, col)
    g.close()
