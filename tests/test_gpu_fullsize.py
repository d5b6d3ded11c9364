"""Parity at the workloads' full sizes: R-MAT s22 (LE and GT) full decode and 1,048,576 walks x 80
against the ingest CSR and the oracle; R-MAT s24 GT batched 2-hop [25,10] requests and MDRW
instances (frontier U[500,4000], target 16384) against the oracle; size-independent properties
of a full 512-instance MDRW batch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ga():
    import grass_amd
    import torch
    assert torch.cuda.is_available()
    grass_amd.lib()
    return grass_amd


@pytest.fixture(scope="module")
def s24(ga, oracle_mod):
    from grass_amd import ingest
    print("[fullsize] ingest s24 gt", flush=True)
    gb, ob, csr = ingest.rmat_graph(24, 16, fmt="gt", seed=42, threads=16)
    g = ga.Graph.from_bytes(gb, ob, "gt")
    og = oracle_mod.Graph(gb, ob, "gt")
    yield g, og, csr
    g.close()


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_s22_decode_and_walks(ga, oracle_mod, fmt):
    from grass_amd import ingest
    print(f"[fullsize] s22 {fmt}", flush=True)
    gb, ob, csr = ingest.rmat_graph(22, 16, fmt=fmt, seed=42, threads=16)
    rowptr, col = csr.arrays()
    g = ga.Graph.from_bytes(gb, ob, fmt)
    rp, cl = g.decode_csr()
    np.testing.assert_array_equal(rp, rowptr)
    np.testing.assert_array_equal(cl, col)
    og = oracle_mod.Graph(gb, ob, fmt)
    W, L = 1 << 20, 80
    out = ga.to_numpy_u32(g.random_walk(W, L, seed=42, instance=0))
    exp, _ = og.random_walk(W, L, seed=42, instance=0)
    np.testing.assert_array_equal(out, exp)
    g.close()


def test_s24_khop_batch(ga, s24):
    g, og, _ = s24
    rng = np.random.default_rng(42)
    seeds = [rng.integers(0, g.n, size=int(np.clip(round(rng.lognormal(6.0, 1.2)), 1, 8192)))
             .astype(np.uint32) for _ in range(256)]
    fans = [[25, 10]] * 256
    print(f"[fullsize] s24 khop seeds={sum(len(s) for s in seeds)}", flush=True)
    res = g.khop_batch(seeds, fans, seed=42)
    exp = og.khop(seeds, fans, seed=42)
    assert res["total"] == len(exp["dst"])
    for r in range(256):
        s, d, _ = ga.Graph.khop_request_edges(res, r)
        e0, e1 = int(exp["edge_off"][r]), int(exp["edge_off"][r + 1])
        np.testing.assert_array_equal(s, exp["src"][e0:e1])
        np.testing.assert_array_equal(d, exp["dst"][e0:e1])


def test_s24_mdrw_instances(ga, s24):
    g, og, _ = s24
    rng = np.random.default_rng(4)
    sizes = [int(x) for x in rng.integers(500, 4001, size=12)]
    print("[fullsize] s24 mdrw 12 instances", flush=True)
    res = g.mdrw(sizes, target=16384, seed=42)
    exp = og.mdrw(sizes, target=16384, seed=42)
    for k in ("step_off", "step_src", "step_dst", "vert_off", "verts", "ind_off", "ind_src", "ind_dst"):
        np.testing.assert_array_equal(res[k], exp[k], err_msg=k)


def test_s24_mdrw_full_batch_properties(ga, s24):
    g, _, csr = s24
    rowptr, col = csr.arrays()
    rng = np.random.default_rng(5)
    sizes = rng.integers(500, 4001, size=512)
    print("[fullsize] s24 mdrw 512 instances", flush=True)
    res = g.mdrw([int(x) for x in sizes], target=16384, seed=42)
    so, vo, io = (res[k].astype(np.int64) for k in ("step_off", "vert_off", "ind_off"))
    np.testing.assert_array_equal(np.diff(so), 16384 - sizes)  # no frontier runs dry on R-MAT
    # every step is an edge of the CSR the graph was compressed from
    src, dst = res["step_src"].astype(np.int64), res["step_dst"].astype(np.int64)
    lo, hi = rowptr[src].astype(np.int64), rowptr[src + 1].astype(np.int64)
    pos = np.array([np.searchsorted(col[a:b], d) + a for a, b, d in
                    zip(lo[::97], hi[::97], dst[::97])])
    assert np.all(col[pos] == dst[::97])
    for i in (0, 255, 511):
        v = res["verts"][vo[i]:vo[i + 1]]
        assert np.all(np.diff(v.astype(np.int64)) > 0)
        vs = set(v.tolist())
        assert set(res["step_dst"][so[i]:so[i + 1]].tolist()) <= vs
        isrc, idst = res["ind_src"][io[i]:io[i + 1]], res["ind_dst"][io[i]:io[i + 1]]
        # induced edges: exactly the CSR rows of the sample restricted to the sample
        exp = [(int(x), int(y)) for x in v[:300] for y in col[rowptr[x]:rowptr[x + 1]] if int(y) in vs]
        got = list(zip(isrc[:len(exp)].tolist(), idst[:len(exp)].tolist()))
        assert got == exp
