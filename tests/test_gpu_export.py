"""Sample export: k-hop mini-batch blocks (each request's vertices numbered by first appearance
over seeds, hop-0 destinations, hop-1 destinations, ...; every batch edge in those local ids),
local ids of MDRW samples, and MDRW results handed to torch through DLPack without a copy."""
import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ga():
    import torch

    import grass_amd
    assert torch.cuda.is_available()
    grass_amd.lib()
    return grass_amd


@pytest.fixture(params=["8", "16"])
def slot(request, monkeypatch):
    """The table layout of the k-hop blocks: packed 8-byte slots (chosen whenever the request,
    vertex and position bits fit one word, as in every test here) or 16-byte slots (forced)."""
    if request.param == "16":
        monkeypatch.setenv("GRASS_EXPORT_SLOT", "16")
    else:
        monkeypatch.delenv("GRASS_EXPORT_SLOT", raising=False)
    return request.param


def _np(t):
    return t.cpu().numpy()


def blocks_by_dict(seeds_list, offsets, src, dst):
    """The numbering restated with a dict per request: first appearance in the sequence seeds,
    hop-0 destinations, hop-1 destinations, ..."""
    R, H = len(seeds_list), offsets.shape[0]
    nodes, node_off = [], [0]
    lsrc = np.full(len(src), -1, np.int64)
    ldst = np.full(len(src), -1, np.int64)
    for r in range(R):
        ids = {}
        for v in seeds_list[r]:
            ids.setdefault(int(v), len(ids))
        for h in range(H):
            for e in range(int(offsets[h, r]), int(offsets[h, r + 1])):
                assert int(src[e]) in ids  # a source is a seed or an earlier destination
                lsrc[e] = ids[int(src[e])]
                ldst[e] = ids.setdefault(int(dst[e]), len(ids))
        nodes.extend(ids.keys())
        node_off.append(len(nodes))
    return np.array(nodes, np.int64), np.array(node_off, np.int64), lsrc, ldst


def _check_blocks(res, seeds_list):
    off = _np(res["offsets"])
    src, dst = _np(res["src"]).view(np.uint32), _np(res["dst"]).view(np.uint32)
    nodes, node_off, lsrc, ldst = blocks_by_dict(seeds_list, off, src, dst)
    np.testing.assert_array_equal(_np(res["node_off"]), node_off)
    np.testing.assert_array_equal(_np(res["nodes"]).view(np.uint32).astype(np.int64), nodes)
    np.testing.assert_array_equal(_np(res["lsrc"]).astype(np.int64), lsrc)
    np.testing.assert_array_equal(_np(res["ldst"]).astype(np.int64), ldst)


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_khop_blocks_small(ga, oracle_mod, fmt, slot):
    print(f"[export khop blocks] rmat12 fmt={fmt} slot={slot}", flush=True)
    c = load_case("rmat12")
    gb, ob = c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes()
    g = ga.Graph.from_bytes(gb, ob, fmt)
    rng = np.random.default_rng(4)
    seeds = [rng.integers(0, g.n, size=k).astype(np.uint32) for k in (1, 40, 300, 7, 2)]
    seeds[3] = np.array([5, 5, 9, 5, 9], np.uint32)  # duplicate seeds keep their first local id
    seeds[4] = np.zeros(0, np.uint32)                # an empty request
    fans = [[25, 10], [10, 10, 5], [64], [3, 3], [4]]
    res = g.khop_minibatch(seeds, fans, seed=5, instances=[1, 2, 3, 4, 5])
    # the batch part is the plain k-hop batch, which equals the oracle
    kexp = oracle_mod.Graph(gb, ob, fmt).khop(seeds, fans, seed=5, instances=[1, 2, 3, 4, 5])
    assert res["total"] == len(kexp["dst"])
    _check_blocks(res, seeds)
    # the distinct seeds come first, in order
    n0 = _np(res["node_off"])
    np.testing.assert_array_equal(_np(res["nodes"])[n0[3]:n0[3] + 2], [5, 9])
    assert n0[5] == n0[4]
    g.close()


def test_khop_blocks_rmat18(ga, slot):
    """A serving-sized batch (64 requests of 200 seeds, [25, 10]) on R-MAT-18, checked against
    the dict numbering and the invariants nodes[local] == global id, one local id per vertex."""
    from grass_amd import ingest
    src, dst = ingest.rmat_edges(18, 16, seed=18, threads=16)
    rowptr, col = ingest.graph_sort(src, dst, threads=16).arrays()
    import torch
    rp = torch.from_numpy(rowptr.astype(np.int64)).cuda()
    cl = torch.from_numpy(col.view(np.int32)).cuda()
    gb, ob = ingest.compress_device(rp, cl, "gt")
    g = ga.Graph.from_bytes(gb, ob, "gt")
    rng = np.random.default_rng(18)
    seeds = [rng.integers(0, g.n, size=200).astype(np.uint32) for _ in range(64)]
    res = g.khop_minibatch(seeds, [[25, 10]] * 64, seed=7)
    print(f"[export khop blocks] rmat18 slot={slot} edges={res['total']} nodes={res['nodes'].numel()}",
          flush=True)
    _check_blocks(res, seeds)
    nodes = _np(res["nodes"]).view(np.uint32)
    n0 = _np(res["node_off"])
    off = _np(res["offsets"])
    for r in (0, 31, 63):
        seg = nodes[n0[r]:n0[r + 1]]
        assert len(np.unique(seg)) == len(seg)
        for h in range(off.shape[0]):
            a, b = off[h, r], off[h, r + 1]
            np.testing.assert_array_equal(seg[_np(res["lsrc"])[a:b]], _np(res["src"]).view(np.uint32)[a:b])
            np.testing.assert_array_equal(seg[_np(res["ldst"])[a:b]], _np(res["dst"]).view(np.uint32)[a:b])
    g.close()


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_mdrw_dlpack_and_local_ids(ga, fmt):
    print(f"[export mdrw dlpack] rmat12 fmt={fmt}", flush=True)
    import torch
    c = load_case("rmat12")
    g = ga.Graph.from_bytes(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    sizes = [5, 60, 200, 1]
    host = g.mdrw(sizes, target=400, seed=9)
    dev = g.mdrw_tensors(sizes, target=400, seed=9)  # the result was freed; the tensors hold it
    torch.cuda.synchronize()
    for k in ga.MDRW_ARRAYS:
        t = dev[k]
        assert t.is_cuda and t.dtype == (torch.int64 if k.endswith("_off") else torch.int32), k
        np.testing.assert_array_equal(_np(t).view(np.uint64 if k.endswith("_off") else np.uint32), host[k],
                                      err_msg=k)
    for pre in ("ind", "step"):
        off = host[pre + "_off"]
        for i in range(len(sizes)):
            V = host["verts"][host["vert_off"][i]:host["vert_off"][i + 1]]
            a, b = int(off[i]), int(off[i + 1])
            np.testing.assert_array_equal(_np(dev[pre + "_lsrc"])[a:b], np.searchsorted(V, host[pre + "_src"][a:b]))
            np.testing.assert_array_equal(_np(dev[pre + "_ldst"])[a:b], np.searchsorted(V, host[pre + "_dst"][a:b]))
            assert (V[_np(dev[pre + "_lsrc"])[a:b]] == host[pre + "_src"][a:b]).all()
    # the memory goes with the last tensor: drop them and run another batch over the same sizes
    del dev
    again = g.mdrw_tensors(sizes, target=400, seed=9, local=False)
    np.testing.assert_array_equal(_np(again["ind_dst"]).view(np.uint32), host["ind_dst"])
    g.close()


def test_khop_blocks_reject_offsets_of_another_batch(ga, slot):
    """Offsets that do not belong to the requests (request 1 claims request 0's hop-0 edges; every
    read stays inside the batch's arrays) give GRASS_EINVAL, and the graph stays usable."""
    import ctypes as C

    import torch
    print("[export khop blocks] mismatched offsets", flush=True)
    c = load_case("rmat12")
    g = ga.Graph.from_bytes(c["gt_graph"].tobytes(), c["gt_offset"].tobytes(), "gt")
    from conftest import csr_of
    rowptr, _ = csr_of(c)
    busy = np.flatnonzero(np.diff(rowptr.astype(np.int64)) >= 2).astype(np.uint32)
    seeds = [busy[:3], busy[-2:]]
    arr, keep = g.khop_request_array(seeds, [[5, 5], [5, 5]])
    cap = g.khop_bound(arr, 2)
    src = torch.empty(cap, dtype=torch.int32, device="cuda")
    dst = torch.empty(cap, dtype=torch.int32, device="cuda")
    off = torch.zeros(6, dtype=torch.int64, device="cuda")
    out = ga.KhopOutput(src.data_ptr(), dst.data_ptr(), cap, off.data_ptr(), 0, 0)
    L = ga.lib()
    assert L.grass_khop_batch(g.handle, arr, 2, 3, C.byref(out), None) == ga.OK
    g.sync()
    o = _np(off)
    assert o[1] > o[0] and o[2] > o[1]
    off[1] = off[0]
    nodes = torch.empty(cap + 5, dtype=torch.int32, device="cuda")
    noff = torch.empty(3, dtype=torch.int64, device="cuda")
    ls = torch.empty(cap, dtype=torch.int32, device="cuda")
    ld = torch.empty(cap, dtype=torch.int32, device="cuda")
    blk = ga.KhopBlocks(nodes.data_ptr(), cap + 5, noff.data_ptr(), ls.data_ptr(), ld.data_ptr(), 0)
    assert L.grass_khop_blocks_build(g.handle, arr, 2, C.byref(out), C.byref(blk), None) == ga.EINVAL
    g.sync()  # no error left behind on the graph
    res = g.khop_minibatch(seeds, [[5, 5], [5, 5]], seed=3)
    _check_blocks(res, seeds)
    g.close()


def test_khop_blocks_reject_malformed_offsets_and_ids(ga, slot):
    """Offsets checked before any pass reads through them: a max_hops larger than the batch wrote
    (rows of garbage past the total), a descending row, a row that ends past the total, and a seed
    that is not a vertex (the seed buffer rewritten after the batch) all give GRASS_EINVAL, under
    both table layouts, and leave the graph usable."""
    import ctypes as C

    import torch
    print(f"[export khop blocks] malformed offsets slot={slot}", flush=True)
    c = load_case("rmat12")
    g = ga.Graph.from_bytes(c["gt_graph"].tobytes(), c["gt_offset"].tobytes(), "gt")
    from conftest import csr_of
    rowptr, _ = csr_of(c)
    busy = np.flatnonzero(np.diff(rowptr.astype(np.int64)) >= 2).astype(np.uint32)
    seeds = [busy[:3], busy[-2:]]
    arr, keep = g.khop_request_array(seeds, [[5, 5], [5, 5]])
    cap = g.khop_bound(arr, 2)
    src = torch.empty(cap, dtype=torch.int32, device="cuda")
    dst = torch.empty(cap, dtype=torch.int32, device="cuda")
    H4 = 4
    off = torch.zeros(H4 * 3, dtype=torch.int64, device="cuda")
    out = ga.KhopOutput(src.data_ptr(), dst.data_ptr(), cap, off.data_ptr(), 0, 0)
    L = ga.lib()
    assert L.grass_khop_batch(g.handle, arr, 2, 3, C.byref(out), None) == ga.OK
    g.sync()
    good = off.clone()
    total = out.total
    nodes = torch.empty(cap + 5, dtype=torch.int32, device="cuda")
    noff = torch.empty(3, dtype=torch.int64, device="cuda")
    ls = torch.empty(cap, dtype=torch.int32, device="cuda")
    ld = torch.empty(cap, dtype=torch.int32, device="cuda")

    def build(max_hops):
        o = ga.KhopOutput(src.data_ptr(), dst.data_ptr(), cap, off.data_ptr(), total, max_hops)
        blk = ga.KhopBlocks(nodes.data_ptr(), cap + 5, noff.data_ptr(), ls.data_ptr(), ld.data_ptr(), 0)
        rc = L.grass_khop_blocks_build(g.handle, arr, 2, C.byref(o), C.byref(blk), None)
        g.sync()  # no error left behind on the graph
        return rc

    assert build(2) == ga.OK
    off[6:] = 1 << 40  # rows 2 and 3 were never written by a 2-hop batch
    assert build(H4) == ga.EINVAL
    off.copy_(good)
    off[0], off[1] = 5, 2  # descending row
    assert build(2) == ga.EINVAL
    off.copy_(good)
    off[5] = total + 100  # hop 1 ends past the batch total
    assert build(2) == ga.EINVAL
    off.copy_(good)
    seed_dev = keep[0]
    saved = seed_dev.clone()
    seed_dev[0] = g.n  # not a vertex of the graph
    assert build(2) == ga.EINVAL
    seed_dev.copy_(saved)
    assert build(2) == ga.OK
    res = g.khop_minibatch(seeds, [[5, 5], [5, 5]], seed=3)
    _check_blocks(res, seeds)
    g.close()


def test_mdrw_dlpack_memory_goes_with_the_last_tensor(ga):
    """The result's device memory stays while any exported tensor lives (after grass_mdrw_free)
    and is returned when the last one goes."""
    import torch

    from grass_amd import ingest
    print("[export mdrw dlpack] memory lifetime", flush=True)
    s, d = ingest.rmat_edges_device(18, 16, seed=8)
    rp, col = ingest.graph_sort_device(s, d)
    del s, d
    g = ga.Graph.from_csr_device(rp, col, "gt")
    del rp, col
    sizes = [4000] * 48
    g.mdrw_tensors(sizes, target=16384, seed=1, local=False)  # the graph's workspaces grow here
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    t = g.mdrw_tensors(sizes, target=16384, seed=1, local=False)
    keep = t["ind_dst"]
    held = sum(x.numel() * x.element_size() for x in t.values())
    assert held > 64 << 20
    del t
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free0 - free1 >= held // 2  # one tensor still holds the whole result
    assert int(keep[-1].item()) >= 0
    del keep
    torch.cuda.synchronize()
    free2 = torch.cuda.mem_get_info()[0]
    assert free2 - free1 >= held // 2  # returned with the last tensor
    g.close()
