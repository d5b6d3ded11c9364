"""Device ingest (grass_rmat_edges_device, grass_graph_sort_device) against the host calls, which
test_ingest pins to the reference graph_sort; the graph_sort fixtures written by the compiled
reference tool directly; and the whole R-MAT -> graph_sort -> encoder pipeline on the GPU against
the host pipeline's files."""
import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch

    import grass_amd
    from grass_amd import ingest
    assert torch.cuda.is_available()
    grass_amd.lib()
    return torch, ingest


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("scale,abc,seed", [(10, (0.57, 0.19, 0.19), 42), (14, (0.45, 0.15, 0.15), 7),
                                            (16, (0.25, 0.25, 0.25), 3)])
def test_rmat_device_equals_host(env, scale, abc, seed):
    print(f"[ingest device] rmat scale={scale} abc={abc} seed={seed}", flush=True)
    torch, ingest = env
    hs, hd = ingest.rmat_edges(scale, 16, *abc, seed=seed, threads=8)
    ds, dd = ingest.rmat_edges_device(scale, 16, *abc, seed=seed)
    np.testing.assert_array_equal(_u32(ds), hs)
    np.testing.assert_array_equal(_u32(dd), hd)


def test_device_call_after_failed_runtime_call(env):
    """A failed HIP call left in the thread (here hipSetDevice on a device that does not exist,
    made by a rejected call) is not reported by the next call's launch check."""
    print("[ingest device] rmat after a failed hipSetDevice", flush=True)
    torch, ingest = env
    import ctypes as C
    src = torch.empty(16 << 10, dtype=torch.int32, device="cuda:0")
    dst = torch.empty_like(src)
    args = (10, 16, 0.57, 0.19, 0.19, 42)
    ptrs = (C.c_void_p(0), C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()))
    L = ingest.lib()
    assert L.grass_rmat_edges_device(*args, 4096, *ptrs) == 5  # GRASS_EDEVICE
    assert L.grass_rmat_edges_device(*args, 0, *ptrs) == 0
    torch.cuda.synchronize()
    hs, hd = ingest.rmat_edges(10, 16, 0.57, 0.19, 0.19, seed=42, threads=4)
    np.testing.assert_array_equal(_u32(src), hs)
    np.testing.assert_array_equal(_u32(dst), hd)


def _check_sort(torch, ingest, src, dst, directed, expect_rowptr, expect_col):
    s = torch.from_numpy(np.ascontiguousarray(src, dtype=np.uint32).view(np.int32)).cuda()
    d = torch.from_numpy(np.ascontiguousarray(dst, dtype=np.uint32).view(np.int32)).cuda()
    rp, col = ingest.graph_sort_device(s, d, directed=directed)
    np.testing.assert_array_equal(rp.cpu().numpy().astype(np.uint64), expect_rowptr)
    np.testing.assert_array_equal(_u32(col), expect_col)


@pytest.mark.parametrize("case", ["gs_small_u", "gs_small_d", "gs_rmat12_u"])
def test_graph_sort_device_reference_fixtures(env, case):
    """The outputs the compiled reference graph_sort wrote for these inputs."""
    print(f"[ingest device] graph_sort fixture {case}", flush=True)
    torch, ingest = env
    c = load_case(case)
    n = int(max(c["out_u"].max(), c["out_v"].max())) + 1
    rowptr = np.zeros(n + 1, np.uint64)
    np.add.at(rowptr, c["out_u"].astype(np.int64) + 1, 1)
    _check_sort(torch, ingest, c["raw_u"], c["raw_v"], bool(c["directed"]), np.cumsum(rowptr).astype(np.uint64),
                c["out_v"].astype(np.uint32))


@pytest.mark.parametrize("directed", [False, True])
@pytest.mark.parametrize("scale", [12, 18])
def test_graph_sort_device_equals_host(env, scale, directed):
    print(f"[ingest device] graph_sort rmat scale={scale} directed={directed}", flush=True)
    torch, ingest = env
    src, dst = ingest.rmat_edges(scale, 16, seed=scale, threads=8)
    rp, col = ingest.graph_sort(src, dst, directed=directed, threads=8).arrays()
    _check_sort(torch, ingest, src, dst, directed, rp, col)


@pytest.mark.parametrize("directed", [False, True])
def test_graph_sort_device_edge_cases(env, directed):
    """Self loops (dropped before numbering), repeated edges, ids that appear only as a
    destination (rows of degree 0 when directed), large sparse ids."""
    torch, ingest = env
    src = np.array([3, 1, 2, 5, 2, 2, 9, 1000000, 7, 1], np.uint32)
    dst = np.array([3, 2, 1, 5, 7, 7, 9, 2, 1000000, 2], np.uint32)
    rp, col = ingest.graph_sort(src, dst, directed=directed, threads=2).arrays()
    _check_sort(torch, ingest, src, dst, directed, rp, col)
    # only self loops: no vertices
    rp0, col0 = ingest.graph_sort(np.array([4, 4], np.uint32), np.array([4, 4], np.uint32),
                                  directed=directed).arrays()
    _check_sort(torch, ingest, [4, 4], [4, 4], directed, rp0, col0)


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_device_pipeline_equals_host_files(env, fmt):
    """R-MAT s20 -> graph_sort -> compressor, all on the GPU: the same files as on the host."""
    print(f"[ingest device] pipeline s20 fmt={fmt}", flush=True)
    torch, ingest = env
    s, d = ingest.rmat_edges_device(20, 16, seed=11)
    rp, col = ingest.graph_sort_device(s, d)
    gb, ob = ingest.compress_device(rp, col, fmt)
    hs, hd = ingest.rmat_edges(20, 16, seed=11, threads=8)
    hg, ho = ingest.graph_sort(hs, hd, threads=8).compress(fmt, threads=8)
    assert gb == hg and ob == ho


def test_graph_sort_device_capacity(env):
    """Too small output buffers: GRASS_ECAPACITY with the sizes reported, nothing written."""
    import ctypes as C
    torch, ingest = env
    import grass_amd
    src = np.array([0, 1, 2, 3], np.uint32)
    dst = np.array([1, 2, 3, 0], np.uint32)
    s = torch.from_numpy(src.view(np.int32)).cuda()
    d = torch.from_numpy(dst.view(np.int32)).cuda()
    rp = torch.zeros(8, dtype=torch.int64, device="cuda")
    col = torch.zeros(16, dtype=torch.int32, device="cuda")
    n, m = C.c_uint64(), C.c_uint64()
    L = ingest.lib()
    rc = L.grass_graph_sort_device(C.c_void_p(s.data_ptr()), C.c_void_p(d.data_ptr()), 4, 0, 0, None,
                                   C.c_void_p(rp.data_ptr()), 2, C.c_void_p(col.data_ptr()), 16, C.byref(n),
                                   C.byref(m))
    assert rc == grass_amd.ECAPACITY and n.value == 4
    rc = L.grass_graph_sort_device(C.c_void_p(s.data_ptr()), C.c_void_p(d.data_ptr()), 4, 0, 0, None,
                                   C.c_void_p(rp.data_ptr()), 8, C.c_void_p(col.data_ptr()), 3, C.byref(n),
                                   C.byref(m))
    assert rc == grass_amd.ECAPACITY and n.value == 4 and m.value == 8
    assert int(col.abs().sum().item()) == 0  # nothing written


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_compress_device_open_and_device_images(env, fmt):
    """A graph encoded on the GPU and loaded from its device images (no host round trip) is the
    graph its host files load to: same sizes, same CSR, same walks."""
    print(f"[ingest device] compress + open on the device fmt={fmt}", flush=True)
    torch, ingest = env
    import grass_amd
    s, d = ingest.rmat_edges_device(16, 16, seed=5)
    rp, col = ingest.graph_sort_device(s, d)
    gb, ob = ingest.compress_device(rp, col, fmt)
    ref = grass_amd.Graph.from_bytes(gb, ob, fmt)
    gd = grass_amd.Graph.from_csr_device(rp, col, fmt)
    gi = grass_amd.Graph.from_device_images(torch.frombuffer(bytearray(gb), dtype=torch.uint8).cuda(),
                                            torch.frombuffer(bytearray(ob), dtype=torch.uint8).cuda(), fmt)
    erp, ecol = rp.cpu().numpy().astype(np.uint64), col.cpu().numpy().view(np.uint32)
    wref = grass_amd.to_numpy_u32(ref.random_walk(4096, 30, seed=3))
    for g in (gd, gi):
        assert g.info()["graph_bytes"] == len(gb) and g.info()["offset_bytes"] == len(ob)
        r2, c2 = g.decode_csr()
        np.testing.assert_array_equal(r2, erp)
        np.testing.assert_array_equal(c2, ecol)
        np.testing.assert_array_equal(grass_amd.to_numpy_u32(g.random_walk(4096, 30, seed=3)), wref)
        g.close()
    ref.close()
