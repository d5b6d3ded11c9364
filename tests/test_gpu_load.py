"""Loading large images: mapped files and host buffers streamed through the pinned chunks with one
or several copy threads (GRASS_LOAD_THREADS) decode to the same CSR -- an R-MAT s21 graph whose
.graph spans several 64 MB chunks and whose size is not a multiple of the copy slices."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def big(tmp_path_factory):
    import torch

    import grass_amd
    from grass_amd import ingest
    assert torch.cuda.is_available()
    grass_amd.lib()
    s, d = ingest.rmat_edges_device(21, 16, seed=21)
    rp, col = ingest.graph_sort_device(s, d)
    out = {}
    for fmt in ("le", "gt"):
        gb, ob = ingest.compress_device(rp, col, fmt)
        prefix = tmp_path_factory.mktemp("load") / f"r21{fmt}"
        ingest.write_files(prefix, gb, ob)
        out[fmt] = (str(prefix), gb, ob)
    return grass_amd, out, rp.cpu().numpy().astype(np.uint64), col.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("threads", ["1", "3", "8"])
@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_load_threads_files_and_buffers(big, monkeypatch, fmt, threads):
    ga, out, erp, ecol = big
    prefix, gb, ob = out[fmt]
    print(f"[load] fmt={fmt} threads={threads} graph={len(gb)} bytes", flush=True)
    size = len(gb)
    assert size > 2 * (64 << 20)  # several staging chunks
    monkeypatch.setenv("GRASS_LOAD_THREADS", threads)
    for g in (ga.Graph.open(prefix, fmt), ga.Graph.from_bytes(gb, ob, fmt)):
        assert g.info()["graph_bytes"] == len(gb)
        rowptr, col = g.decode_csr()
        np.testing.assert_array_equal(rowptr, erp)
        np.testing.assert_array_equal(col, ecol)
        g.close()
