"""The GT payload stream the loader builds (block heads and gaps copied from .graph, the absolute
id stored in place of every P-th in-block gap) at several checkpoint intervals P: the full decode
equals the compressor's input CSR, walks, k-hop and MDRW equal the oracle, and the walk byte count
equals the oracle's model of the same interval."""
import numpy as np
import pytest

from conftest import csr_of, load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ga():
    import grass_amd
    import torch
    assert torch.cuda.is_available()
    grass_amd.lib()
    return grass_amd


@pytest.mark.parametrize("interval", [2, 8, 64, 1024])
@pytest.mark.parametrize("case", ["degrees", "hub", "rmat12"])
def test_gt_stream_intervals(ga, oracle_mod, monkeypatch, case, interval):
    print(f"[gt stream] case={case} interval={interval}", flush=True)
    monkeypatch.setenv("GRASS_GT_CHECKPOINT", str(interval))
    c = load_case(case)
    gb, ob = c["gt_graph"].tobytes(), c["gt_offset"].tobytes()
    g = ga.Graph.from_bytes(gb, ob, "gt")
    assert g.info()["gt_checkpoint"] == interval
    erp, ecol = csr_of(c)
    rowptr, col = g.decode_csr()
    np.testing.assert_array_equal(rowptr, erp)
    np.testing.assert_array_equal(col, ecol)

    og = oracle_mod.Graph(gb, ob, "gt")
    og.set_gt_checkpoint(interval)
    out = ga.to_numpy_u32(g.random_walk(2000, 30, seed=31, instance=interval))
    exp, exp_bytes = og.random_walk(2000, 30, seed=31, instance=interval)
    np.testing.assert_array_equal(out, exp)
    assert g.random_walk_bytes(2000, 30, seed=31, instance=interval) == exp_bytes

    rng = np.random.default_rng(interval)
    seeds = [rng.integers(0, g.n, size=k).astype(np.uint32) for k in (1, 40, 300)]
    fans = [[25, 10], [10, 10, 5], [64]]
    res = g.khop_batch(seeds, fans, seed=5, instances=[1, 2, 3])
    kexp = og.khop(seeds, fans, seed=5, instances=[1, 2, 3])
    assert res["total"] == len(kexp["dst"])
    for r in range(len(seeds)):
        s, d, _ = ga.Graph.khop_request_edges(res, r)
        e0, e1 = int(kexp["edge_off"][r]), int(kexp["edge_off"][r + 1])
        np.testing.assert_array_equal(s, kexp["src"][e0:e1])
        np.testing.assert_array_equal(d, kexp["dst"][e0:e1])

    sizes = [5, 60, 200]
    mres = g.mdrw(sizes, target=400, seed=9)
    mexp = og.mdrw(sizes, target=400, seed=9)
    for k in ("step_src", "step_dst", "verts", "ind_off", "ind_src", "ind_dst"):
        np.testing.assert_array_equal(mres[k], mexp[k], err_msg=k)
    g.close()
