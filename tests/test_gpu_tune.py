"""Launch-shape tuning (grass_tuning, grass_autotune): every candidate of every knob gives the
oracle's results -- walks, k-hop edges and the full decode -- and the autotuner keeps a valid
choice; invalid settings are rejected without changing the current one."""
import itertools

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


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_every_candidate_gives_the_oracle_result(ga, oracle_mod, fmt):
    c = load_case("rmat12")
    gb, ob = c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes()
    g, og = ga.Graph.from_bytes(gb, ob, fmt), oracle_mod.Graph(gb, ob, fmt)
    rowptr, col = csr_of(c)
    rng = np.random.default_rng(6)
    reqs = [{"walkers": int(rng.integers(1, 300)), "length": int(rng.integers(1, 70)), "instance": i}
            for i in range(20)]
    exp_w = [og.random_walk(r["walkers"], r["length"], seed=8, instance=r["instance"])[0] for r in reqs]
    seeds = [rng.integers(0, g.n, size=int(k)).astype(np.uint32) for k in (5, 60, 300)]
    fans = [[25, 10], [6], [10, 10, 5]]
    exp_k = og.khop(seeds, fans, seed=8)
    for st, cl in itertools.product((8, 16, 32), (1, 2, 4)):
        print(f"[tune] {fmt} stage={st} claim={cl}", flush=True)
        g.set_tuning(rw_stage=st, rw_claim=cl)
        outs = g.random_walk_batch(reqs, seed=8)
        for o, e in zip(outs, exp_w):
            np.testing.assert_array_equal(ga.to_numpy_u32(o), e)
        np.testing.assert_array_equal(ga.to_numpy_u32(g.random_walk(300, 33, seed=8, instance=1)),
                                      og.random_walk(300, 33, seed=8, instance=1)[0])
    for sc in (0, 1, 2):
        g.set_tuning(khop_lane_scale=sc)
        res = g.khop_batch(seeds, fans, seed=8)
        assert res["total"] == len(exp_k["dst"])
        for r in range(len(seeds)):
            s, d, _ = ga.Graph.khop_request_edges(res, r)
            e0, e1 = int(exp_k["edge_off"][r]), int(exp_k["edge_off"][r + 1])
            np.testing.assert_array_equal(d, exp_k["dst"][e0:e1])
    for e in (512, 1024, 2048, 4096, 8192):
        g.set_tuning(decode_edges_per_wave=e)
        rp, cl = g.decode_csr()
        np.testing.assert_array_equal(rp, rowptr)
        np.testing.assert_array_equal(cl, col)
    g.close()


def test_autotune_and_validation(ga, oracle_mod):
    c = load_case("rmat12")
    gb, ob = c["gt_graph"].tobytes(), c["gt_offset"].tobytes()
    g, og = ga.Graph.from_bytes(gb, ob, "gt"), oracle_mod.Graph(gb, ob, "gt")
    chosen, rep = g.autotune()
    assert len(rep["rw"]) == 9 and len(rep["khop"]) == 3 and len(rep["decode"]) == 5
    assert all(ms > 0 for *_, ms in rep["rw"] + rep["khop"] + rep["decode"])
    assert chosen == g.tuning()
    assert chosen["rw_stage"] in (8, 16, 32) and chosen["rw_claim"] in (1, 2, 4)
    # results under the chosen tuning
    out = ga.to_numpy_u32(g.random_walk(1000, 40, seed=3, instance=9))
    np.testing.assert_array_equal(out, og.random_walk(1000, 40, seed=3, instance=9)[0])
    before = g.tuning()
    for bad in ({"rw_stage": 12}, {"rw_claim": 3}, {"khop_lane_scale": 4}, {"decode_edges_per_wave": 100}):
        with pytest.raises(ga.GrassError) as e:
            g.set_tuning(**bad)
        assert e.value.status == ga.EINVAL
        assert g.tuning() == before
    g.close()
