"""GPU parity of batched k-hop sampling (without replacement, per-hop dedup) against the CPU
oracle on the same Philox counters, through the C ABI; and "alone == in the batch"."""
import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu

FMTS = ("le", "gt")


@pytest.fixture(scope="module")
def ga():
    import grass_amd
    import torch
    assert torch.cuda.is_available()
    grass_amd.lib()
    return grass_amd


def _pair(ga, oracle_mod, case, fmt):
    c = load_case(case)
    gb, ob = c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes()
    return ga.Graph.from_bytes(gb, ob, fmt), oracle_mod.Graph(gb, ob, fmt)


def _compare(ga, res, exp, R, hops):
    for r in range(R):
        s, d, counts = ga.Graph.khop_request_edges(res, r)
        e0, e1 = int(exp["edge_off"][r]), int(exp["edge_off"][r + 1])
        np.testing.assert_array_equal(s, exp["src"][e0:e1], err_msg=f"request {r} src")
        np.testing.assert_array_equal(d, exp["dst"][e0:e1], err_msg=f"request {r} dst")
        assert counts[: hops[r]] == [int(x) for x in exp["hop_count"][r, : hops[r]]]


@pytest.mark.parametrize("slot", ["8", "16"])
@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", ["rmat12", "degrees", "hub"])
def test_khop_matches_oracle(ga, oracle_mod, case, fmt, slot, monkeypatch):
    """Both dedup tables: 8-byte packed slots (the default when the request, vertex and position
    bits fit a word) and the 16-byte generation-tagged ones (GRASS_KHOP_SLOT=16)."""
    print(f"[khop] case={case} fmt={fmt} slot={slot}", flush=True)
    if slot == "16":
        monkeypatch.setenv("GRASS_KHOP_SLOT", "16")
    g, og = _pair(ga, oracle_mod, case, fmt)
    rng = np.random.default_rng(11)
    sizes = [1, 3, 17, 64, 200, 1000, 0, 5]
    seeds = [rng.integers(0, g.n, size=k).astype(np.uint32) for k in sizes]
    seeds[5][:10] = seeds[5][0]  # duplicate seeds
    fans = [[25, 10], [25, 10], [15, 10], [10, 10, 5], [3], [25, 10], [4, 4], [64, 2]]
    inst = [100 + i for i in range(len(seeds))]
    res = g.khop_batch(seeds, fans, seed=2024, instances=inst)
    exp = og.khop(seeds, fans, seed=2024, instances=inst)
    assert res["total"] == len(exp["dst"])
    _compare(ga, res, exp, len(seeds), [len(f) for f in fans])
    g.close()


@pytest.mark.parametrize("fmt", FMTS)
def test_khop_alone_equals_batch(ga, oracle_mod, fmt):
    g, og = _pair(ga, oracle_mod, "rmat12", fmt)
    rng = np.random.default_rng(3)
    seeds = [rng.integers(0, g.n, size=int(rng.integers(1, 300))).astype(np.uint32) for _ in range(30)]
    fans = [[25, 10] if i % 3 else [10, 10, 5] for i in range(30)]
    inst = list(range(500, 530))
    batch = g.khop_batch(seeds, fans, seed=9, instances=inst)
    for r in (0, 7, 29):
        alone = g.khop_batch([seeds[r]], [fans[r]], seed=9, instances=[inst[r]])
        bs, bd, bc = ga.Graph.khop_request_edges(batch, r)
        a_s, a_d, ac = ga.Graph.khop_request_edges(alone, 0)
        np.testing.assert_array_equal(bs, a_s)
        np.testing.assert_array_equal(bd, a_d)
    exp = og.khop(seeds, fans, seed=9, instances=inst)
    _compare(ga, batch, exp, 30, [len(f) for f in fans])
    g.close()


def test_khop_bad_seed_and_capacity(ga):
    c = load_case("small_u")
    g = ga.Graph.from_bytes(c["le_graph"].tobytes(), c["le_offset"].tobytes(), "le")
    with pytest.raises(ga.GrassError) as e:
        g.khop_batch([np.array([0, g.n + 5], np.uint32)], [[3]], seed=1)
    assert e.value.status == ga.EINVAL
    g.sync()
    # capacity too small -> GRASS_ECAPACITY with the needed total
    import ctypes as C
    import torch
    arr, keep = g.khop_request_array([np.arange(20, dtype=np.uint32)], [[5, 5]])
    src = torch.empty(4, dtype=torch.int32, device="cuda")
    dst = torch.empty(4, dtype=torch.int32, device="cuda")
    off = torch.zeros(2 * 2, dtype=torch.int64, device="cuda")
    out = ga.KhopOutput(src.data_ptr(), dst.data_ptr(), 4, off.data_ptr(), 0, 0)
    rc = ga.lib().grass_khop_batch(g.handle, arr, 1, 1, C.byref(out), None)
    assert rc == ga.ECAPACITY and out.total > 4
    g.close()
