"""Serving session on the GPU: every result of a mixed stream of random-walk and k-hop requests
equals the oracle's result for that request alone; cancellation, release, memory-based
admission and trace replay."""
import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    import torch
    assert torch.cuda.is_available()
    import grass_amd
    from grass_amd import serve
    return grass_amd, serve


def _graphs(grass_amd, oracle_mod, fmt):
    c = load_case("rmat12")
    gb, ob = c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes()
    return grass_amd.Graph.from_bytes(gb, ob, fmt), oracle_mod.Graph(gb, ob, fmt)


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_mixed_stream_equals_alone(env, oracle_mod, fmt):
    grass_amd, serve = env
    g, og = _graphs(grass_amd, oracle_mod, fmt)
    rng = np.random.default_rng(17)
    reqs, specs = [], []
    for i in range(120):
        if i % 2 == 0:
            W, L = int(rng.integers(1, 500)), int(rng.integers(1, 60))
            starts = rng.integers(0, g.n, size=W).astype(np.uint32) if i % 4 == 0 else None
            reqs.append(serve.rw_request(W, L, 7000 + i, starts))
            specs.append(("rw", W, L, starts, 7000 + i))
        else:
            seeds = rng.integers(0, g.n, size=int(rng.integers(1, 200))).astype(np.uint32)
            fan = [15, 10] if i % 3 else [10, 10, 5]
            reqs.append(serve.khop_request(seeds, fan, 7000 + i))
            specs.append(("kh", seeds, fan, 7000 + i))
    with serve.Session(g, memory_budget=256 << 20, max_batch=16, max_inflight=3, seed=99) as s:
        tickets = [s.submit(r) for r in reqs]
        for t in tickets:
            assert s.wait(t) == 0
        for t, sp in zip(tickets, specs):
            out = s.fetch(t)
            if sp[0] == "rw":
                exp, _ = og.random_walk(sp[1], sp[2], seed=99, instance=sp[4], starts=sp[3])
                np.testing.assert_array_equal(out, exp)
            else:
                exp = og.khop([sp[1]], [sp[2]], seed=99, instances=[sp[3]])
                src, dst, counts = out
                np.testing.assert_array_equal(src, exp["src"])
                np.testing.assert_array_equal(dst, exp["dst"])
                assert counts == [int(x) for x in exp["hop_count"][0, : len(sp[2])]]
            s.release(t)
        st = s.stats()
        assert st["completed"] == 120 and st["groups"] >= 8 and st["memory_in_use"] == 0
    g.close()


def test_admission_cancel_release(env):
    grass_amd, serve = env
    c = load_case("rmat12")
    g = grass_amd.Graph.from_bytes(c["le_graph"].tobytes(), c["le_offset"].tobytes(), "le")
    with serve.Session(g, memory_budget=1 << 20, max_batch=4, max_inflight=1, seed=1) as s:
        # larger than the whole arena: rejected at submit
        rc, _ = s.try_submit(serve.rw_request(100000, 50, 1))
        assert rc == grass_amd.ECAPACITY
        # fills most of the arena; the next request must wait until the first is released
        t1 = s.submit(serve.rw_request(4000, 50, 2))
        assert s.wait(t1) == 0
        t2 = s.submit(serve.rw_request(4000, 50, 3))
        assert s.wait(t2, 200000) == grass_amd.ETIMEOUT  # held back by memory admission
        t3 = s.submit(serve.rw_request(10, 5, 4))
        assert s.cancel(t3) == 0
        assert s.poll(t3) == grass_amd.ECANCELLED
        s.release(t1)
        assert s.wait(t2) == 0
        s.release(t2)
        st = s.stats()
        assert st["rejected"] == 1 and st["cancelled"] == 1 and st["memory_in_use"] == 0
    g.close()


def test_oversized_and_invalid_requests(env):
    """Sizes that would overflow or exceed the arena are rejected before any id is read; malformed
    requests are EINVAL; the session keeps serving afterwards."""
    grass_amd, serve = env
    c = load_case("rmat12")
    g = grass_amd.Graph.from_bytes(c["le_graph"].tobytes(), c["le_offset"].tobytes(), "le")
    with serve.Session(g, memory_budget=8 << 20, max_batch=8, max_inflight=2, seed=3) as s:
        big = serve.rw_request(0xFFFFFFFF, 0xFFFFFFFE, 1)  # walkers x (length + 1) x 4 > 2^64
        assert s.try_submit(big)[0] == grass_amd.ECAPACITY
        many = serve.khop_request(np.zeros(4, np.uint32), [10], 2)
        many.num_seeds = 1 << 40  # far more seeds than the arena holds: never read
        assert s.try_submit(many)[0] == grass_amd.ECAPACITY
        bad_fan = serve.khop_request(np.zeros(4, np.uint32), [65], 3)
        assert s.try_submit(bad_fan)[0] == grass_amd.EINVAL
        bad_id = serve.khop_request(np.array([g.n], np.uint32), [5], 4)
        assert s.try_submit(bad_id)[0] == grass_amd.EINVAL
        assert s.try_submit(serve.rw_request(0, 5, 5))[0] == grass_amd.EINVAL
        t = s.submit(serve.rw_request(100, 10, 6))
        assert s.wait(t) == 0
        s.release(t)
        assert s.stats()["rejected"] == 2
    g.close()


def test_replay_latencies(env):
    grass_amd, serve = env
    c = load_case("rmat12")
    g = grass_amd.Graph.from_bytes(c["gt_graph"].tobytes(), c["gt_offset"].tobytes(), "gt")
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(300):
        if i % 2:
            reqs.append(serve.rw_request(int(rng.integers(1, 256)), int(rng.integers(10, 40)), i))
        else:
            reqs.append(serve.khop_request(rng.integers(0, g.n, size=int(rng.integers(1, 64))),
                                           [15, 10], i))
    arrival = np.cumsum(rng.exponential(200.0, size=300))  # ~5k requests/s
    with serve.Session(g, memory_budget=512 << 20, max_batch=64, max_inflight=3) as s:
        lat, wall = s.replay(reqs, arrival)
        assert np.all(lat > 0) and wall > 0
        st = s.stats()
        assert st["completed"] == 300 and st["memory_in_use"] == 0
    g.close()
