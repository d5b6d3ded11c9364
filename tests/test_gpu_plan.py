"""Captured batches (grass_plan_*): every replay of the HIP graph equals the oracle -- read back on
the launch stream with no device-wide synchronisation, with eager work on the same graph between
replays, and after the seed buffers are rewritten in place."""
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


def _check(ga, plan, og, rw, seeds, fans, inst, seed):
    for r, o in zip(rw, plan.walks):
        exp, _ = og.random_walk(r["walkers"], r["length"], seed=seed, instance=r["instance"],
                                starts=r.get("starts"))
        np.testing.assert_array_equal(ga.to_numpy_u32(o), exp)
    res = plan.khop_result()
    exp = og.khop(seeds, fans, seed=seed, instances=inst)
    assert res["total"] == len(exp["dst"])
    for r in range(len(seeds)):
        s, d, _ = ga.Graph.khop_request_edges(res, r)
        e0, e1 = int(exp["edge_off"][r]), int(exp["edge_off"][r + 1])
        np.testing.assert_array_equal(s, exp["src"][e0:e1])
        np.testing.assert_array_equal(d, exp["dst"][e0:e1])


@pytest.mark.parametrize("fmt", FMTS)
def test_plan_replays_equal_oracle(ga, oracle_mod, fmt):
    import torch
    print(f"[plan] fmt={fmt}", flush=True)
    c = load_case("rmat12")
    gb, ob = c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes()
    g, og = ga.Graph.from_bytes(gb, ob, fmt), oracle_mod.Graph(gb, ob, fmt)
    rng = np.random.default_rng(12)
    rw = []
    for i in range(12):
        r = {"walkers": int(rng.integers(1, 700)), "length": int(rng.integers(1, 60)), "instance": 40 + i}
        if i % 4 == 0:
            r["starts"] = rng.integers(0, g.n, size=r["walkers"]).astype(np.uint32)
        rw.append(r)
    seeds = [rng.integers(0, g.n, size=int(k)).astype(np.uint32) for k in (1, 30, 250, 900, 7)]
    fans = [[25, 10], [15, 10], [10, 10, 5], [25, 10], [3]]
    inst = [90, 91, 92, 93, 94]
    stream = torch.cuda.Stream()  # explicit: launches and the torch reads share it
    torch.cuda.set_stream(stream)
    plan = g.plan(rw, seeds, fans, seed=31, khop_instances=inst)
    for rep in range(3):
        plan.launch()  # on torch's current stream; the reads below are ordered after it
        _check(ga, plan, og, rw, seeds, fans, inst, 31)
        g.sync()
        # eager work on the same graph between replays
        g.khop_batch([rng.integers(0, g.n, size=50).astype(np.uint32)], [[5, 5]], seed=3)
        g.random_walk(100, 10, seed=3, instance=5)
        g.sync()
    # new seed vertices written into the captured buffers
    seeds2 = [rng.integers(0, g.n, size=len(s)).astype(np.uint32) for s in seeds]
    for t, s in zip(plan.khop_seeds, seeds2):
        t.copy_(torch.from_numpy(s.view(np.int32)).to(t.device))
    torch.cuda.synchronize()
    plan.launch()
    _check(ga, plan, og, rw, seeds2, fans, inst, 31)
    g.sync()
    plan.close()
    g.close()
    torch.cuda.set_stream(torch.cuda.default_stream())


MDRW_KEYS = ("step_off", "step_src", "step_dst", "vert_off", "verts", "ind_off", "ind_src", "ind_dst")


def _clobber(ga, g, plan):
    """Overwrites every output array of an MDRW plan with 0xFF bytes, so that a replay that did not
    rewrite them shows."""
    o = ga.MdrwOutput()
    assert ga.lib().grass_plan_mdrw_output(plan._h, ga.C.byref(o)) == ga.OK
    I = plan.I
    for ptr, nbytes in ((o.step_off_dev, 8 * (I + 1)), (o.vert_off_dev, 8 * (I + 1)), (o.ind_off_dev, 8 * (I + 1)),
                        (o.step_src_dev, 4 * o.steps), (o.step_dst_dev, 4 * o.steps), (o.verts_dev, 4 * o.verts),
                        (o.ind_src_dev, 4 * o.induced), (o.ind_dst_dev, 4 * o.induced)):
        if nbytes:
            junk = np.full(nbytes, 0xFF, np.uint8)
            assert ga.lib().grass_copy_to_device(g.handle, ptr, junk.ctypes.data, nbytes) == ga.OK


@pytest.mark.parametrize("fmt", FMTS)
def test_mdrw_plan_replays_equal_oracle(ga, oracle_mod, fmt):
    import torch
    c = load_case("rmat12")
    gb, ob = c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes()
    g, og = ga.Graph.from_bytes(gb, ob, fmt), oracle_mod.Graph(gb, ob, fmt)
    rng = np.random.default_rng(21)
    cases = [([8, 100, 700, 1, 2500], 3000, [5, 6, 7, 8, 9]),
             # more instances than one induced-edge slice holds (512): the capture holds both slices
             ([int(x) for x in rng.integers(1, 30, size=600)], 80, list(range(1000, 1600)))]
    stream = torch.cuda.Stream()
    for sizes, target, inst in cases:
        print(f"[plan mdrw] fmt={fmt} instances={len(sizes)} target={target}", flush=True)
        exp = og.mdrw(sizes, target=target, seed=77, instances=inst)
        plan = g.mdrw_plan(sizes, target, seed=77, instances=inst)
        for rep in range(3):
            _clobber(ga, g, plan)
            plan.launch(stream=stream)
            stream.synchronize()
            res = plan.result()
            for k in MDRW_KEYS:
                np.testing.assert_array_equal(res[k], exp[k], err_msg=f"replay {rep}: {k}")
            # eager work on the same graph between replays
            g.mdrw([50, 60], target=500, seed=1)
            g.random_walk(100, 10, seed=3, instance=5)
            g.sync()
        plan.close()
    g.close()
