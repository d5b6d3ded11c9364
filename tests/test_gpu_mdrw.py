"""GPU parity of batched multi-dimensional random walks (frontier sampling) against the CPU
oracle on the same Philox counters, through the C ABI: traversed edges, sorted vertex sets and
induced edge lists must be identical; each instance alone equals its part of a batch."""
import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu

FMTS = ("le", "gt")
KEYS = ("step_off", "step_src", "step_dst", "vert_off", "verts", "ind_off", "ind_src", "ind_dst")


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


def _same(res, exp):
    for k in KEYS:
        np.testing.assert_array_equal(res[k], exp[k], err_msg=k)


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", ["rmat12", "degrees", "small_u"])
def test_mdrw_matches_oracle(ga, oracle_mod, case, fmt):
    print(f"[mdrw] case={case} fmt={fmt}", flush=True)
    g, og = _pair(ga, oracle_mod, case, fmt)
    # 1 entry, sub-wave, one block per lane, ragged blocks, frontier == target, frontier > target
    sizes = [1, 2, 7, 63, 64, 65, 129, 300, 1000, 1499, 1500, 2000]
    inst = [7 + 3 * i for i in range(len(sizes))]
    res = g.mdrw(sizes, target=1500, seed=99, instances=inst)
    exp = og.mdrw(sizes, target=1500, seed=99, instances=inst)
    assert int(res["step_off"][-1]) > 0
    _same(res, exp)
    g.close()


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("count", [130, 600])
def test_mdrw_many_instances(ga, oracle_mod, fmt, count):
    """Induced lists are built per slice of 512 instances with one membership bit per instance:
    130 instances use three mask words, 600 two slices (the second of 88 instances)."""
    print(f"[mdrw] many instances count={count} fmt={fmt}", flush=True)
    g, og = _pair(ga, oracle_mod, "rmat12", fmt)
    rng = np.random.default_rng(count)
    sizes = [int(x) for x in rng.integers(1, 300, size=count)]
    res = g.mdrw(sizes, target=400, seed=count)
    exp = og.mdrw(sizes, target=400, seed=count)
    _same(res, exp)
    g.close()


@pytest.mark.parametrize("fmt", FMTS)
def test_mdrw_isolated_frontiers(ga, oracle_mod, fmt):
    """Most vertices of `hub` have no edge: many small frontiers stop at step 0 (degree sum 0)."""
    g, og = _pair(ga, oracle_mod, "hub", fmt)
    sizes = [1] * 40 + [3] * 20 + [50, 400]
    res = g.mdrw(sizes, target=600, seed=5)
    exp = og.mdrw(sizes, target=600, seed=5)
    steps = np.diff(exp["step_off"].astype(np.int64))
    assert (steps == 0).any() and (steps > 0).any()
    _same(res, exp)
    g.close()


@pytest.mark.parametrize("fmt", FMTS)
def test_mdrw_largest_shapes(ga, oracle_mod, fmt):
    """GRASS_MDRW_MAX_FRONTIER entries and a GRASS_MDRW_MAX_TARGET sample (largest LDS use)."""
    g, og = _pair(ga, oracle_mod, "rmat12", fmt)
    sizes = [ga.MDRW_MAX_FRONTIER, 4000, 500]
    res = g.mdrw(sizes, target=ga.MDRW_MAX_TARGET, seed=17, instances=[3, 4, 5])
    exp = og.mdrw(sizes, target=ga.MDRW_MAX_TARGET, seed=17, instances=[3, 4, 5])
    _same(res, exp)
    g.close()


@pytest.mark.parametrize("fmt", FMTS)
def test_mdrw_alone_equals_batch(ga, oracle_mod, fmt):
    g, og = _pair(ga, oracle_mod, "rmat12", fmt)
    rng = np.random.default_rng(8)
    sizes = [int(x) for x in rng.integers(1, 900, size=24)]
    inst = list(range(500, 524))
    batch = g.mdrw(sizes, target=1000, seed=4, instances=inst)
    for i in (0, 5, 23):
        alone = g.mdrw([sizes[i]], target=1000, seed=4, instances=[inst[i]])
        for pre in ("step", "vert", "ind"):
            off = batch[f"{pre}_off"]
            a, b = int(off[i]), int(off[i + 1])
            names = {"step": ("step_src", "step_dst"), "vert": ("verts",),
                     "ind": ("ind_src", "ind_dst")}[pre]
            for k in names:
                np.testing.assert_array_equal(alone[k], batch[k][a:b], err_msg=f"{k} inst {i}")
    g.close()


def test_mdrw_rejects_bad_shapes(ga, oracle_mod):
    g, _ = _pair(ga, oracle_mod, "rmat12", "le")
    for sizes, target in (([0], 10), ([ga.MDRW_MAX_FRONTIER + 1], 10), ([5], ga.MDRW_MAX_TARGET + 1)):
        with pytest.raises(ga.GrassError) as ei:
            g.mdrw(sizes, target=target, seed=1)
        assert ei.value.status == ga.EINVAL
    res = g.mdrw([], target=100, seed=1)
    assert len(res["verts"]) == 0 and list(res["step_off"]) == [0]
    g.close()


@pytest.mark.parametrize("fmt", FMTS)
def test_mdrw_64bit_output_slots(ga, oracle_mod, fmt, monkeypatch):
    """The induced write pass keeps 32-bit output slots below 2^32 induced edges and 64-bit ones
    above; GRASS_MDRW_SLOTS64=1 selects the 64-bit variant on a small batch, which must give the
    same lists."""
    print(f"[mdrw] 64-bit output slots fmt={fmt}", flush=True)
    monkeypatch.setenv("GRASS_MDRW_SLOTS64", "1")
    g, og = _pair(ga, oracle_mod, "rmat12", fmt)
    sizes = [1, 64, 65, 300, 1000, 1500]
    res = g.mdrw(sizes, target=1500, seed=5)
    exp = og.mdrw(sizes, target=1500, seed=5)
    assert int(res["ind_off"][-1]) > 0
    _same(res, exp)
    g.close()
