"""GPU parity of the loader, the decode-to-CSR kernel and the random-walk kernels, through the
C ABI (grass_amd is a ctypes mirror of include/grass.h).  Bit-exact against the golden CSR
(the reference compressors' input) and against the CPU oracle on the same Philox counters."""
import numpy as np
import pytest

from conftest import CASES, csr_of, load_case

pytestmark = pytest.mark.gpu

FMTS = ("le", "gt")


@pytest.fixture(scope="module")
def ga():
    import grass_amd
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    grass_amd.lib()
    return grass_amd


def _open(ga, case, fmt):
    c = load_case(case)
    return ga.Graph.from_bytes(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt), c


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", CASES)
def test_load_and_decode_golden(ga, case, fmt):
    print(f"[decode] case={case} fmt={fmt}", flush=True)
    g, c = _open(ga, case, fmt)
    erp, ecol = csr_of(c)
    assert g.n == int(c["n"]) and g.m == len(ecol)
    assert g.max_degree == int(np.diff(erp.astype(np.int64)).max())
    rowptr, col = g.decode_csr()
    np.testing.assert_array_equal(rowptr, erp)
    np.testing.assert_array_equal(col, ecol)
    g.close()


def test_open_files_and_errors(ga, tmp_path):
    c = load_case("degrees")
    for fmt in FMTS:
        (tmp_path / f"g{fmt}.graph").write_bytes(c[f"{fmt}_graph"].tobytes())
        (tmp_path / f"g{fmt}.offset").write_bytes(c[f"{fmt}_offset"].tobytes())
        g = ga.Graph.open(str(tmp_path / f"g{fmt}"), fmt)
        assert g.n == int(c["n"])
        g.close()
    with pytest.raises(ga.GrassError) as e:
        ga.Graph.open(str(tmp_path / "missing"), "le")
    assert e.value.status == ga.EIO
    gb, ob = c["le_graph"].tobytes(), c["le_offset"].tobytes()
    bad = [
        (gb, b"\x00" + ob[1:]),                 # ubl = 0 (the reference divides by it)
        (gb[:-1], ob),                           # truncated .graph
        (gb + b"\x00", ob),                      # .graph longer than its records
        (b"\x00" * len(gb), ob),                 # all-zero gamma windows
        (gb, ob[:1] + b"\xff" * (len(ob) - 1)),  # record lengths beyond the .graph
    ]
    for gbad, obad in bad:
        with pytest.raises(ga.GrassError) as e:
            ga.Graph.from_bytes(gbad, obad, "le")
        assert e.value.status == ga.EFORMAT


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", ["degrees", "rmat12", "hub", "small_u"])
def test_random_walk_matches_oracle(ga, oracle_mod, case, fmt):
    print(f"[rw] case={case} fmt={fmt}", flush=True)
    g, c = _open(ga, case, fmt)
    og = oracle_mod.Graph(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    W, L = 5000, 40
    out = ga.to_numpy_u32(g.random_walk(W, L, seed=1234, instance=7))
    exp, _ = og.random_walk(W, L, seed=1234, instance=7)
    np.testing.assert_array_equal(out, exp)
    starts = np.random.default_rng(0).integers(0, g.n, size=777).astype(np.uint32)
    out = ga.to_numpy_u32(g.random_walk(777, 13, seed=99, instance=2, starts=starts))
    exp, _ = og.random_walk(777, 13, seed=99, instance=2, starts=starts)
    np.testing.assert_array_equal(out, exp)
    g.close()


@pytest.mark.parametrize("fmt", FMTS)
def test_random_walk_batch_equals_alone(ga, oracle_mod, fmt):
    g, c = _open(ga, "rmat12", fmt)
    og = oracle_mod.Graph(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    rng = np.random.default_rng(5)
    reqs = []
    for i in range(40):
        r = {"walkers": int(rng.integers(1, 300)), "length": int(rng.integers(1, 100)),
             "instance": 1000 + i}
        if i % 3 == 0:
            r["starts"] = rng.integers(0, g.n, size=r["walkers"]).astype(np.uint32)
        reqs.append(r)
    outs = g.random_walk_batch(reqs, seed=77)
    for r, o in zip(reqs, outs):
        exp, _ = og.random_walk(r["walkers"], r["length"], seed=77, instance=r["instance"],
                                starts=r.get("starts"))
        np.testing.assert_array_equal(ga.to_numpy_u32(o), exp)
        alone = g.random_walk(r["walkers"], r["length"], seed=77, instance=r["instance"],
                              starts=r.get("starts"))
        np.testing.assert_array_equal(ga.to_numpy_u32(alone), exp)
    g.close()


def test_bad_start_reported(ga):
    g, c = _open(ga, "small_u", "le")
    g.random_walk(4, 3, seed=1, starts=np.array([0, 1, g.n, 2], np.uint32), sync=False)
    with pytest.raises(ga.GrassError) as e:
        g.sync()
    assert e.value.status == ga.EINVAL
    g.sync()  # cleared
    g.close()


@pytest.mark.parametrize("fmt", FMTS)
def test_rmat16_decode_and_walks(ga, oracle_mod, fmt):
    """The CPU-plain config: R-MAT s16, ef 16, graph_sort, 4096 walks x 80 from Philox starts."""
    from grass_amd import ingest
    gb, ob, csr = ingest.rmat_graph(16, 16, fmt=fmt, seed=42)
    rowptr, col = csr.arrays()
    g = ga.Graph.from_bytes(gb, ob, fmt)
    rp, cl = g.decode_csr()
    np.testing.assert_array_equal(rp, rowptr)
    np.testing.assert_array_equal(cl, col)
    og = oracle_mod.Graph(gb, ob, fmt)
    out = ga.to_numpy_u32(g.random_walk(4096, 80, seed=42, instance=0))
    exp, _ = og.random_walk(4096, 80, seed=42, instance=0)
    np.testing.assert_array_equal(out, exp)
    g.close()


@pytest.mark.parametrize("fmt", FMTS)
@pytest.mark.parametrize("case", ["degrees", "rmat12", "hub"])
def test_walk_bytes_match_oracle(ga, oracle_mod, case, fmt):
    """The algorithmic byte count the roofline uses is the oracle's count of the same walks."""
    g, c = _open(ga, case, fmt)
    og = oracle_mod.Graph(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    og.set_gt_checkpoint(g.info()["gt_checkpoint"])
    _, exp = og.random_walk(3000, 25, seed=5, instance=1)
    assert g.random_walk_bytes(3000, 25, seed=5, instance=1) == exp
    starts = np.random.default_rng(1).integers(0, g.n, size=500).astype(np.uint32)
    _, exp_s = og.random_walk(500, 9, seed=5, instance=2, starts=starts)
    assert g.random_walk_bytes(500, 9, seed=5, instance=2, starts=starts) == exp_s
    reqs = [{"walkers": 3000, "length": 25, "instance": 1},
            {"walkers": 500, "length": 9, "instance": 2, "starts": starts}]
    assert g.random_walk_batch_bytes(reqs, seed=5) == exp + exp_s
    g.close()
