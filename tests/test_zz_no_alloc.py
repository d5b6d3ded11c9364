"""GPU parity of grass_random_walk_batch_noalloc (the walk batch whose working memory is the
library's fixed device array) against grass_random_walk_batch, through the C ABI: two calls in
a row on one graph, the second with a smaller batch, each bit-exact against the allocating entry
point on the same requests; and its argument errors."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ga():
    import grass_amd
    import torch
    assert torch.cuda.is_available(), "GPU tests need a ROCm device"
    grass_amd.lib()
    return grass_amd


def _requests(rng, n, count):
    reqs = []
    for i in range(count):
        walkers = int(rng.integers(0, 700)) if i % 5 else 0  # some requests without walkers
        r = {"walkers": walkers, "length": int(rng.integers(0, 90)), "instance": 1000 + i}
        if i % 3 == 0 and walkers:
            r["starts"] = rng.integers(0, n, size=walkers).astype(np.uint32)
        reqs.append(r)
    return reqs


def _run(ga, g, reqs, seed, noalloc):
    import torch
    outs = [torch.full((r["walkers"], r["length"] + 1), -7, dtype=torch.int32, device=g._cuda())
            for r in reqs]
    arr, keep = g.rw_request_array(reqs, outs)
    stream = torch.cuda.current_stream().cuda_stream
    fn = ga.lib().grass_random_walk_batch_noalloc if noalloc else ga.lib().grass_random_walk_batch
    rc = fn(g._h, arr, len(reqs), seed, C.c_void_p(stream))
    assert rc == ga.OK, ga.strerror(rc)
    g.sync()
    del keep
    return [ga.to_numpy_u32(o) for o in outs]


@pytest.mark.parametrize("fmt", ["le", "gt"])
def test_noalloc_batch_equals_allocating_batch(ga, fmt):
    c = load_case("rmat12")
    g = ga.Graph.from_bytes(c[f"{fmt}_graph"].tobytes(), c[f"{fmt}_offset"].tobytes(), fmt)
    rng = np.random.default_rng(7)
    for count, seed in ((300, 42), (37, 9)):  # the second call is smaller than the first
        print(f"[noalloc] fmt={fmt} requests={count} seed={seed}", flush=True)
        reqs = _requests(rng, g.n, count)
        got = _run(ga, g, reqs, seed, noalloc=True)
        exp = _run(ga, g, reqs, seed, noalloc=False)
        for r, a, b in zip(reqs, got, exp):
            assert a.shape == (r["walkers"], r["length"] + 1)
            np.testing.assert_array_equal(a, b)
    g.close()


def test_noalloc_argument_errors(ga):
    import torch
    c = load_case("example")
    g = ga.Graph.from_bytes(c["gt_graph"].tobytes(), c["gt_offset"].tobytes(), "gt")
    L = ga.lib()
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    big = [{"walkers": 1, "length": 1}] * (8192 + 1)
    arr, _ = g.rw_request_array(big)
    assert L.grass_random_walk_batch_noalloc(g._h, arr, len(big), 1, stream) == ga.ECAPACITY
    arr, _ = g.rw_request_array([{"walkers": 5, "length": 3}])  # no output buffer
    assert L.grass_random_walk_batch_noalloc(g._h, arr, 1, 1, stream) == ga.EINVAL
    assert L.grass_random_walk_batch_noalloc(None, arr, 1, 1, stream) == ga.EINVAL
    assert L.grass_random_walk_batch_noalloc(g._h, None, 1, 1, stream) == ga.EINVAL
    assert L.grass_random_walk_batch_noalloc(g._h, None, 0, 1, stream) == ga.OK
    g.close()
