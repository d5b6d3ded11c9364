"""Benchmark of the GraSS sampling path on one MI355X.

One step = one pass of the sampling hot path over one batch of synthetic requests on a
compressed R-MAT graph (a/b/c/d = .57/.19/.19/.05, edge factor 16, seed 42, graph_sort, LE or GT
compressor), inputs resident in HBM before the timed region:
  * random-walk requests: walkers U[1, 4096], length U[10, 100], Philox starts;
  * 2-hop requests (fanout [25, 10], without replacement, per-hop dedup), seed counts
    lognormal(6, 1.2) clipped to [1, 8192].
`value` is sampled edges/s of the batched path (each kind's requests in one batched call);
`plain_*` is the same batch issued one launch per request; `graph_*` the same batch replayed
from a captured HIP graph.  `roofline` is the dominant kernel's algorithmic bytes per launch over
its HIP-event-timed duration (`traffic` from the committed rocprofv3 summary of the same
workload when there is one), `cpu_baseline` the CPU oracle on a bounded sample of the same
requests.  Also reported: serving saturation QPS and p50/p99 at 70 % of it, the full decode to
CSR, and a batch of 512 multi-dimensional random walks.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--scale 24] [--fmt gt]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "sampled edges/s on compressed R-MAT-24 (RW + 2-hop), batched vs plain; p99 ms"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def rw_requests(rng, count, n_max_walkers=4096):
    reqs = []
    for i in range(count):
        reqs.append({"walkers": int(rng.integers(1, n_max_walkers + 1)),
                     "length": int(rng.integers(10, 101)), "instance": i})
    return reqs


def serving_mix(rng, n_req, n_vertices, max_seeds=4096):
    """Serving mix: 50 % random walks (length U[10,100], 1-4096 walkers), 35 % 2-hop [15,10],
    15 % 3-hop [10,10,5]; k-hop seed counts lognormal(6, 1.2) clipped to [1, max_seeds]."""
    from grass_amd import serve
    reqs, kinds = [], rng.random(n_req)
    for i, k in enumerate(kinds):
        if k < 0.5:
            reqs.append(serve.rw_request(int(rng.integers(1, 4097)), int(rng.integers(10, 101)), i))
        else:
            ns = int(np.clip(round(rng.lognormal(6.0, 1.2)), 1, max_seeds))
            fan = [15, 10] if k < 0.85 else [10, 10, 5]
            reqs.append(serve.khop_request(rng.integers(0, n_vertices, size=ns), fan, i))
    return reqs


def serving_bench(g, seed, rng, seconds=3.0, burst=6000):
    """Saturation QPS from burst replays (all requests due at once; the median of three), then p50/p99 latency of an
    open-loop Poisson replay at 70 % of it for `seconds` (long enough that a single host stall of a
    few milliseconds moves p99 little)."""
    from grass_amd import serve
    with serve.Session(g, memory_budget=24 << 30, max_batch=512, seed=seed) as s:
        warm = serving_mix(rng, 500, g.n)
        s.replay(warm, np.zeros(len(warm)))
        sats = []
        for _ in range(3):  # the median of three bursts: one burst reads a few percent either way
            reqs = serving_mix(rng, burst, g.n)
            e0 = s.stats()["sampled_edges"]
            lat, wall = s.replay(reqs, np.zeros(len(reqs)))
            sats.append((len(reqs) / (wall * 1e-6), (s.stats()["sampled_edges"] - e0) / (wall * 1e-6)))
        sats.sort()
        sat_qps, sat_eps = sats[1]
        qps = 0.7 * sat_qps
        n = int(qps * seconds)
        reqs = serving_mix(rng, n, g.n)
        arrival = np.cumsum(rng.exponential(1e6 / qps, size=n))
        e0 = s.stats()["sampled_edges"]
        lat, wall = s.replay(reqs, arrival)
        ok = lat[lat >= 0]
        out = {"saturation_qps": round(sat_qps, 1), "saturation_edges_per_s": sat_eps,
               "saturation_bursts_qps": [round(x[0], 1) for x in sats],
               "load_qps": round(qps, 1), "requests": n,
               "p50_ms": round(float(np.percentile(ok, 50)) / 1e3, 3),
               "p99_ms": round(float(np.percentile(ok, 99)) / 1e3, 3),
               "edges_per_s_at_load": (s.stats()["sampled_edges"] - e0) / (wall * 1e-6),
               "rejected": int((lat < 0).sum()), "groups": s.stats()["groups"],
               "mix": "50% RW len U[10,100] walkers U[1,4096]; 35% 2-hop [15,10]; "
                      "15% 3-hop [10,10,5]; seeds lognormal(6,1.2) clip [1,4096]"}
    return out


def random_line_ceiling(buffer_bytes, lanes, steps):
    """The GPU's rate of random 64-byte lines (libgrass_calib.so, walk-shaped chains: two dependent
    random lines per step, `lanes` chains of `steps` steps over a buffer of the resident graph's
    size), as {lines_per_s, GBps (of 64-byte lines), buffer_bytes}; None when the calibration
    library is not built."""
    import ctypes as C
    import torch
    path = os.path.join(REPO, "grass_amd", "libgrass_calib.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.grass_calib_gather.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint64, C.c_uint64,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    lines = max(buffer_bytes // 64, 1)
    buf = torch.zeros(lines * 16, dtype=torch.int32, device="cuda")
    sink = torch.zeros(8, dtype=torch.int64, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    best = None
    for it in range(4):  # first run warms up
        ms, nl = C.c_double(), C.c_uint64()
        rc = L.grass_calib_gather(C.c_void_p(buf.data_ptr()), lines, 2, steps, lanes, 11 + it, C.c_void_p(stream),
                                  C.c_void_p(sink.data_ptr()), C.byref(ms), C.byref(nl))
        if rc:
            return None
        if it:
            rate = nl.value / (ms.value * 1e-3)
            best = rate if best is None else max(best, rate)
    del buf
    torch.cuda.empty_cache()
    return {"lines_per_s": best, "GBps": round(best * 64 / 1e9, 1), "buffer_bytes": lines * 64,
            "shape": f"{lanes} chains x {steps} steps x 2 dependent random lines"}


def sync_costs(iters=1000):
    """What one grid-wide barrier inside a launch costs against one kernel boundary on a stream
    (libgrass_calib.so), at one and at two workgroups per CU: the k-hop pipeline's per-hop phases
    each need the whole previous phase, so a single-launch form would pay one barrier where the
    pipeline pays one boundary."""
    import ctypes as C
    import torch
    path = os.path.join(REPO, "grass_amd", "libgrass_calib.so")
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.grass_calib_sync_cost.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                        C.c_void_p]
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    words = 1 << 20
    buf = torch.zeros(words + 4, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    out = {}
    for per_cu in (1, 2):
        bar, bnd = C.c_double(), C.c_double()
        rc = L.grass_calib_sync_cost(C.c_void_p(buf.data_ptr()), words, cus * per_cu, iters, C.c_void_p(stream),
                                     C.byref(bar), C.byref(bnd))
        if rc:
            return None
        out[f"{per_cu}_per_cu"] = {"workgroups": cus * per_cu, "grid_barrier_us": round(bar.value, 2),
                                   "kernel_boundary_us": round(bnd.value, 2)}
    return out


def plain_four_queues(g, reqs, kseeds, kfans, seed, reps=3):
    """The step's requests one at a time (a launch group of one request each: one walk launch, or
    one k-hop pipeline), four groups in flight on their own streams and no host synchronisation
    per request: a serving session with max_batch 1 and max_inflight 4 given every request at
    once.  Isolates what batching gains over per-request launches that already overlap."""
    from grass_amd import serve
    reqs_s = [serve.rw_request(r["walkers"], r["length"], r["instance"]) for r in reqs]
    reqs_s += [serve.khop_request(k, f, 10000 + i) for i, (k, f) in enumerate(zip(kseeds, kfans))]
    with serve.Session(g, memory_budget=24 << 30, max_batch=1, max_inflight=4, seed=seed) as s:
        s.replay(reqs_s, np.zeros(len(reqs_s)))  # warm-up: sizes every slot's workspaces
        walls, edges = [], []
        for _ in range(reps):
            e0 = s.stats()["sampled_edges"]
            lat, wall = s.replay(reqs_s, np.zeros(len(reqs_s)))
            assert (lat >= 0).all()
            walls.append(wall)
            edges.append(s.stats()["sampled_edges"] - e0)
    wall = float(np.median(walls))
    return {"value": edges[0] / (wall * 1e-6), "ms_per_step": round(wall / 1e3, 3), "queues": 4,
            "edges": int(edges[0])}


def serving_rmat26(seed, rng):
    """The serving workload's own graph: R-MAT scale 26 (edge factor 16, seed 42) built, encoded
    (GT) and loaded on the GPU, then the same saturation + 70 % Poisson measurement."""
    import torch
    import grass_amd
    from grass_amd import ingest
    torch.cuda.synchronize()
    t = time.time()
    src, dst = ingest.rmat_edges_device(26, 16, seed=seed)
    rp, col = ingest.graph_sort_device(src, dst)
    del src, dst
    g = grass_amd.Graph.from_csr_device(rp, col, "gt")
    torch.cuda.synchronize()
    build_s = time.time() - t
    del rp, col
    torch.cuda.empty_cache()
    try:
        out = serving_bench(g, seed, rng)
        out.update({"scale": 26, "n": g.n, "m": g.m, "device_bytes": g.device_bytes,
                    "build_s": round(build_s, 2)})
    finally:
        g.close()
    return out


def profile_figures(scale, fmt):
    """(HBM bytes per rw_batch_kernel launch from rocprofv3 FETCH_SIZE + WRITE_SIZE, TCC hit rate,
    the profiled average launch in ms) from the committed summary of this exact workload;
    (None, None, None) when there is none."""
    import re
    path = os.path.join(REPO, "profiles", f"rw_batch_kernel_rmat{scale}_{fmt}.md")
    if not os.path.exists(path):
        return None, None, None
    text = open(path).read()
    t = re.search(r"^traffic_bytes_per_launch:\s*(\d+)", text, re.M)
    h = re.search(r"^tcc_hit_rate:\s*([0-9.]+)", text, re.M)
    d = re.search(r"^profiled_launch_us:\s*([0-9.]+)", text, re.M)
    return ((int(t.group(1)) if t else None), (float(h.group(1)) if h else None),
            (float(d.group(1)) / 1e3 if d else None))


def mdrw_bench(g, seed, instances, sptr, target=16384):
    """Batched multi-dimensional random walks: `instances` frontiers of U[500, 4000] vertices,
    16384-vertex samples, degree-proportional frontier moves, induced edge lists (one
    grass_mdrw call: walks, vertex sets, induced edges; outputs stay in HBM).  The frontier
    sizes come from their own generator, so the workload does not depend on the other phases."""
    import grass_amd
    C = grass_amd.C
    L = grass_amd.lib()
    sizes = np.random.default_rng([seed, 3]).integers(500, 4001, size=instances)
    arr = (grass_amd.MdrwRequest * instances)(*[grass_amd.MdrwRequest(i, int(m)) for i, m in enumerate(sizes)])

    def run():
        res = C.c_void_p()
        t = time.perf_counter()
        grass_amd._check(L.grass_mdrw(g.handle, arr, instances, target, seed, C.c_void_p(sptr),
                                      C.byref(res)), "grass_mdrw")
        dt = time.perf_counter() - t
        o = grass_amd.MdrwOutput()
        L.grass_mdrw_output_get(res, C.byref(o))
        L.grass_mdrw_free(res)
        return dt, o

    run()  # warm-up
    dt, o = run()
    return {"instances": instances, "frontier": "U[500,4000]", "target": target, "ms": round(dt * 1e3, 2),
            "steps": o.steps, "sampled_vertices": o.verts, "induced_edges": o.induced,
            "steps_per_s": o.steps / dt, "induced_edges_per_s": o.induced / dt}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--fmt", default="gt", choices=["le", "gt"])
    ap.add_argument("--rw-requests", type=int, default=256)
    ap.add_argument("--khop-requests", type=int, default=256)
    ap.add_argument("--threads", type=int, default=16, help="host threads for ingest")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-decode", action="store_true", help="skip the full decode to CSR")
    ap.add_argument("--no-graph", action="store_true", help="skip the captured-graph replay")
    ap.add_argument("--no-mdrw", action="store_true", help="skip the multi-dimensional random walks")
    ap.add_argument("--mdrw-instances", type=int, default=512)
    ap.add_argument("--no-tune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving (p50/p99) runs")
    ap.add_argument("--no-s26", action="store_true",
                    help="skip the serving run on the R-MAT scale-26 graph (built on the GPU)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run one 256 MiB device copy (a known-size kernel that scales rocprof "
                         "FETCH_SIZE / WRITE_SIZE readings)")
    args = ap.parse_args()
    if args.rw_requests < 1:
        ap.error("--rw-requests must be at least 1: the roofline is taken on the walk kernel")

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): requests shard "
              "across GPUs that each hold a replica of the graph, which is not measured here.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import grass_amd
    from grass_amd import ingest

    assert torch.cuda.is_available(), "bench.py needs a ROCm GPU"
    dev = torch.device("cuda:0")
    torch.cuda.set_device(0)
    seed = 42

    if args.calib_copy:
        # known-size kernels: a fill writes 256 MiB, a sum reads 256 MiB, a copy does both
        a = torch.ones(64 << 20, dtype=torch.int32, device=dev)
        s = a.sum()
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b, s
        log("calibration: fill 256 MiB, sum over 256 MiB, copy 256 MiB")

    t0 = time.time()
    src, dst = ingest.rmat_edges(args.scale, 16, seed=seed, threads=args.threads)
    log(f"R-MAT s{args.scale}: {len(src)} edges in {time.time() - t0:.1f} s")
    csr = ingest.graph_sort(src, dst, threads=args.threads)
    del src, dst
    log(f"graph_sort: n={csr.n} m={csr.m} at {time.time() - t0:.1f} s")
    IL = ingest.lib()

    def encoded(call):
        """(graph bytes, offset bytes, seconds of the C ABI call alone) of one compressor call."""
        hg, ho = grass_amd.C.c_void_p(), grass_amd.C.c_void_p()
        t1 = time.time()
        grass_amd._check(call(grass_amd.C.byref(hg), grass_amd.C.byref(ho)), "compress")
        dt = time.time() - t1
        try:
            return ingest._blob_bytes(hg), ingest._blob_bytes(ho), dt
        finally:
            IL.grass_blob_free(hg)
            IL.grass_blob_free(ho)

    fmt_code = 0 if args.fmt == "le" else 1
    gb, ob, t_host_enc = encoded(lambda hg, ho: IL.grass_compress(csr._h, fmt_code, args.threads, hg, ho))
    t_ingest = time.time() - t0
    # the same images from the device encoder (CSR uploaded first); its time includes the images'
    # copy back to host memory
    rp_h, col_h = csr.arrays()
    rp_d = torch.from_numpy(rp_h.astype(np.int64)).to(dev)
    col_d = torch.from_numpy(col_h.view(np.int32)).to(dev)
    del rp_h, col_h
    torch.cuda.synchronize()
    t_dev_enc = None
    for _ in range(2):  # the first call also loads the kernels and allocates the pinned chunks
        gb_d, ob_d, t_dev_enc = encoded(lambda hg, ho: IL.grass_compress_device(
            grass_amd.C.c_void_p(rp_d.data_ptr()), grass_amd.C.c_void_p(col_d.data_ptr()), csr.n, fmt_code,
            0, None, hg, ho))
    encode = {"host_ms": round(t_host_enc * 1e3, 1), "host_threads": args.threads,
              "device_ms": round(t_dev_enc * 1e3, 1), "identical": gb_d == gb and ob_d == ob}
    del gb_d, ob_d, rp_d, col_d
    torch.cuda.empty_cache()
    log(f"encode: {encode}")
    # the whole ingest on the device: R-MAT, graph_sort, encoder (images back in host memory)
    torch.cuda.synchronize()
    t1 = time.time()
    ds, dd = ingest.rmat_edges_device(args.scale, 16, seed=seed)
    torch.cuda.synchronize()
    t_rmat = time.time() - t1
    drp, dcol = ingest.graph_sort_device(ds, dd)
    torch.cuda.synchronize()
    t_sort = time.time() - t1 - t_rmat
    del ds, dd
    gb_d, ob_d, t_enc2 = encoded(lambda hg, ho: IL.grass_compress_device(
        grass_amd.C.c_void_p(drp.data_ptr()), grass_amd.C.c_void_p(dcol.data_ptr()), drp.numel() - 1, fmt_code,
        0, None, hg, ho))
    encode["device_pipeline"] = {"rmat_ms": round(t_rmat * 1e3, 1), "graph_sort_ms": round(t_sort * 1e3, 1),
                                 "encode_ms": round(t_enc2 * 1e3, 1),
                                 "total_s": round(time.time() - t1, 3),
                                 "host_pipeline_s": round(t_ingest, 2),
                                 "identical": gb_d == gb and ob_d == ob}
    del gb_d, ob_d
    # the same CSR encoded and loaded without the images leaving the GPU
    torch.cuda.synchronize()
    t1 = time.time()
    gdev = grass_amd.Graph.from_csr_device(drp, dcol, args.fmt)
    encode["device_pipeline"]["encode_and_load_ms"] = round((time.time() - t1) * 1e3, 1)
    gdev.close()
    del drp, dcol
    torch.cuda.empty_cache()
    log(f"device ingest: {encode['device_pipeline']}")
    log(f"ingest s{args.scale} {args.fmt}: n={csr.n} m={csr.m} graph={len(gb)/1e6:.1f} MB "
        f"offset={len(ob)/1e6:.1f} MB in {t_ingest:.1f} s")
    t0 = time.time()
    g = grass_amd.Graph.from_bytes(gb, ob, args.fmt)
    t_load = time.time() - t0
    log(f"load to HBM: {t_load:.2f} s ({(len(gb) + len(ob)) / t_load / 1e9:.2f} GB/s incl. validation)")

    tuning = None
    if not args.no_tune:
        t0 = time.time()
        tuning, trep = g.autotune()
        log(f"autotune ({time.time() - t0:.2f} s): {tuning}  {trep}")

    rng = np.random.default_rng(seed)
    reqs = rw_requests(rng, args.rw_requests)
    outs = [torch.empty((r["walkers"], r["length"] + 1), dtype=torch.int32, device=dev) for r in reqs]
    arr, keep = g.rw_request_array(reqs, outs)
    rw_edges = sum(r["walkers"] * r["length"] for r in reqs)
    # one explicit stream for every launch and every timing event (a null handle would select
    # the graph's own stream inside the library)
    stream = torch.cuda.Stream()
    torch.cuda.set_stream(stream)
    sptr = stream.cuda_stream
    L = grass_amd.lib()
    C = grass_amd.C

    # 2-hop requests: seed counts lognormal(6, 1.2) clipped to [1, 8192], fanout [25, 10]
    R2 = args.khop_requests
    kseeds = [rng.integers(0, g.n, size=int(np.clip(round(x), 1, 8192))).astype(np.uint32)
              for x in rng.lognormal(6.0, 1.2, size=R2)]
    kfans = [[25, 10]] * R2
    karr, kkeep = g.khop_request_array(kseeds, kfans, instances=list(range(10000, 10000 + R2)))
    kcap = g.khop_bound(karr, R2) if R2 else 0
    ksrc = torch.empty(max(kcap, 1), dtype=torch.int32, device=dev)
    kdst = torch.empty(max(kcap, 1), dtype=torch.int32, device=dev)
    koff = torch.zeros(max(2 * (R2 + 1), 1), dtype=torch.int64, device=dev)
    kout = grass_amd.KhopOutput(ksrc.data_ptr(), kdst.data_ptr(), kcap, koff.data_ptr(), 0, 0)
    # one request at a time (plain): its own output slice of the same buffers
    kplain = []
    for i in range(R2):
        a1 = (grass_amd.KhopRequest * 1)(karr[i])
        kplain.append(a1)
    khop_edges = 0

    def khop_batched():
        grass_amd._check(L.grass_khop_batch(g.handle, karr, R2, seed, C.byref(kout), C.c_void_p(sptr)),
                         "khop batch")

    def khop_plain():
        for a1 in kplain:
            o1 = grass_amd.KhopOutput(ksrc.data_ptr(), kdst.data_ptr(), kcap, koff.data_ptr(), 0, 0)
            grass_amd._check(L.grass_khop_batch(g.handle, a1, 1, seed, C.byref(o1), C.c_void_p(sptr)),
                             "khop plain")

    def batched_step():
        grass_amd._check(L.grass_random_walk_batch(g.handle, arr, len(reqs), seed,
                                                   C.c_void_p(sptr)), "batch")
        if R2:
            khop_batched()

    def plain_step():
        for r, o in zip(reqs, outs):
            grass_amd._check(L.grass_random_walk(g.handle, None, r["walkers"], r["length"], seed,
                                                 r["instance"], C.c_void_p(o.data_ptr()),
                                                 C.c_void_p(sptr)), "plain")
        if R2:
            khop_plain()

    def timed(fn, steps, warmup, events=False):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        evs = []
        t = time.perf_counter()
        for _ in range(steps):
            if events:
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                fn()
                e1.record(stream)
                evs.append((e0, e1))
            else:
                fn()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t
        ev_ms = [a.elapsed_time(b) for a, b in evs]
        return wall, ev_ms

    # correctness spot check of the batched path against the plain path (same outputs)
    batched_step()
    torch.cuda.synchronize()
    ref0 = outs[0].clone()
    khop_edges = kout.total
    kref = (kdst[:khop_edges].clone(), koff.clone()) if R2 else None
    if not args.no_plain:
        plain_step()
        torch.cuda.synchronize()
        assert torch.equal(ref0, outs[0]), "batched and plain walk outputs differ"
    g.sync()
    # sampled walk edges: entries that are not the padding a walk gets after a vertex without
    # neighbours (none on an undirected graph), minus the start column
    rw_edges = int(sum(int((o != g.n).sum()) - o.shape[0] for o in outs))
    log(f"2-hop batch: {R2} requests, {sum(len(s) for s in kseeds)} seeds, {khop_edges} edges")
    export = None
    if R2:
        # mini-batch blocks of the 2-hop batch (local numbering per request), timed on the stream;
        # the batch runs again first, as the plain requests above reused its output buffers
        khop_batched()
        ncap = sum(len(x) for x in kseeds) + khop_edges
        bn = torch.empty(max(ncap, 1), dtype=torch.int32, device=dev)
        bo = torch.empty(R2 + 1, dtype=torch.int64, device=dev)
        bls = torch.empty(max(khop_edges, 1), dtype=torch.int32, device=dev)
        bld = torch.empty(max(khop_edges, 1), dtype=torch.int32, device=dev)
        blk = grass_amd.KhopBlocks(bn.data_ptr(), ncap, bo.data_ptr(), bls.data_ptr(), bld.data_ptr(), 0)

        def blocks():
            grass_amd._check(L.grass_khop_blocks_build(g.handle, karr, R2, C.byref(kout), C.byref(blk),
                                                       C.c_void_p(sptr)), "khop blocks")
        _, ev_bl = timed(blocks, 5, 1, events=True)
        export = {"khop_blocks_ms": round(float(np.mean(ev_bl)), 3), "nodes": blk.nodes,
                  "edges": khop_edges}
        del bn, bo, bls, bld
        log(f"export: {export}")

    # the two request kinds are timed separately (HIP events on the launching stream), then together
    _, ev_rw = timed(lambda: L.grass_random_walk_batch(g.handle, arr, len(reqs), seed, C.c_void_p(sptr)),
                     args.steps, args.warmup, events=True)
    ev_kh = [0.0]
    if R2:
        _, ev_kh = timed(khop_batched, args.steps, args.warmup, events=True)
    wall_b, ev_b = timed(batched_step, args.steps, args.warmup, events=True)
    edges_step = rw_edges + khop_edges
    ms_step = wall_b / args.steps * 1e3
    value = edges_step * args.steps / wall_b
    log(f"batched: {ms_step:.3f} ms/step, {value/1e9:.3f} G edges/s "
        f"(rw {np.mean(ev_rw):.3f} ms for {rw_edges} edges, 2-hop {np.mean(ev_kh):.3f} ms for {khop_edges})")
    plain = None
    plain4 = None
    if not args.no_plain:
        wall_p, _ = timed(plain_step, max(1, args.steps // 2), 1)
        plain = edges_step * max(1, args.steps // 2) / wall_p
        log(f"plain: {wall_p/max(1, args.steps // 2)*1e3:.3f} ms/step, {plain/1e9:.3f} G edges/s")
        plain4 = plain_four_queues(g, reqs, kseeds, kfans, seed)
        log(f"plain on 4 queues: {plain4['value']/1e9:.3f} G edges/s ({plain4['ms_per_step']:.3f} ms/step)")

    # roofline of the dominant kernel (the persistent random-walk kernel): its launch timed with
    # HIP events on the launching stream, replayed from a captured graph so that no host
    # synchronisation sits inside the timed region (the replay holds the descriptor upload, the
    # queue reset and the kernel)
    bytes_launch = g.random_walk_batch_bytes(reqs, seed)
    rw_plan = g.plan(reqs, [], [], seed)
    _, ev_rk = timed(lambda: rw_plan.launch(stream=stream), args.steps, args.warmup, events=True)
    rw_plan.close()
    del rw_plan
    kern_ms = float(np.mean(ev_rk))
    achieved = bytes_launch / (kern_ms * 1e-3) / 1e9
    roof = {"kernel": "rw_batch_kernel", "bound": "hbm", "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "bytes_per_launch": bytes_launch,
            "bytes_per_edge": round(bytes_launch / rw_edges, 3), "launch_ms": round(kern_ms, 4)}

    # the committed summary is of the default request mix only
    default_mix = args.rw_requests == 256 and args.khop_requests == 256
    traffic, tcc_hit, prof_ms = profile_figures(args.scale, args.fmt) if default_mix else (None, None, None)
    # the counters describe the kernel that was profiled: a launch time more than 5 % away from the
    # profiled one means the kernel changed since, and the figures are left out
    if prof_ms is None or abs(kern_ms - prof_ms) > 0.05 * prof_ms:
        if traffic is not None:
            roof["traffic_note"] = (f"profiled launch {prof_ms} ms differs from this run's by more than 5 %: "
                                    "counters not reused")
        traffic, tcc_hit = None, None
    else:
        roof["profiled_launch_ms"] = prof_ms
    # what random 64-byte lines the memory system delivers, for the walk kernel's traffic (its
    # accesses are random lines, not a stream): same buffer size as the resident graph and about
    # the walk's chains and steps
    walkers = sum(r["walkers"] for r in reqs)
    ceil = random_line_ceiling(g.device_bytes, walkers, max(1, round(rw_edges / max(walkers, 1))))
    if ceil:
        roof["random_line_ceiling"] = ceil
    sync = sync_costs()
    if tcc_hit is not None:
        roof["tcc_hit_rate"] = tcc_hit
    if traffic is not None:
        roof["traffic"] = traffic
        roof["traffic_per_edge"] = round(traffic / rw_edges, 3)
        roof["traffic_GBps"] = round(traffic / (kern_ms * 1e-3) / 1e9, 2)
        if ceil:
            roof["traffic_frac_of_random_line_ceiling"] = round(traffic / (kern_ms * 1e-3) / 1e9 / ceil["GBps"], 4)

    extra = {}
    if sync:
        extra["sync_costs"] = sync
    if not args.no_graph:
        plan = g.plan(reqs, kseeds, kfans, seed, khop_instances=list(range(10000, 10000 + R2)))
        _, ev_g = timed(lambda: plan.launch(stream=stream), args.steps, args.warmup, events=True)
        pres = plan.khop_result()
        assert pres["total"] == khop_edges and torch.equal(plan.walks[0], ref0), \
            "graph replay differs from the eager batch"
        gms = float(np.mean(ev_g))
        extra["graph"] = {"ms_per_step": round(gms, 4), "value": edges_step / (gms * 1e-3),
                          "eager_ms_per_step": round(float(np.mean(ev_b)), 4)}
        log(f"graph replay: {gms:.3f} ms/step ({edges_step / (gms * 1e-3) / 1e9:.3f} G edges/s)")
        plan.close()
        del plan
        # a small steady-state step (16 walk requests of <= 256 walkers, 16 2-hop requests of
        # <= 64 seeds), where launch overhead matters: eager calls against the captured replay
        sreqs = [dict(r, walkers=min(r["walkers"], 256)) for r in reqs[:16]]
        sseeds = [k[:64] for k in kseeds[:16]]
        souts = [torch.empty((r["walkers"], r["length"] + 1), dtype=torch.int32, device=dev) for r in sreqs]
        sarr, skeep = g.rw_request_array(sreqs, souts)
        skarr, skkeep = g.khop_request_array(sseeds, kfans[:16], instances=list(range(10000, 10016)))
        skcap = g.khop_bound(skarr, 16)
        sks = torch.empty(max(skcap, 1), dtype=torch.int32, device=dev)
        skd = torch.empty(max(skcap, 1), dtype=torch.int32, device=dev)
        sko = torch.zeros(2 * 17, dtype=torch.int64, device=dev)
        skout = grass_amd.KhopOutput(sks.data_ptr(), skd.data_ptr(), skcap, sko.data_ptr(), 0, 0)

        def small_eager():
            grass_amd._check(L.grass_random_walk_batch(g.handle, sarr, 16, seed, C.c_void_p(sptr)), "small rw")
            grass_amd._check(L.grass_khop_batch(g.handle, skarr, 16, seed, C.byref(skout), C.c_void_p(sptr)),
                             "small khop")
        _, ev_se = timed(small_eager, 20, 5, events=True)
        splan = g.plan(sreqs, sseeds, kfans[:16], seed, khop_instances=list(range(10000, 10016)))
        _, ev_sg = timed(lambda: splan.launch(stream=stream), 20, 5, events=True)
        assert splan.khop_result()["total"] == skout.total
        splan.close()
        del splan
        extra["graph"]["small_step"] = {
            "rw_requests": 16, "khop_requests": 16,
            "edges": int(sum(r["walkers"] * r["length"] for r in sreqs) + skout.total),
            "eager_ms": round(float(np.median(ev_se)), 4), "graph_ms": round(float(np.median(ev_sg)), 4)}
        log(f"small step: eager {np.median(ev_se):.3f} ms, graph replay {np.median(ev_sg):.3f} ms")
    p99 = None
    if not args.no_serve:
        del outs, ksrc, kdst  # the session arena takes their HBM
        torch.cuda.empty_cache()
        extra["serving"] = serving_bench(g, seed, rng)
        p99 = extra["serving"]["p99_ms"]
        log(f"serving: {extra['serving']}")
    if not args.no_mdrw:
        extra["mdrw"] = mdrw_bench(g, seed, args.mdrw_instances, sptr)
        log(f"mdrw: {extra['mdrw']}")
    if not args.no_decode:
        rowptr = torch.empty(g.n + 1, dtype=torch.int64, device=dev)
        col = torch.empty(g.m, dtype=torch.int32, device=dev)

        def dec():
            grass_amd._check(L.grass_decode_csr(g.handle, grass_amd.C.c_void_p(rowptr.data_ptr()),
                                                grass_amd.C.c_void_p(col.data_ptr()),
                                                grass_amd.C.c_void_p(sptr)), "decode")
        _, ev_d = timed(dec, 5, 3, events=True)
        dms = float(np.mean(ev_d))
        dbytes = len(gb) + len(ob) + 8 * (g.n + 1) + 4 * g.m
        extra["decode"] = {"ms": round(dms, 3), "decoded_bytes_per_s": (len(gb) + len(ob)) / (dms * 1e-3),
                           "algorithmic_GBps": round(dbytes / (dms * 1e-3) / 1e9, 2),
                           "frac": round(dbytes / (dms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"decode CSR: {dms:.3f} ms  {extra['decode']}")
        del rowptr, col
        torch.cuda.empty_cache()

    if not args.no_serve and not args.no_s26:
        extra["serving_s26"] = serving_rmat26(seed, np.random.default_rng([seed, 26]))
        log(f"serving s26: {extra['serving_s26']}")
    csr_bytes = 8 * (g.n + 1) + 4 * g.m
    extra["resident"] = {"device_bytes": g.device_bytes, "bytes_per_edge": round(g.device_bytes / g.m, 3),
                         "csr_u64_u32_bytes": csr_bytes, "over_csr": round(g.device_bytes / csr_bytes, 4),
                         "file_bytes": len(gb) + len(ob)}

    cpu = None
    if not args.no_cpu:
        import oracle
        og = oracle.Graph(gb, ob, args.fmt)
        # the host's cores are shared between its GPUs: one GPU's share is OMP_NUM_THREADS (16 on
        # these boxes), whatever nproc counts for the whole host
        nproc = os.cpu_count() or 1
        try:
            affinity = len(os.sched_getaffinity(0))
        except (AttributeError, OSError):
            affinity = nproc
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        cores = max(1, min(affinity, omp or 16))
        # bounded sample of the same step: a prefix of the walk requests and a prefix of the 2-hop
        # requests, each sized to half of ~cpu-seconds of oracle work
        t = time.perf_counter()
        pe = 0
        for r in reqs[:4]:
            og.random_walk(r["walkers"], r["length"], seed, instance=r["instance"], nthreads=cores)
            pe += r["walkers"] * r["length"]
        rw_rate = pe / (time.perf_counter() - t)
        t = time.perf_counter()
        pk = og.khop(kseeds[:4], kfans[:4], seed=seed, instances=list(range(10000, 10004)), nthreads=cores)
        kh_rate = max(len(pk["dst"]), 1) / (time.perf_counter() - t)
        half = args.cpu_seconds / 2
        rs, re_ = [], 0
        for r in reqs:
            if re_ >= rw_rate * half:
                break
            rs.append(r)
            re_ += r["walkers"] * r["length"]
        nk = 1
        while nk < R2 and sum(len(x) for x in kseeds[:nk]) * 260 < kh_rate * half:
            nk += 1
        def sample_once():
            t0 = time.perf_counter()
            for r in rs:
                og.random_walk(r["walkers"], r["length"], seed, instance=r["instance"], nthreads=cores)
            kres = og.khop(kseeds[:nk], kfans[:nk], seed=seed, instances=list(range(10000, 10000 + nk)),
                           nthreads=cores) if R2 else {"dst": []}
            return time.perf_counter() - t0, len(kres["dst"])

        sample_once()  # warm-up
        runs = [sample_once() for _ in range(5)]
        dt = float(np.median([x[0] for x in runs]))
        ke = runs[0][1]
        model = "unknown"
        try:
            with open("/proc/cpuinfo") as f:
                model = next(l.split(":", 1)[1].strip() for l in f if l.startswith("model name"))
        except (OSError, StopIteration):
            pass
        cpu = {"value": (re_ + ke) / dt, "unit": "edges/s", "cores": cores, "kind": "port",
               "sample": f"first {len(rs)} of the {len(reqs)} random-walk requests ({re_} edges) and "
                         f"first {nk} of the {R2} 2-hop requests ({ke} edges), the oracle's C "
                         f"restatement with OpenMP over walkers / requests; median of 5 runs after "
                         f"a warm-up, {dt:.2f} s each",
               "cpu_model": model, "nproc": nproc, "affinity_cpus": affinity, "omp_num_threads": omp or None}
        # one core: the first walk requests that take about a second, median of 5 runs
        t = time.perf_counter()
        r1, e1 = [], 0
        for r in reqs:
            og.random_walk(r["walkers"], r["length"], seed, instance=r["instance"], nthreads=1)
            r1.append(r)
            e1 += r["walkers"] * r["length"]
            if time.perf_counter() - t > 1.0:
                break
        t1 = []
        for _ in range(5):
            t = time.perf_counter()
            for r in r1:
                og.random_walk(r["walkers"], r["length"], seed, instance=r["instance"], nthreads=1)
            t1.append(time.perf_counter() - t)
        cpu["single_core_value"] = e1 / float(np.median(t1))
        # the full decode to CSR on the same cores (decoded .graph + .offset bytes per second)
        t = time.perf_counter()
        og.decode_csr(nthreads=cores)
        cpu["decode_bytes_per_s"] = (len(gb) + len(ob)) / (time.perf_counter() - t)
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": METRIC, "value": value, "unit": "edges/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32",
        "data": f"synthetic R-MAT s{args.scale} ef16 (.57,.19,.19,.05) seed 42, graph_sort, "
                f"{args.fmt.upper()} compressor; Philox-drawn walk starts, uniform 2-hop seeds",
        "config": {"workload": f"batched random-walk + 2-hop [25,10] requests on R-MAT-{args.scale} "
                               f"({args.fmt.upper()})", "scale": args.scale, "edge_factor": 16,
                   "format": args.fmt, "n": g.n, "m": g.m, "rw_requests": len(reqs),
                   "rw_edges_per_step": rw_edges, "khop_requests": R2,
                   "khop_seeds": int(sum(len(s) for s in kseeds)),
                   "khop_edges_per_step": khop_edges},
        "rw_ms": round(float(np.mean(ev_rw)), 4), "khop_ms": round(float(np.mean(ev_kh)), 4),
        "plain_value": plain, "batched_over_plain": (value / plain) if plain else None,
        "plain_4q": plain4, "batched_over_plain_4q": (value / plain4["value"]) if plain4 else None,
        "p99_ms": p99,
        "roofline": roof, "cpu_baseline": cpu,
        "ingest_s": round(t_ingest, 2), "load_s": round(t_load, 3), "encode": encode, "export": export,
        "tuning": tuning,
        "gt_checkpoint": g.gt_checkpoint,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)
    g.close()


if __name__ == "__main__":
    main()
