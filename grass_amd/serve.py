"""Serving sessions (include/grass.h, grass_session_*): continuous batching of random-walk and
k-hop requests on a resident graph, with memory-based admission and trace replay for latency
measurement."""
import ctypes as C

import numpy as np

from . import KHOP_MAX_HOPS, _check, lib as _core_lib

REQ_RW, REQ_KHOP = 0, 1


class Request(C.Structure):
    _fields_ = [("kind", C.c_int32), ("instance", C.c_uint32), ("walkers", C.c_uint64),
                ("length", C.c_uint32), ("starts", C.c_void_p), ("seeds", C.c_void_p),
                ("num_seeds", C.c_uint64), ("hops", C.c_uint32),
                ("fanout", C.c_uint32 * KHOP_MAX_HOPS)]


class Config(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("max_batch", C.c_uint32),
                ("max_inflight", C.c_uint32), ("seed", C.c_uint64),
                ("max_group_edges", C.c_uint64)]


class Result(C.Structure):
    _fields_ = [("kind", C.c_int32), ("status", C.c_int32), ("walks_dev", C.c_void_p),
                ("walkers", C.c_uint64), ("length", C.c_uint32), ("hops", C.c_uint32),
                ("hop_src_dev", C.c_void_p * KHOP_MAX_HOPS),
                ("hop_dst_dev", C.c_void_p * KHOP_MAX_HOPS),
                ("hop_edges", C.c_uint64 * KHOP_MAX_HOPS),
                ("submit_us", C.c_double), ("launch_us", C.c_double), ("done_us", C.c_double)]


class Counters(C.Structure):
    _fields_ = [("submitted", C.c_uint64), ("completed", C.c_uint64), ("rejected", C.c_uint64),
                ("cancelled", C.c_uint64), ("groups", C.c_uint64), ("sampled_edges", C.c_uint64),
                ("memory_in_use", C.c_uint64), ("memory_peak", C.c_uint64)]


_ready = False


def lib():
    global _ready
    L = _core_lib()
    if not _ready:
        P, U64, U32, I32, I64 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_int64
        sig = {
            "grass_session_create": (I32, [P, P, P]),
            "grass_session_submit": (I32, [P, P, P]),
            "grass_session_poll": (I32, [P, U64]),
            "grass_session_wait": (I32, [P, U64, I64]),
            "grass_session_result": (I32, [P, U64, P]),
            "grass_session_fetch": (I32, [P, U64, P, U64, P]),
            "grass_session_release": (I32, [P, U64]),
            "grass_session_cancel": (I32, [P, U64]),
            "grass_session_stats": (I32, [P, P]),
            "grass_session_destroy": (I32, [P]),
            "grass_session_replay": (I32, [P, P, P, U32, P, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _ready = True
    return L


def rw_request(walkers, length, instance, starts=None):
    """A random-walk request; `starts` (uint32 array) is kept alive on the returned object."""
    r = Request()
    r.kind, r.instance, r.walkers, r.length = REQ_RW, instance, walkers, length
    if starts is not None:
        a = np.ascontiguousarray(starts, dtype=np.uint32)
        r.starts = a.ctypes.data
        r._keep = a
    return r


def khop_request(seeds, fanout, instance):
    r = Request()
    a = np.ascontiguousarray(seeds, dtype=np.uint32)
    r.kind, r.instance, r.seeds, r.num_seeds, r.hops = REQ_KHOP, instance, a.ctypes.data, len(a), len(fanout)
    for i, f in enumerate(fanout):
        r.fanout[i] = f
    r._keep = a
    return r


class Session:
    def __init__(self, graph, memory_budget=4 << 30, max_batch=512, max_inflight=0, seed=42,
                 max_group_edges=0):
        """max_inflight / max_group_edges = 0 take the library defaults (2 groups, 6M edges)."""
        self.graph = graph
        cfg = Config(memory_budget, max_batch, max_inflight, seed, max_group_edges)
        h = C.c_void_p()
        _check(lib().grass_session_create(graph.handle, C.byref(cfg), C.byref(h)),
               "grass_session_create")
        self._h = h

    def close(self):
        if self._h:
            lib().grass_session_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, req):
        t = C.c_uint64()
        _check(lib().grass_session_submit(self._h, C.byref(req), C.byref(t)), "submit")
        return t.value

    def try_submit(self, req):
        """(status, ticket) without raising."""
        t = C.c_uint64()
        rc = lib().grass_session_submit(self._h, C.byref(req), C.byref(t))
        return rc, t.value

    def poll(self, ticket):
        return lib().grass_session_poll(self._h, ticket)

    def wait(self, ticket, timeout_us=-1):
        return lib().grass_session_wait(self._h, ticket, timeout_us)

    def result(self, ticket):
        r = Result()
        rc = lib().grass_session_result(self._h, ticket, C.byref(r))
        return rc, r

    def fetch(self, ticket):
        """Host copy of a finished request: walk matrix, or (src, dst) of all its hops."""
        rc, res = self.result(ticket)
        _check(rc, "result")
        nbytes = C.c_uint64()
        lib().grass_session_fetch(self._h, ticket, None, 0, C.byref(nbytes))
        buf = np.zeros(max(nbytes.value // 4, 1), dtype=np.uint32)
        _check(lib().grass_session_fetch(self._h, ticket, buf.ctypes.data, buf.nbytes,
                                         C.byref(nbytes)), "fetch")
        if res.kind == REQ_RW:
            return buf[: res.walkers * (res.length + 1)].reshape(res.walkers, res.length + 1)
        e = nbytes.value // 8
        counts = [int(res.hop_edges[h]) for h in range(res.hops)]
        return buf[:e], buf[e:2 * e], counts

    def release(self, ticket):
        _check(lib().grass_session_release(self._h, ticket), "release")

    def cancel(self, ticket):
        return lib().grass_session_cancel(self._h, ticket)

    def stats(self):
        c = Counters()
        _check(lib().grass_session_stats(self._h, C.byref(c)), "stats")
        return {k: getattr(c, k) for k, _ in Counters._fields_}

    def replay(self, requests, arrival_us, queue=False):
        """Open-loop replay; returns (latency_us, wall_us), or (latency_us, queue_us, wall_us)
        with queue=True (queue_us: arrival to the launch of the request's group)."""
        n = len(requests)
        arr = (Request * max(n, 1))(*requests)
        arrival = np.ascontiguousarray(arrival_us, dtype=np.float64)
        lat = np.zeros(max(n, 1), dtype=np.float64)
        qu = np.zeros(max(n, 1), dtype=np.float64)
        wall = C.c_double()
        _check(lib().grass_session_replay(self._h, arr, arrival.ctypes.data, n, lat.ctypes.data,
                                          qu.ctypes.data if queue else None, C.byref(wall)),
               "replay")
        if queue:
            return lat[:n], qu[:n], wall.value
        return lat[:n], wall.value
