"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by the
cpu_baseline leg of bench.py, to check results against.  The product (grass_amd) never
imports it.
The C restatement and its reference citations are in grass_oracle.c.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
LE, GT = 0, 1
MAX_HOPS = 8
_FMT = {"le": LE, "gt": GT, LE: LE, GT: GT}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P, U64, U32, I32, I64 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_int64
        sig = {
            "or_pack_words": (None, [P, U64, P]),
            "or_parse_offsets": (I32, [P, U64, P, P, P]),
            "or_open": (I32, [P, U64, P, U64, I32, P]),
            "or_close": (None, [P]),
            "or_set_gt_checkpoint": (None, [P, U32]),
            "or_num_vertices": (U64, [P]),
            "or_total_bits": (U64, [P]),
            "or_record_start": (U64, [P, U64]),
            "or_header": (I32, [P, U64, P]),
            "or_degree": (I64, [P, U64]),
            "or_get": (I64, [P, U64, U64]),
            "or_payload_bits": (I64, [P, U64, U64]),
            "or_decode_csr": (I32, [P, P, P, I32]),
            "or_philox4x32_10": (None, [P, P, P]),
            "or_uniform": (U64, [U64, U32, U32, U32, U32, U64]),
            "or_rw": (I32, [P, U64, U32, P, U64, U32, P, I32, P]),
            "or_khop": (I32, [P, U64, U32, P, P, P, P, P, I32, P]),
            "or_khop_result_sizes": (None, [P, P, P, P, P]),
            "or_khop_result_edges": (None, [P, P, P]),
            "or_khop_result_free": (None, [P]),
            "or_mdrw": (I32, [P, U64, U32, P, P, U32, I32, P]),
            "or_mdrw_result_sizes": (None, [P, P, P, P, P]),
            "or_mdrw_result_data": (None, [P, P, P, P, P, P]),
            "or_mdrw_result_free": (None, [P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise OracleError(f"{what} failed with oracle status {rc}")


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def uniform(seed, c0, c1, c2, c3, bound):
    return int(lib().or_uniform(seed, c0, c1, c2, c3, bound))


def parse_offsets(offset_bytes):
    b = np.ascontiguousarray(np.frombuffer(bytes(offset_bytes), dtype=np.uint8))
    ubl = C.c_uint32()
    nf = C.c_uint64()
    _check(lib().or_parse_offsets(_ptr(b), len(b), C.byref(ubl), C.byref(nf), None), "parse")
    lens = np.zeros(max(nf.value, 1), dtype=np.uint64)
    _check(lib().or_parse_offsets(_ptr(b), len(b), C.byref(ubl), C.byref(nf), _ptr(lens)), "parse")
    return ubl.value, nf.value, lens[: nf.value]


class Graph:
    """A compressed graph opened by the oracle (in memory)."""

    def __init__(self, graph_bytes, offset_bytes, fmt):
        g = np.ascontiguousarray(np.frombuffer(bytes(graph_bytes), dtype=np.uint8))
        o = np.ascontiguousarray(np.frombuffer(bytes(offset_bytes), dtype=np.uint8))
        h = C.c_void_p()
        _check(lib().or_open(_ptr(g), len(g), _ptr(o), len(o), _FMT[fmt], C.byref(h)), "open")
        self._h = h
        self.fmt = _FMT[fmt]
        self.n = int(lib().or_num_vertices(h))
        self.total_bits = int(lib().or_total_bits(h))

    @classmethod
    def from_files(cls, prefix, fmt):
        with open(prefix + ".graph", "rb") as f:
            gb = f.read()
        with open(prefix + ".offset", "rb") as f:
            ob = f.read()
        return cls(gb, ob, fmt)

    def set_gt_checkpoint(self, interval):
        """Byte model only: count GT reads through a checkpoint every `interval` positions."""
        lib().or_set_gt_checkpoint(self._h, interval)

    def close(self):
        if self._h:
            lib().or_close(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def record_start(self, v):
        return int(lib().or_record_start(self._h, v))

    def header(self, v):
        out = np.zeros(9, dtype=np.int64)
        _check(lib().or_header(self._h, v, _ptr(out)), "header")
        keys = ("d", "f1", "f2", "first_pos", "C", "q", "R", "payload", "header_bits")
        return dict(zip(keys, (int(x) for x in out)))

    def degree(self, v):
        return int(lib().or_degree(self._h, v))

    def get(self, v, k):
        return int(lib().or_get(self._h, v, k))

    def payload_bits(self, v, k):
        return int(lib().or_payload_bits(self._h, v, k))

    def decode_csr(self, nthreads=0):
        rowptr = np.zeros(self.n + 1, dtype=np.uint64)
        _check(lib().or_decode_csr(self._h, _ptr(rowptr), None, nthreads), "decode")
        col = np.zeros(max(int(rowptr[-1]), 1), dtype=np.uint32)
        _check(lib().or_decode_csr(self._h, _ptr(rowptr), _ptr(col), nthreads), "decode")
        return rowptr, col[: int(rowptr[-1])]

    def random_walk(self, walkers, length, seed, instance=0, starts=None, nthreads=0):
        out = np.zeros((walkers, length + 1), dtype=np.uint32)
        st = None if starts is None else np.ascontiguousarray(starts, dtype=np.uint32)
        nbytes = C.c_uint64()
        _check(lib().or_rw(self._h, seed, instance, _ptr(st), walkers, length, _ptr(out), nthreads,
                           C.byref(nbytes)), "rw")
        return out, nbytes.value

    def khop(self, seeds_list, fanouts_list, seed, instances=None, nthreads=0):
        """seeds_list: list of uint32 arrays; fanouts_list: list of fanout lists."""
        R = len(seeds_list)
        off = np.zeros(R + 1, dtype=np.uint64)
        for i, s in enumerate(seeds_list):
            off[i + 1] = off[i] + len(s)
        seeds = np.ascontiguousarray(np.concatenate([np.asarray(s, dtype=np.uint32) for s in seeds_list])
                                     if R else np.zeros(0, np.uint32))
        inst = np.ascontiguousarray(np.arange(R, dtype=np.uint32) if instances is None else
                                    np.asarray(instances, dtype=np.uint32))
        hops = np.array([len(f) for f in fanouts_list], dtype=np.uint32)
        fan = np.zeros(R * MAX_HOPS + 1, dtype=np.uint32)
        for i, f in enumerate(fanouts_list):
            fan[i * MAX_HOPS: i * MAX_HOPS + len(f)] = f
        res = C.c_void_p()
        _check(lib().or_khop(self._h, seed, R, _ptr(off), _ptr(seeds), _ptr(inst), _ptr(hops),
                             _ptr(fan), nthreads, C.byref(res)), "khop")
        try:
            eoff = np.zeros(R + 1, dtype=np.uint64)
            hc = np.zeros(R * MAX_HOPS + 1, dtype=np.uint64)
            tot = C.c_uint64()
            nbytes = C.c_uint64()
            lib().or_khop_result_sizes(res, _ptr(eoff), _ptr(hc), C.byref(tot), C.byref(nbytes))
            src = np.zeros(tot.value + 1, dtype=np.uint32)
            dst = np.zeros(tot.value + 1, dtype=np.uint32)
            lib().or_khop_result_edges(res, _ptr(src), _ptr(dst))
        finally:
            lib().or_khop_result_free(res)
        t = tot.value
        return {"edge_off": eoff, "hop_count": hc[: R * MAX_HOPS].reshape(R, MAX_HOPS),
                "src": src[:t], "dst": dst[:t], "bytes": nbytes.value}

    def mdrw(self, frontier_sizes, target, seed, instances=None, nthreads=0):
        I = len(frontier_sizes)
        fs = np.ascontiguousarray(frontier_sizes, dtype=np.uint32)
        inst = np.ascontiguousarray(np.arange(I, dtype=np.uint32) if instances is None else
                                    np.asarray(instances, dtype=np.uint32))
        res = C.c_void_p()
        _check(lib().or_mdrw(self._h, seed, I, _ptr(inst), _ptr(fs), target, nthreads,
                             C.byref(res)), "mdrw")
        try:
            so = np.zeros(I + 1, dtype=np.uint64)
            vo = np.zeros(I + 1, dtype=np.uint64)
            io = np.zeros(I + 1, dtype=np.uint64)
            nbytes = C.c_uint64()
            lib().or_mdrw_result_sizes(res, _ptr(so), _ptr(vo), _ptr(io), C.byref(nbytes))
            ss = np.zeros(int(so[-1]) + 1, dtype=np.uint32)
            sd = np.zeros(int(so[-1]) + 1, dtype=np.uint32)
            vv = np.zeros(int(vo[-1]) + 1, dtype=np.uint32)
            isrc = np.zeros(int(io[-1]) + 1, dtype=np.uint32)
            idst = np.zeros(int(io[-1]) + 1, dtype=np.uint32)
            lib().or_mdrw_result_data(res, _ptr(ss), _ptr(sd), _ptr(vv), _ptr(isrc), _ptr(idst))
        finally:
            lib().or_mdrw_result_free(res)
        return {"step_off": so, "vert_off": vo, "ind_off": io,
                "step_src": ss[: int(so[-1])], "step_dst": sd[: int(so[-1])],
                "verts": vv[: int(vo[-1])], "ind_src": isrc[: int(io[-1])],
                "ind_dst": idst[: int(io[-1])], "bytes": nbytes.value}
