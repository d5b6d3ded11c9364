"""grass_amd: MI355X-native sampler for GraSS compressed graphs.

Python mirror of the C ABI in include/grass.h (libgrass_amd.so, loaded with ctypes).  Device
buffers are torch tensors (plumbing only: allocation, streams and copies); every kernel is the
library's own HIP code.  There is no CPU fallback: without the built library, or without a
GPU, the calls raise.

The reference exposes its path through the rw / ks / mrw apps (`-g <cgc_path> [-n runs]`) and
App templates; the names below follow that surface:
  Graph.open(prefix, fmt)           loadGraph + transferToGPUs   (LE/src/grass.cu:1015-1113)
  Graph.decode_csr()                BCGVertex::get_vertex, all k (LE/src/bcg.cuh:196-261)
  Graph.random_walk(...)            DeepWalkApp (LE/src/apps/randomwalk/randomWalks.cu:62-79)
  Graph.random_walk_batch(...)      many random-walk requests, one persistent launch
  Graph.khop_batch(...)             KHopApp (LE/src/apps/khop/khop.cu:1-69), batched
  Graph.mdrw(...)                   MultiRWApp (LE/src/apps/multiRW/multiRW.cu:22-97), batched
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgrass_amd.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "grass.h")

LE, GT = 0, 1
_FMT = {"le": LE, "gt": GT, LE: LE, GT: GT}

OK, EINVAL, EIO, EFORMAT, ENOMEM, EDEVICE, ECAPACITY, ECANCELLED, ETIMEOUT, ENOTREADY = range(10)


class GrassError(RuntimeError):
    def __init__(self, status, what):
        self.status = status
        super().__init__(f"{what}: {strerror(status)} (status {status})")


class _Info(C.Structure):
    _fields_ = [("n", C.c_uint64), ("m", C.c_uint64), ("max_degree", C.c_uint64),
                ("graph_bytes", C.c_uint64), ("offset_bytes", C.c_uint64),
                ("total_bits", C.c_uint64), ("ubl", C.c_uint32), ("format", C.c_int32),
                ("device", C.c_int32), ("device_bytes", C.c_uint64),
                ("gt_checkpoint", C.c_uint32)]


class RwRequest(C.Structure):
    _fields_ = [("walkers", C.c_uint64), ("length", C.c_uint32), ("instance", C.c_uint32),
                ("starts_dev", C.c_void_p), ("out_dev", C.c_void_p)]


KHOP_MAX_HOPS = 8


class KhopRequest(C.Structure):
    _fields_ = [("seeds_dev", C.c_void_p), ("num_seeds", C.c_uint64), ("instance", C.c_uint32),
                ("hops", C.c_uint32), ("fanout", C.c_uint32 * KHOP_MAX_HOPS)]


class KhopOutput(C.Structure):
    _fields_ = [("src_dev", C.c_void_p), ("dst_dev", C.c_void_p), ("capacity", C.c_uint64),
                ("offsets_dev", C.c_void_p), ("total", C.c_uint64), ("max_hops", C.c_uint32)]


class MdrwRequest(C.Structure):
    _fields_ = [("instance", C.c_uint32), ("frontier", C.c_uint32)]


class MdrwOutput(C.Structure):
    _fields_ = [("count", C.c_uint32), ("step_off_dev", C.c_void_p), ("step_src_dev", C.c_void_p),
                ("step_dst_dev", C.c_void_p), ("vert_off_dev", C.c_void_p),
                ("verts_dev", C.c_void_p), ("ind_off_dev", C.c_void_p),
                ("ind_src_dev", C.c_void_p), ("ind_dst_dev", C.c_void_p),
                ("steps", C.c_uint64), ("verts", C.c_uint64), ("induced", C.c_uint64)]


class KhopBlocks(C.Structure):
    _fields_ = [("nodes_dev", C.c_void_p), ("node_capacity", C.c_uint64), ("node_off_dev", C.c_void_p),
                ("lsrc_dev", C.c_void_p), ("ldst_dev", C.c_void_p), ("nodes", C.c_uint64)]


# grass_mdrw_array: the arrays of an MDRW result, in grass_mdrw_to_dlpack's numbering
MDRW_ARRAYS = ("step_off", "step_src", "step_dst", "vert_off", "verts", "ind_off", "ind_src", "ind_dst")
_DLTENSOR = b"dltensor"  # capsule name a DLPack consumer expects (kept alive with the module)


def _from_dlpack(ptr):
    """torch tensor over a DLManagedTensor* without a copy; torch calls its deleter."""
    torch = _torch()
    new = C.pythonapi.PyCapsule_New
    new.restype = C.py_object
    new.argtypes = [C.c_void_p, C.c_char_p, C.c_void_p]
    return torch.utils.dlpack.from_dlpack(new(ptr, _DLTENSOR, None))


class Tuning(C.Structure):
    _fields_ = [("rw_stage", C.c_uint32), ("rw_claim", C.c_uint32), ("khop_lane_scale", C.c_uint32),
                ("decode_edges_per_wave", C.c_uint32)]


class TuneReport(C.Structure):
    _fields_ = [("rw_count", C.c_uint32), ("rw_stage", C.c_uint32 * 9), ("rw_claim", C.c_uint32 * 9),
                ("rw_ms", C.c_double * 9), ("khop_count", C.c_uint32),
                ("khop_lane_scale", C.c_uint32 * 3), ("khop_ms", C.c_double * 3),
                ("decode_count", C.c_uint32), ("decode_edges_per_wave", C.c_uint32 * 5),
                ("decode_ms", C.c_double * 5)]


TUNE_RW, TUNE_KHOP, TUNE_DECODE = 1, 2, 4

MDRW_MAX_FRONTIER = 8192
MDRW_MAX_TARGET = 16384

_lib = None


def _bind_torch_runtime():
    """Imports torch (when installed) before the library is loaded.  torch's wheel carries its own
    HIP runtime and loads it by file name (libamdhip64.so); the library asks for the soname
    libamdhip64.so.7, which the dynamic loader then resolves to torch's copy.  Loaded the other
    way round, the library binds the system runtime and torch loads a second HIP and HSA runtime
    beside it, and kernels launched through one runtime on memory allocated by the other fail."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Loads libgrass_amd.so (after torch, so that both use one HIP runtime); raises if it has not
    been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C grass_amd` "
                               "(or __graft_entry__.build())")
        _bind_torch_runtime()
        L = C.CDLL(LIB_PATH)
        P, U64, U32, I32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
        sig = {
            "grass_strerror": (C.c_char_p, [I32]),
            "grass_open": (I32, [C.c_char_p, I32, I32, P]),
            "grass_open_memory": (I32, [P, U64, P, U64, I32, I32, P]),
            "grass_open_device_images": (I32, [P, U64, P, U64, I32, I32, P]),
            "grass_compress_device_open": (I32, [P, P, U64, I32, I32, P, P]),
            "grass_close": (I32, [P]),
            "grass_info": (I32, [P, P]),
            "grass_device_arrays": (I32, [P, P, P, P]),
            "grass_sync": (I32, [P, P]),
            "grass_decode_csr": (I32, [P, P, P, P]),
            "grass_random_walk": (I32, [P, P, U64, U32, U64, U32, P, P]),
            "grass_random_walk_batch": (I32, [P, P, U32, U64, P]),
            "grass_random_walk_batch_noalloc": (I32, [P, P, U32, U64, P]),
            "grass_random_walk_bytes": (I32, [P, P, U64, U32, U64, U32, P, P]),
            "grass_random_walk_batch_bytes": (I32, [P, P, U32, U64, P, P]),
            "grass_khop_bound": (I32, [P, P, U32, P]),
            "grass_khop_batch": (I32, [P, P, U32, U64, P, P]),
            "grass_mdrw": (I32, [P, P, U32, U32, U64, P, P]),
            "grass_mdrw_output_get": (I32, [P, P]),
            "grass_mdrw_fetch": (I32, [P, P, P, P, P, P, P, P, P]),
            "grass_mdrw_free": (None, [P]),
            "grass_khop_blocks_build": (I32, [P, P, U32, P, P, P]),
            "grass_mdrw_local_edges": (I32, [P, P, I32, P, P, P]),
            "grass_mdrw_to_dlpack": (I32, [P, I32, P]),
            "grass_plan_create": (I32, [P, P, U32, P, U32, P, U64, P]),
            "grass_plan_launch": (I32, [P, P]),
            "grass_plan_khop_total": (I32, [P, P]),
            "grass_plan_destroy": (I32, [P]),
            "grass_plan_create_mdrw": (I32, [P, P, U32, U32, U64, P]),
            "grass_plan_mdrw_output": (I32, [P, P]),
            "grass_copy_to_host": (I32, [P, P, P, U64]),
            "grass_copy_to_device": (I32, [P, P, P, U64]),
            "grass_get_tuning": (I32, [P, P]),
            "grass_set_tuning": (I32, [P, P]),
            "grass_autotune": (I32, [P, U32, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def strerror(status):
    return lib().grass_strerror(status).decode()


def _check(rc, what):
    if rc != OK:
        raise GrassError(rc, what)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("grass_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _stream_ptr(stream):
    if stream is None:
        return _torch().cuda.current_stream().cuda_stream
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else stream


def _dptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def to_numpy_u32(t):
    return t.cpu().numpy().view(np.uint32)


class Graph:
    """A compressed graph resident in HBM."""

    def __init__(self, handle, device):
        self._h = handle
        self.device = device
        info = _Info()
        _check(lib().grass_info(self._h, C.byref(info)), "grass_info")
        self.n = info.n
        self.m = info.m
        self.max_degree = info.max_degree
        self.graph_bytes = info.graph_bytes
        self.offset_bytes = info.offset_bytes
        self.total_bits = info.total_bits
        self.ubl = info.ubl
        self.fmt = info.format
        self.device_bytes = info.device_bytes
        self.gt_checkpoint = info.gt_checkpoint

    def info(self):
        """grass_graph_info as a dict."""
        i = _Info()
        _check(lib().grass_info(self._h, C.byref(i)), "grass_info")
        return {k: getattr(i, k) for k, _ in _Info._fields_}

    @classmethod
    def open(cls, prefix, fmt="le", device=0):
        h = C.c_void_p()
        _check(lib().grass_open(os.fsencode(prefix), _FMT[fmt], device, C.byref(h)), "grass_open")
        return cls(h, device)

    @classmethod
    def from_csr_device(cls, rowptr, col, fmt="le", device=0, stream=None):
        """Encodes a CSR held in device tensors (rowptr int64 [n+1], col int32 [m], rows sorted
        and free of duplicates) and loads the result, all on the GPU (grass_compress_device_open)."""
        h = C.c_void_p()
        _check(lib().grass_compress_device_open(C.c_void_p(rowptr.data_ptr()), C.c_void_p(col.data_ptr()),
                                                rowptr.numel() - 1, _FMT[fmt], device,
                                                C.c_void_p(_stream_ptr(stream)), C.byref(h)),
               "grass_compress_device_open")
        return cls(h, device)

    @classmethod
    def from_device_images(cls, graph_t, offset_t, fmt="le", device=0):
        """Loads .graph/.offset images held in device tensors (uint8), as from_bytes does."""
        h = C.c_void_p()
        _check(lib().grass_open_device_images(C.c_void_p(graph_t.data_ptr()), graph_t.numel(),
                                              C.c_void_p(offset_t.data_ptr()), offset_t.numel(), _FMT[fmt],
                                              device, C.byref(h)), "grass_open_device_images")
        return cls(h, device)

    @classmethod
    def from_bytes(cls, graph_bytes, offset_bytes, fmt="le", device=0):
        gb = np.ascontiguousarray(np.frombuffer(graph_bytes, dtype=np.uint8))
        ob = np.ascontiguousarray(np.frombuffer(offset_bytes, dtype=np.uint8))
        h = C.c_void_p()
        _check(lib().grass_open_memory(gb.ctypes.data, gb.nbytes, ob.ctypes.data, ob.nbytes,
                                       _FMT[fmt], device, C.byref(h)), "grass_open_memory")
        return cls(h, device)

    def close(self):
        if self._h:
            lib().grass_close(self._h)
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

    @property
    def handle(self):
        return self._h

    def sync(self, stream=None):
        _check(lib().grass_sync(self._h, C.c_void_p(_stream_ptr(stream))), "grass_sync")

    def device_arrays(self):
        w, s, d = C.c_void_p(), C.c_void_p(), C.c_void_p()
        _check(lib().grass_device_arrays(self._h, C.byref(w), C.byref(s), C.byref(d)), "arrays")
        return w.value, s.value, d.value

    def _cuda(self):
        return f"cuda:{self.device}"

    def decode_csr(self, with_col=True, stream=None):
        """Full decode; returns host (rowptr uint64[n+1], col uint32[m])."""
        torch = _torch()
        rowptr = torch.empty(self.n + 1, dtype=torch.int64, device=self._cuda())
        col = torch.empty(max(self.m, 1), dtype=torch.int32, device=self._cuda()) if with_col else None
        _check(lib().grass_decode_csr(self._h, _dptr(rowptr), _dptr(col),
                                      C.c_void_p(_stream_ptr(stream))), "grass_decode_csr")
        self.sync(stream)
        rp = rowptr.cpu().numpy().view(np.uint64)
        c = to_numpy_u32(col)[: self.m] if with_col else None
        return rp, c

    def random_walk(self, walkers, length, seed, instance=0, starts=None, out=None, stream=None,
                    sync=True):
        """Walk matrix [walkers][length+1] as a device int32 tensor (uint32 ids)."""
        torch = _torch()
        if out is None:
            out = torch.empty((walkers, length + 1), dtype=torch.int32, device=self._cuda())
        st = None
        if starts is not None:
            st = starts if hasattr(starts, "data_ptr") else torch.from_numpy(
                np.ascontiguousarray(starts, dtype=np.uint32).view(np.int32)).to(self._cuda())
        _check(lib().grass_random_walk(self._h, _dptr(st), walkers, length, seed, instance,
                                       _dptr(out), C.c_void_p(_stream_ptr(stream))),
               "grass_random_walk")
        if sync:
            self.sync(stream)
        return out

    def random_walk_bytes(self, walkers, length, seed, instance=0, starts=None, stream=None):
        """Algorithmic HBM bytes of the walks random_walk would produce."""
        torch = _torch()
        st = None
        if starts is not None:
            st = starts if hasattr(starts, "data_ptr") else torch.from_numpy(
                np.ascontiguousarray(starts, dtype=np.uint32).view(np.int32)).to(self._cuda())
        nbytes = C.c_uint64()
        _check(lib().grass_random_walk_bytes(self._h, _dptr(st), walkers, length, seed, instance,
                                             C.c_void_p(_stream_ptr(stream)), C.byref(nbytes)),
               "grass_random_walk_bytes")
        return nbytes.value

    def rw_request_array(self, requests, outs=None):
        """ctypes array of grass_rw_request (keeps device buffers alive in the returned list)."""
        torch = _torch()
        keep, arr = [], (RwRequest * max(len(requests), 1))()
        for i, r in enumerate(requests):
            st = r.get("starts")
            if st is not None and not hasattr(st, "data_ptr"):
                st = torch.from_numpy(np.ascontiguousarray(st, dtype=np.uint32).view(np.int32)).to(
                    self._cuda())
            keep.append(st)
            o = None if outs is None else outs[i]
            arr[i] = RwRequest(r["walkers"], r["length"], r.get("instance", i),
                               None if st is None else st.data_ptr(),
                               None if o is None else o.data_ptr())
        return arr, keep

    def random_walk_batch_bytes(self, requests, seed, stream=None):
        arr, keep = self.rw_request_array(requests)
        nbytes = C.c_uint64()
        _check(lib().grass_random_walk_batch_bytes(self._h, arr, len(requests), seed,
                                                   C.c_void_p(_stream_ptr(stream)),
                                                   C.byref(nbytes)),
               "grass_random_walk_batch_bytes")
        return nbytes.value

    def khop_request_array(self, seeds_list, fanouts_list, instances=None):
        """ctypes array of grass_khop_request; the returned list keeps device seeds alive."""
        torch = _torch()
        R = len(seeds_list)
        arr = (KhopRequest * max(R, 1))()
        keep = []
        for i, (s, f) in enumerate(zip(seeds_list, fanouts_list)):
            if hasattr(s, "data_ptr"):
                t = s
            else:
                t = torch.from_numpy(np.ascontiguousarray(s, dtype=np.uint32).view(np.int32)).to(
                    self._cuda())
            keep.append(t)
            fan = (C.c_uint32 * KHOP_MAX_HOPS)(*([int(x) for x in f] + [0] * (KHOP_MAX_HOPS - len(f))))
            inst = i if instances is None else int(instances[i])
            arr[i] = KhopRequest(t.data_ptr() if t.numel() else None, t.numel(), inst, len(f), fan)
        return arr, keep

    def khop_bound(self, arr, count):
        b = C.c_uint64()
        _check(lib().grass_khop_bound(self._h, arr, count, C.byref(b)), "grass_khop_bound")
        return b.value

    def khop_batch(self, seeds_list, fanouts_list, seed, instances=None, stream=None):
        """Runs the batch; returns host results {"offsets": [H][R+1], "src", "dst", "total"}."""
        torch = _torch()
        R = len(seeds_list)
        arr, keep = self.khop_request_array(seeds_list, fanouts_list, instances)
        cap = self.khop_bound(arr, R)
        H = max([len(f) for f in fanouts_list] + [0])
        src = torch.empty(max(cap, 1), dtype=torch.int32, device=self._cuda())
        dst = torch.empty(max(cap, 1), dtype=torch.int32, device=self._cuda())
        off = torch.zeros(max(H * (R + 1), 1), dtype=torch.int64, device=self._cuda())
        out = KhopOutput(src.data_ptr(), dst.data_ptr(), cap, off.data_ptr(), 0, 0)
        _check(lib().grass_khop_batch(self._h, arr, R, seed, C.byref(out),
                                      C.c_void_p(_stream_ptr(stream))), "grass_khop_batch")
        self.sync(stream)
        t = out.total
        return {"offsets": off.cpu().numpy().view(np.uint64)[: H * (R + 1)].reshape(H, R + 1),
                "src": to_numpy_u32(src)[:t], "dst": to_numpy_u32(dst)[:t], "total": t,
                "max_hops": out.max_hops}

    def mdrw(self, frontier_sizes, target, seed, instances=None, stream=None):
        """Frontier sampling, one instance per frontier size; returns host arrays
        {"step_off", "step_src", "step_dst", "vert_off", "verts", "ind_off", "ind_src",
        "ind_dst"} (the oracle's layout)."""
        I = len(frontier_sizes)
        arr = (MdrwRequest * max(I, 1))()
        for i, m in enumerate(frontier_sizes):
            arr[i] = MdrwRequest(int(instances[i]) if instances is not None else i, int(m))
        res = C.c_void_p()
        _check(lib().grass_mdrw(self._h, arr, I, target, seed, C.c_void_p(_stream_ptr(stream)),
                                C.byref(res)), "grass_mdrw")
        try:
            o = MdrwOutput()
            _check(lib().grass_mdrw_output_get(res, C.byref(o)), "grass_mdrw_output_get")
            so, vo, io = (np.zeros(I + 1, np.uint64) for _ in range(3))
            ss, sd = np.zeros(o.steps + 1, np.uint32), np.zeros(o.steps + 1, np.uint32)
            vv = np.zeros(o.verts + 1, np.uint32)
            isrc, idst = np.zeros(o.induced + 1, np.uint32), np.zeros(o.induced + 1, np.uint32)
            ptr = (lambda a: a.ctypes.data)
            _check(lib().grass_mdrw_fetch(res, ptr(so), ptr(ss), ptr(sd), ptr(vo), ptr(vv), ptr(io),
                                          ptr(isrc), ptr(idst)), "grass_mdrw_fetch")
        finally:
            lib().grass_mdrw_free(res)
        return {"step_off": so, "vert_off": vo, "ind_off": io, "step_src": ss[: o.steps],
                "step_dst": sd[: o.steps], "verts": vv[: o.verts], "ind_src": isrc[: o.induced],
                "ind_dst": idst[: o.induced]}

    def khop_minibatch(self, seeds_list, fanouts_list, seed, instances=None, stream=None):
        """A k-hop batch and its mini-batch blocks, as device tensors: "src", "dst" (int32, batch
        edges hop-major), "offsets" (int64 [H][R+1]), "nodes" (int32: each request's vertices by
        first appearance -- seeds, hop-0 destinations, hop-1 destinations, ...), "node_off"
        (int64 [R+1]) and "lsrc", "ldst" (int32 local ids of every batch edge in its request)."""
        torch = _torch()
        R = len(seeds_list)
        arr, keep = self.khop_request_array(seeds_list, fanouts_list, instances)
        cap = self.khop_bound(arr, R)
        H = max([len(f) for f in fanouts_list] + [0])
        dev = self._cuda()
        src = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        dst = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        off = torch.zeros(max(H * (R + 1), 1), dtype=torch.int64, device=dev)
        out = KhopOutput(src.data_ptr(), dst.data_ptr(), cap, off.data_ptr(), 0, 0)
        sp = C.c_void_p(_stream_ptr(stream))
        _check(lib().grass_khop_batch(self._h, arr, R, seed, C.byref(out), sp), "grass_khop_batch")
        total = out.total
        ncap = sum(int(k.numel()) for k in keep) + total
        nodes = torch.empty(max(ncap, 1), dtype=torch.int32, device=dev)
        node_off = torch.empty(R + 1, dtype=torch.int64, device=dev)
        lsrc = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        ldst = torch.empty(max(total, 1), dtype=torch.int32, device=dev)
        blk = KhopBlocks(nodes.data_ptr(), ncap, node_off.data_ptr(), lsrc.data_ptr(), ldst.data_ptr(), 0)
        _check(lib().grass_khop_blocks_build(self._h, arr, R, C.byref(out), C.byref(blk), sp),
               "grass_khop_blocks_build")
        self.sync(stream)
        return {"src": src[:total], "dst": dst[:total], "offsets": off[: H * (R + 1)].view(H, R + 1),
                "nodes": nodes[: blk.nodes], "node_off": node_off, "lsrc": lsrc[:total],
                "ldst": ldst[:total], "total": total}

    def mdrw_tensors(self, frontier_sizes, target, seed, instances=None, local=True, stream=None):
        """Frontier sampling with the result handed to torch through DLPack (no copy): a dict of
        device tensors named as mdrw()'s arrays (offsets int64, ids int32) and, with local, the
        local ids "ind_lsrc", "ind_ldst", "step_lsrc", "step_ldst" (ranks in the instance's verts)."""
        torch = _torch()
        I = len(frontier_sizes)
        arr = (MdrwRequest * max(I, 1))()
        for i, m in enumerate(frontier_sizes):
            arr[i] = MdrwRequest(int(instances[i]) if instances is not None else i, int(m))
        res = C.c_void_p()
        sp = C.c_void_p(_stream_ptr(stream))
        _check(lib().grass_mdrw(self._h, arr, I, target, seed, sp, C.byref(res)), "grass_mdrw")
        try:
            out = {}
            for k, name in enumerate(MDRW_ARRAYS):
                mt = C.c_void_p()
                _check(lib().grass_mdrw_to_dlpack(res, k, C.byref(mt)), "grass_mdrw_to_dlpack")
                out[name] = _from_dlpack(mt.value)
            if local:
                for which, pre in ((0, "ind"), (1, "step")):
                    n = out[pre + "_src"].numel()
                    ls = torch.empty(max(n, 1), dtype=torch.int32, device=self._cuda())
                    ld = torch.empty(max(n, 1), dtype=torch.int32, device=self._cuda())
                    _check(lib().grass_mdrw_local_edges(self._h, res, which, _dptr(ls), _dptr(ld), sp),
                           "grass_mdrw_local_edges")
                    out[pre + "_lsrc"], out[pre + "_ldst"] = ls[:n], ld[:n]
                self.sync(stream)
        finally:
            lib().grass_mdrw_free(res)  # the tensors keep the result's memory alive
        return out

    def tuning(self):
        t = Tuning()
        _check(lib().grass_get_tuning(self._h, C.byref(t)), "grass_get_tuning")
        return {k: getattr(t, k) for k, _ in Tuning._fields_}

    def set_tuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        t = Tuning(**cur)
        _check(lib().grass_set_tuning(self._h, C.byref(t)), "grass_set_tuning")

    def autotune(self, what=0):
        """Times every candidate of the selected knobs (TUNE_* mask, 0 = all) and keeps the fastest;
        returns (chosen tuning, per-candidate timings)."""
        t, r = Tuning(), TuneReport()
        _check(lib().grass_autotune(self._h, what, C.byref(t), C.byref(r)), "grass_autotune")
        rep = {"rw": [(r.rw_stage[i], r.rw_claim[i], r.rw_ms[i]) for i in range(r.rw_count)],
               "khop": [(r.khop_lane_scale[i], r.khop_ms[i]) for i in range(r.khop_count)],
               "decode": [(r.decode_edges_per_wave[i], r.decode_ms[i]) for i in range(r.decode_count)]}
        return {k: getattr(t, k) for k, _ in Tuning._fields_}, rep

    def plan(self, rw_requests, khop_seeds, khop_fanouts, seed, khop_instances=None):
        """Captures a fixed batch (see Plan); the buffers live as long as the plan."""
        return Plan(self, rw_requests, khop_seeds, khop_fanouts, seed, khop_instances)

    def mdrw_plan(self, frontier_sizes, target, seed, instances=None):
        """Captures one MDRW batch (see MdrwPlan)."""
        return MdrwPlan(self, frontier_sizes, target, seed, instances)

    @staticmethod
    def khop_request_edges(res, r):
        """Request r's edges, hop by hop concatenated (request-major view of a hop-major batch)."""
        off = res["offsets"]
        parts_s, parts_d, counts = [], [], []
        for h in range(off.shape[0]):
            a, b = int(off[h, r]), int(off[h, r + 1])
            parts_s.append(res["src"][a:b])
            parts_d.append(res["dst"][a:b])
            counts.append(b - a)
        cat = (lambda p: np.concatenate(p) if p else np.zeros(0, np.uint32))
        return cat(parts_s), cat(parts_d), counts

    def random_walk_batch(self, requests, seed, stream=None, sync=True):
        """requests: list of dicts {walkers, length, instance, starts (optional)}.
        Returns the list of device walk matrices."""
        torch = _torch()
        outs, keep, arr = [], [], (RwRequest * max(len(requests), 1))()
        for i, r in enumerate(requests):
            o = torch.empty((r["walkers"], r["length"] + 1), dtype=torch.int32, device=self._cuda())
            st = r.get("starts")
            if st is not None and not hasattr(st, "data_ptr"):
                st = torch.from_numpy(np.ascontiguousarray(st, dtype=np.uint32).view(np.int32)).to(
                    self._cuda())
            keep.append(st)
            outs.append(o)
            arr[i] = RwRequest(r["walkers"], r["length"], r.get("instance", i),
                               None if st is None else st.data_ptr(), o.data_ptr())
        _check(lib().grass_random_walk_batch(self._h, arr, len(requests), seed,
                                             C.c_void_p(_stream_ptr(stream))),
               "grass_random_walk_batch")
        if sync:
            self.sync(stream)
        return outs


def header_functions(path=HEADER):
    """Names of the functions include/grass.h declares."""
    import re
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(grass_[a-z_0-9]+)\s*\(", text)))


class Plan:
    """A batch of random-walk and k-hop requests captured into a HIP graph (grass_plan_*).
    `walks` are the walk matrices, `khop_seeds` the device seed buffers (rewrite their contents
    between launches for new mini-batches), `khop_src/dst/offsets` the k-hop output."""

    def __init__(self, graph, rw_requests, khop_seeds, khop_fanouts, seed, khop_instances=None):
        torch = _torch()
        self.graph = graph
        dev = graph._cuda()
        self.walks, self._keep = [], []
        self._rw = (RwRequest * max(len(rw_requests), 1))()
        for i, r in enumerate(rw_requests):
            o = torch.empty((r["walkers"], r["length"] + 1), dtype=torch.int32, device=dev)
            st = r.get("starts")
            if st is not None and not hasattr(st, "data_ptr"):
                st = torch.from_numpy(np.ascontiguousarray(st, dtype=np.uint32).view(np.int32)).to(dev)
            self._keep.append(st)
            self.walks.append(o)
            self._rw[i] = RwRequest(r["walkers"], r["length"], r.get("instance", i),
                                    st.data_ptr() if st is not None else None, o.data_ptr())
        self._kh, keep = graph.khop_request_array(khop_seeds, khop_fanouts, khop_instances)
        self.khop_seeds = keep
        R = len(khop_seeds)
        cap = graph.khop_bound(self._kh, R) if R else 0
        H = max([len(f) for f in khop_fanouts] + [0])
        self.R, self.H = R, H
        self.khop_src = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        self.khop_dst = torch.empty(max(cap, 1), dtype=torch.int32, device=dev)
        self.khop_offsets = torch.zeros(max(H * (R + 1), 1), dtype=torch.int64, device=dev)
        self._kout = KhopOutput(self.khop_src.data_ptr(), self.khop_dst.data_ptr(), cap,
                                self.khop_offsets.data_ptr(), 0, 0)
        h = C.c_void_p()
        _check(lib().grass_plan_create(graph.handle, self._rw, len(rw_requests), self._kh, R,
                                       C.byref(self._kout), seed, C.byref(h)), "grass_plan_create")
        self._h = h

    def launch(self, stream=None):
        _check(lib().grass_plan_launch(self._h, C.c_void_p(_stream_ptr(stream))), "grass_plan_launch")

    def khop_result(self):
        """Host view of the last completed launch's k-hop output (khop_batch's layout)."""
        t = C.c_uint64()
        _check(lib().grass_plan_khop_total(self._h, C.byref(t)), "grass_plan_khop_total")
        n = t.value
        off = self.khop_offsets.cpu().numpy().view(np.uint64)[: self.H * (self.R + 1)]
        return {"offsets": off.reshape(self.H, self.R + 1), "src": to_numpy_u32(self.khop_src)[:n],
                "dst": to_numpy_u32(self.khop_dst)[:n], "total": n}

    def close(self):
        if getattr(self, "_h", None):
            lib().grass_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MdrwPlan:
    """One MDRW batch captured into a HIP graph (grass_plan_create_mdrw): every launch replays it
    into the same device outputs."""

    def __init__(self, graph, frontier_sizes, target, seed, instances=None):
        self.graph = graph
        I = len(frontier_sizes)
        self.I = I
        arr = (MdrwRequest * max(I, 1))()
        for i, m in enumerate(frontier_sizes):
            arr[i] = MdrwRequest(int(instances[i]) if instances is not None else i, int(m))
        h = C.c_void_p()
        _check(lib().grass_plan_create_mdrw(graph.handle, arr, I, target, seed, C.byref(h)),
               "grass_plan_create_mdrw")
        self._h = h

    def launch(self, stream=None):
        _check(lib().grass_plan_launch(self._h, C.c_void_p(_stream_ptr(stream))), "grass_plan_launch")

    def result(self):
        """Host copies of the outputs in Graph.mdrw's layout (synchronous copies: call them after
        the launch stream has passed the replay)."""
        o = MdrwOutput()
        _check(lib().grass_plan_mdrw_output(self._h, C.byref(o)), "grass_plan_mdrw_output")
        I = self.I

        def get(ptr, count, dtype):
            a = np.zeros(max(count, 1), dtype=dtype)
            if count:
                _check(lib().grass_copy_to_host(self.graph.handle, a.ctypes.data, ptr,
                                                count * a.itemsize), "grass_copy_to_host")
            return a[:count]
        return {"step_off": get(o.step_off_dev, I + 1, np.uint64),
                "vert_off": get(o.vert_off_dev, I + 1, np.uint64),
                "ind_off": get(o.ind_off_dev, I + 1, np.uint64),
                "step_src": get(o.step_src_dev, o.steps, np.uint32),
                "step_dst": get(o.step_dst_dev, o.steps, np.uint32),
                "verts": get(o.verts_dev, o.verts, np.uint32),
                "ind_src": get(o.ind_src_dev, o.induced, np.uint32),
                "ind_dst": get(o.ind_dst_dev, o.induced, np.uint32)}

    def close(self):
        if getattr(self, "_h", None):
            lib().grass_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
