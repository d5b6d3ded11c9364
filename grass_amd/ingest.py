"""Ingest through libgrass_amd.so (include/grass_ingest.h): R-MAT edge lists, graph_sort
preprocessing and the LE / GT compressors, on the host and (the *_device calls) on a GPU."""
import ctypes as C
import os

import numpy as np

from . import _check, lib as _core_lib, _FMT

_ready = False


def lib():
    global _ready
    L = _core_lib()
    if not _ready:
        P, U64, U32, I32, D = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int, C.c_double
        sig = {
            "grass_rmat_edges": (I32, [U32, U32, D, D, D, U64, P, P, I32]),
            "grass_graph_sort": (I32, [P, P, U64, I32, I32, P]),
            "grass_csr_from_sorted_edges": (I32, [P, P, U64, P]),
            "grass_csr_info": (I32, [P, P, P]),
            "grass_csr_arrays": (I32, [P, P, P]),
            "grass_csr_free": (None, [P]),
            "grass_compress": (I32, [P, I32, I32, P, P]),
            "grass_compress_device": (I32, [P, P, U64, I32, I32, P, P, P]),
            "grass_rmat_edges_device": (I32, [U32, U32, C.c_double, C.c_double, C.c_double, U64, I32, P, P,
                                              P]),
            "grass_graph_sort_device": (I32, [P, P, U64, I32, I32, P, P, U64, P, U64, P, P]),
            "grass_blob_data": (I32, [P, P, P]),
            "grass_blob_free": (None, [P]),
            "grass_write_files": (I32, [C.c_char_p, P, P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _ready = True
    return L


def rmat_edges(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=42, threads=0):
    """R-MAT edge list (uint32 src, dst) of edge_factor << scale edges."""
    m = edge_factor << scale
    src = np.empty(m, dtype=np.uint32)
    dst = np.empty(m, dtype=np.uint32)
    _check(lib().grass_rmat_edges(scale, edge_factor, a, b, c, seed, src.ctypes.data,
                                  dst.ctypes.data, threads), "grass_rmat_edges")
    return src, dst


class Csr:
    """Host CSR owned by the library."""

    def __init__(self, handle):
        self._h = handle
        n, m = C.c_uint64(), C.c_uint64()
        _check(lib().grass_csr_info(handle, C.byref(n), C.byref(m)), "grass_csr_info")
        self.n, self.m = n.value, m.value

    def arrays(self):
        """Copies of (rowptr uint64[n+1], col uint32[m])."""
        rp, cl = C.c_void_p(), C.c_void_p()
        _check(lib().grass_csr_arrays(self._h, C.byref(rp), C.byref(cl)), "grass_csr_arrays")
        rowptr = np.ctypeslib.as_array(C.cast(rp, C.POINTER(C.c_uint64)), (self.n + 1,)).copy()
        col = (np.ctypeslib.as_array(C.cast(cl, C.POINTER(C.c_uint32)), (self.m,)).copy()
               if self.m else np.zeros(0, np.uint32))
        return rowptr, col

    def compress(self, fmt="le", threads=0):
        """(.graph bytes, .offset bytes) written by the LE or GT compressor."""
        g, o = C.c_void_p(), C.c_void_p()
        _check(lib().grass_compress(self._h, _FMT[fmt], threads, C.byref(g), C.byref(o)),
               "grass_compress")
        try:
            return _blob_bytes(g), _blob_bytes(o)
        finally:
            lib().grass_blob_free(g)
            lib().grass_blob_free(o)

    def close(self):
        if self._h:
            lib().grass_csr_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def compress_device(rowptr, col, fmt="le", device=0, stream=None):
    """(.graph bytes, .offset bytes) of the LE or GT compressor, encoded on the GPU from a CSR in
    device memory: rowptr (int64/uint64, n + 1) and col (int32/uint32, m) torch tensors on `device`,
    rows sorted without duplicates as graph_sort writes them."""
    from . import _stream_ptr
    n = rowptr.numel() - 1
    g, o = C.c_void_p(), C.c_void_p()
    _check(lib().grass_compress_device(C.c_void_p(rowptr.data_ptr()), C.c_void_p(col.data_ptr()), n,
                                       _FMT[fmt], device, C.c_void_p(_stream_ptr(stream)), C.byref(g),
                                       C.byref(o)), "grass_compress_device")
    try:
        return _blob_bytes(g), _blob_bytes(o)
    finally:
        lib().grass_blob_free(g)
        lib().grass_blob_free(o)


def rmat_edges_device(scale, edge_factor=16, a=0.57, b=0.19, c=0.19, seed=1, device=0, stream=None):
    """grass_rmat_edges on the GPU: (src, dst) int32 device tensors of edge_factor << scale ids."""
    from . import _stream_ptr, _torch
    torch = _torch()
    m = edge_factor << scale
    src = torch.empty(m, dtype=torch.int32, device=f"cuda:{device}")
    dst = torch.empty(m, dtype=torch.int32, device=f"cuda:{device}")
    _check(lib().grass_rmat_edges_device(scale, edge_factor, a, b, c, seed, device,
                                         C.c_void_p(_stream_ptr(stream)), C.c_void_p(src.data_ptr()),
                                         C.c_void_p(dst.data_ptr())), "grass_rmat_edges_device")
    return src, dst


def graph_sort_device(src, dst, directed=False, device=0, stream=None):
    """grass_graph_sort on the GPU over device tensors of ids: (rowptr int64 [n+1], col int32 [m'])
    device tensors."""
    from . import _stream_ptr, _torch
    torch = _torch()
    m = src.numel()
    n, mo = C.c_uint64(), C.c_uint64()
    cap_r = (int(torch.maximum(src.max(), dst.max()).item()) + 2) if m else 1
    cap_c = max(m if directed else 2 * m, 1)
    rowptr = torch.empty(cap_r, dtype=torch.int64, device=src.device)
    col = torch.empty(cap_c, dtype=torch.int32, device=src.device)
    _check(lib().grass_graph_sort_device(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), m,
                                         int(directed), device, C.c_void_p(_stream_ptr(stream)),
                                         C.c_void_p(rowptr.data_ptr()), cap_r, C.c_void_p(col.data_ptr()),
                                         cap_c, C.byref(n), C.byref(mo)), "grass_graph_sort_device")
    return rowptr[: n.value + 1], col[: mo.value]


def _blob_bytes(h):
    """The blob as bytes (copied through numpy: ctypes.string_at takes a C int size and cannot
    copy 2 GiB or more)."""
    p, n = C.c_void_p(), C.c_uint64()
    _check(lib().grass_blob_data(h, C.byref(p), C.byref(n)), "grass_blob_data")
    if not n.value:
        return b""
    view = np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_uint8)), (n.value,))
    return view.tobytes()


def graph_sort(src, dst, directed=False, threads=0):
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    h = C.c_void_p()
    _check(lib().grass_graph_sort(src.ctypes.data, dst.ctypes.data, len(src), int(directed),
                                  threads, C.byref(h)), "grass_graph_sort")
    return Csr(h)


def csr_from_sorted_edges(src, dst):
    src = np.ascontiguousarray(src, dtype=np.uint32)
    dst = np.ascontiguousarray(dst, dtype=np.uint32)
    h = C.c_void_p()
    _check(lib().grass_csr_from_sorted_edges(src.ctypes.data, dst.ctypes.data, len(src),
                                             C.byref(h)), "grass_csr_from_sorted_edges")
    return Csr(h)


def write_files(prefix, graph_bytes, offset_bytes):
    for ext, data in ((".graph", graph_bytes), (".offset", offset_bytes)):
        with open(os.fspath(prefix) + ext, "wb") as f:
            f.write(data)


def rmat_graph(scale, edge_factor=16, fmt="le", seed=42, threads=0, directed=False):
    """R-MAT -> graph_sort -> compressor: returns (graph_bytes, offset_bytes, csr)."""
    src, dst = rmat_edges(scale, edge_factor, seed=seed, threads=threads)
    csr = graph_sort(src, dst, directed=directed, threads=threads)
    del src, dst
    gb, ob = csr.compress(fmt, threads)
    return gb, ob, csr
