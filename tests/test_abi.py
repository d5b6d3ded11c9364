"""CPU checks of the C ABI library: it loads without a GPU and exports every function the
headers under include/ declare; status strings; host-side argument errors."""
import ctypes as C
import glob
import os

import pytest

from conftest import REPO


def _declared():
    import re
    names = set()
    for h in glob.glob(os.path.join(REPO, "include", "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(grass_[a-z_0-9]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    import grass_amd
    L = grass_amd.lib()
    names = _declared()
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_strerror_and_null_handles():
    import grass_amd
    L = grass_amd.lib()
    assert grass_amd.strerror(0) == "success"
    assert "malformed" in grass_amd.strerror(grass_amd.EFORMAT)
    assert L.grass_info(None, None) == grass_amd.EINVAL
    assert L.grass_close(None) == grass_amd.OK
    assert L.grass_decode_csr(None, None, None, None) == grass_amd.EINVAL
    h = C.c_void_p()
    assert L.grass_open(None, 0, 0, C.byref(h)) == grass_amd.EINVAL
    assert L.grass_open(b"/nonexistent/x", 0, 0, C.byref(h)) == grass_amd.EIO
    assert L.grass_open(b"/nonexistent/x", 7, 0, C.byref(h)) in (grass_amd.EINVAL, grass_amd.EIO)


def test_export_argument_errors():
    import grass_amd
    L = grass_amd.lib()
    out = C.c_void_p()
    assert L.grass_mdrw_to_dlpack(None, 0, C.byref(out)) == grass_amd.EINVAL
    assert L.grass_khop_blocks_build(None, None, 0, None, None, None) == grass_amd.EINVAL
    assert L.grass_mdrw_local_edges(None, None, 0, None, None, None) == grass_amd.EINVAL
    assert L.grass_compress_device(None, None, 0, 0, 0, None, None, None) == grass_amd.EINVAL


def test_device_ingest_argument_errors():
    """Argument checks that come before any device work."""
    from grass_amd import ingest
    import grass_amd
    L = ingest.lib()
    assert L.grass_rmat_edges_device(0, 16, 0.57, 0.19, 0.19, 1, 0, None, None, None) == grass_amd.EINVAL
    assert L.grass_rmat_edges_device(10, 16, 0.6, 0.3, 0.3, 1, 0, None, None, None) == grass_amd.EINVAL
    n, m = C.c_uint64(), C.c_uint64()
    assert L.grass_graph_sort_device(None, None, 5, 0, 0, None, None, 0, None, 0, C.byref(n),
                                     C.byref(m)) == grass_amd.EINVAL


def test_device_open_argument_errors():
    import grass_amd
    L = grass_amd.lib()
    h = C.c_void_p()
    assert L.grass_open_device_images(None, 0, None, 0, 0, 0, C.byref(h)) == grass_amd.EINVAL
    assert L.grass_compress_device_open(None, None, 0, 0, 0, None, C.byref(h)) == grass_amd.EINVAL


def test_one_hip_runtime_when_the_library_loads_before_torch():
    """A process whose first grass_amd call is a host one (no torch imported yet) still ends up
    with a single HIP runtime: the library binds the copy torch loads."""
    import subprocess
    import sys
    code = ("from grass_amd import ingest\n"
            "ingest.rmat_edges(4, 2, seed=1, threads=1)\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read().split('\\n')\n"
            "print(sorted({l.split()[-1] for l in maps if 'libamdhip64' in l}))\n")
    p = subprocess.run([sys.executable, "-c", code], cwd=REPO, capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    paths = eval(p.stdout.strip().splitlines()[-1])
    assert len(paths) == 1, paths
