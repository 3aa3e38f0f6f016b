"""The C-ABI library loads and exports every entry point its header declares
(no compute calls: this runs without a GPU)."""
import ctypes as C
import os
import re

import voro_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(vgpu_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(voro_amd.LIB_PATH), "run __graft_entry__.build() first"
    lib = C.CDLL(voro_amd.LIB_PATH)
    names = declared("voro_gpu.h")
    assert len(names) >= 15
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_status_strings_without_device():
    L = voro_amd.lib()
    assert L.vgpu_status_string(0) == b"ok"
    assert L.vgpu_status_string(2) == b"capacity exceeded"


def test_create_rejects_bad_geometry_without_touching_the_device():
    import numpy as np
    L = voro_amd.lib()
    h = C.c_void_p()
    g = np.array([1.0, 0.0, 0, 1, 0, 1])
    n = np.array([2, 2, 2], np.int32)
    st = L.vgpu_create(0, g.ctypes.data_as(C.POINTER(C.c_double)), n.ctypes.data_as(C.POINTER(C.c_int)), None, 8,
                       C.byref(h))
    assert st == voro_amd.INVALID_ARGUMENT
