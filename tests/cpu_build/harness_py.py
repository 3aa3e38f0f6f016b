"""Builds and drives tests/cpu_build/harness.cpp: the engine's device code
compiled for the CPU (interleave 1).  Test infrastructure only."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "_build", "libharness.so")


def build():
    os.makedirs(os.path.join(HERE, "_build"), exist_ok=True)
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall",
                    "-o", SO, os.path.join(HERE, "harness.cpp")], check=True)


def run(O, spec, ids, xyz, r=None, list_depth=0, wave=False):
    """Same result dict as oracle.run, plus per-cell stats and tier counts.
    list_depth > 0 computes every cell through a candidate list of that
    worklist depth first (the learned path's kernel), the rest on the plain
    tiers; res["list_fallbacks"] counts the cells whose list was too short.
    wave: every cell is first tried by the wavefront-cell tiers;
    res["wave_done"] / res["wave2_done"] / res["wave4_done"] count the cells
    the 64-, 128- and 254-vertex ones finished."""
    lib = C.CDLL(SO)
    lib.hx_use_list(int(list_depth))
    lib.hx_use_wave(1 if wave else 0)
    lib.hx_list_fallbacks.restype = C.c_longlong
    lib.hx_wave_done.restype = C.c_longlong
    lib.hx_wave2_done.restype = C.c_longlong
    lib.hx_wave4_done.restype = C.c_longlong
    f = lib.hx_compute_neighbor
    f.restype = C.c_longlong
    n = len(ids)
    cap = max(1, n * 40)
    res = {k: np.zeros(n, np.int32) for k in ("id", "ijk", "q", "exists", "nfaces")}
    res["vol"] = np.zeros(n)
    res["cent"] = np.zeros((n, 3))
    res["pos"] = np.zeros((n, 3))
    res["nbr_off"] = np.zeros(n + 1, np.int64)
    res["nbr"] = np.zeros(cap, np.int32)
    o = O.Out()
    for k in ("id", "ijk", "q", "exists", "nfaces"):
        setattr(o, k, O._ptr(res[k], C.c_int))
    o.vol = O._ptr(res["vol"], C.c_double)
    o.cent = O._ptr(res["cent"], C.c_double)
    o.pos = O._ptr(res["pos"], C.c_double)
    o.nbr_off = O._ptr(res["nbr_off"], C.c_longlong)
    o.nbr = O._ptr(res["nbr"], C.c_int)
    o.nbr_cap = cap
    st = np.zeros(4 * max(n, 1), np.int32)
    tc = np.zeros(3, np.int64)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    m = f(C.byref(spec), n, O._ptr(ids, C.c_int), O._ptr(xyz, C.c_double),
          None if rr is None else O._ptr(rr, C.c_double), C.byref(o), O._ptr(st, C.c_int), O._ptr(tc, C.c_longlong))
    if m < 0:
        raise RuntimeError("harness failed (%d)" % m)
    out = {k: v[:m] for k, v in res.items() if k not in ("nbr_off", "nbr")}
    out["nbr_off"] = res["nbr_off"][:m + 1]
    out["nbr"] = res["nbr"][:int(res["nbr_off"][m])]
    out["stats"] = st[:4 * m].reshape(-1, 4)
    out["list_fallbacks"] = int(lib.hx_list_fallbacks())
    out["wave_done"] = int(lib.hx_wave_done())
    out["wave2_done"] = int(lib.hx_wave2_done())
    out["wave4_done"] = int(lib.hx_wave4_done())
    out["tiers"] = tc
    return out
