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


def run(O, spec, ids, xyz, r=None, list_depth=0, wave=False, resume=True):
    """Same result dict as oracle.run, plus per-cell stats and tier counts.
    list_depth > 0 computes every cell through a candidate list of that
    worklist depth first (the learned path's kernel), the rest on the plain
    tiers; res["list_fallbacks"] counts the cells whose list was too short.
    wave: every cell is first tried by the wavefront-cell tiers;
    res["wave_done"] / res["wave2_done"] / res["wave4_done"] count the cells
    the 64-, 128- and 254-vertex ones finished.  resume: a wavefront cell
    whose list ends is continued from there (else computed again)."""
    lib = C.CDLL(SO)
    lib.hx_use_list(int(list_depth))
    lib.hx_list_resume(1 if resume else 0)
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


def geometry_stats(ref_geo, exists, tol):
    """geom_dev.h's cell_geometry_stats over the tables of a reference dump
    (oracle.geometry_reference): the statistics the engine's statistics
    pass computes from the tables it keeps."""
    lib = C.CDLL(SO)
    f = lib.hx_geometry_stats
    f.restype = C.c_longlong
    g = ref_geo
    m = len(g["surface"])
    nv, nf, nfv = len(g["orders"]), len(g["face_order"]), len(g["face_verts"])
    out = {"vertices": np.zeros((max(nv, 1), 3)), "face_area": np.zeros(max(nf, 1)), "face_perim": np.zeros(max(nf, 1)),
           "normal": np.zeros((max(nf, 1), 3)), "face_order": np.zeros(max(nf, 1), np.int32),
           "face_verts": np.zeros(max(nfv, 1), np.int32), "surface": np.zeros(m), "edge_dist": np.zeros(m),
           "n_edges": np.zeros(m, np.int32), "maxr2": np.zeros(m)}
    P = lambda a, t: a.ctypes.data_as(C.POINTER(t))
    ex = np.ascontiguousarray(exists, dtype=np.int32)
    k = f(m, P(ex, C.c_int), P(g["vert_off"], C.c_longlong), P(np.ascontiguousarray(g["pts"]), C.c_double),
          P(g["orders"], C.c_int), P(g["edge_off"], C.c_longlong), P(g["edges"], C.c_int), P(g["edge_rel"], C.c_int),
          P(g["face_off"], C.c_longlong), P(g["fv_off"], C.c_longlong), C.c_double(tol),
          P(out["vertices"], C.c_double), P(out["face_area"], C.c_double), P(out["face_perim"], C.c_double),
          P(out["normal"], C.c_double), P(out["face_order"], C.c_int), P(out["face_verts"], C.c_int),
          P(out["surface"], C.c_double), P(out["edge_dist"], C.c_double), P(out["n_edges"], C.c_int),
          P(out["maxr2"], C.c_double))
    if k < 0:
        raise RuntimeError("statistics changed the tables")
    out["vertices"], out["normal"] = out["vertices"][:nv], out["normal"][:nf]
    for key in ("face_area", "face_perim", "face_order"):
        out[key] = out[key][:nf]
    out["face_verts"] = out["face_verts"][:nfv]
    return out
