"""ctypes front end for the test oracles (test infrastructure only).

* ``restatement()`` -- oracle/_oracle/libvoro_oracle.so, the project's CPU
  restatement of the reference algorithm (oracle/voro_oracle.cc).
* ``reference()``   -- oracle/_ref/libvororef.so, the unmodified reference
  sources compiled by oracle/Makefile (absent where the reference checkout is
  not available and nothing was prebuilt).

Only tests, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use
this module, as the checker.  The product never imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_oracle", "libvoro_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "libvororef.so")
REF_ROOT = "/root/reference"

KIND = {"container": 0, "container_poly": 1, "container_periodic": 2, "container_periodic_poly": 3}


class Spec(C.Structure):
    _fields_ = [("kind", C.c_int), ("geom", C.c_double * 6), ("grid", C.c_int * 3),
                ("periodic", C.c_int * 3), ("init_mem", C.c_int), ("n_spheres", C.c_int),
                ("spheres", C.POINTER(C.c_double)), ("n_walls", C.c_int), ("walls", C.POINTER(C.c_double))]


class Out(C.Structure):
    _fields_ = [("id", C.POINTER(C.c_int)), ("ijk", C.POINTER(C.c_int)), ("q", C.POINTER(C.c_int)),
                ("exists", C.POINTER(C.c_int)), ("vol", C.POINTER(C.c_double)),
                ("cent", C.POINTER(C.c_double)), ("nfaces", C.POINTER(C.c_int)),
                ("nbr_off", C.POINTER(C.c_longlong)), ("nbr", C.POINTER(C.c_int)),
                ("nbr_cap", C.c_longlong), ("pos", C.POINTER(C.c_double))]


class Stats(C.Structure):
    _fields_ = [("n_planes", C.POINTER(C.c_int)), ("wl_exit", C.POINTER(C.c_int)),
                ("phase", C.POINTER(C.c_int)), ("max_p", C.POINTER(C.c_int)),
                ("max_order", C.POINTER(C.c_int))]


def build(reference=True, quiet=True):
    """Builds the restatement and, when the reference checkout is present,
    the reference library.  Never needed on hosts that got prebuilt files."""
    targets = ["oracle"]
    if reference and os.path.isdir(os.path.join(REF_ROOT, "src")):
        targets.append("ref")
    subprocess.run(["make", "-s", "-j8", "REF=" + REF_ROOT] + targets, cwd=HERE, check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


_LIBS = {}


def _load(path, fn):
    if path not in _LIBS:
        if not os.path.exists(path):
            return None
        lib = C.CDLL(path)
        f = getattr(lib, fn)
        f.restype = C.c_longlong
        _LIBS[path] = (lib, f)
    return _LIBS[path][1]


def restatement():
    f = _load(ORACLE_SO, "orc_compute_neighbor")
    if f is None:
        raise RuntimeError("oracle restatement not built: run oracle.build()")
    return f


def reference():
    """The reference-compiled oracle, or None where it is unavailable."""
    return _load(REF_SO, "ref_compute_neighbor")


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def make_spec(kind, geom, grid, periodic=(0, 0, 0), spheres=None, init_mem=8, walls=None):
    """walls: further walls after the spheres, each (type, w_id, *args) with
    type 0 sphere (xc,yc,zc,rc), 1 plane (xc,yc,zc,ac), 2 cylinder
    (xc,yc,zc,xa,ya,za,rc), 3 cone (xc,yc,zc,xa,ya,za,ang)."""
    s = Spec()
    s.kind = KIND[kind] if isinstance(kind, str) else int(kind)
    for t in range(6):
        s.geom[t] = float(geom[t])
    for t in range(3):
        s.grid[t] = int(grid[t])
        s.periodic[t] = int(periodic[t])
    s.init_mem = init_mem
    keep = None
    if spheres is not None and len(spheres):
        keep = np.ascontiguousarray(np.asarray(spheres, dtype=np.float64).reshape(-1, 5))
        s.n_spheres = keep.shape[0]
        s.spheres = _ptr(keep, C.c_double)
    else:
        s.n_spheres = 0
    wk = None
    if walls:
        wk = np.zeros((len(walls), 9))
        for t, w in enumerate(walls):
            wk[t, :len(w)] = w
        s.n_walls = len(walls)
        s.walls = _ptr(wk, C.c_double)
    else:
        s.n_walls = 0
    s._keep = (keep, wk)
    return s


def default_threads():
    """Host threads for the threaded reference runs (at most 16)."""
    try:
        return max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        return max(1, min(16, os.cpu_count() or 1))


def run(which, spec, ids, xyz, r=None, stats=False, nbr_per_cell=40, threads=None):
    """Runs one oracle; returns a dict of numpy arrays in c_loop_all order.
    which = "reference" with threads > 1 runs the reference library on that
    many threads (z-slabs of blocks, one container copy each; rectangular
    kinds), with outputs identical to the single-threaded loop."""
    f = restatement() if which == "restatement" else reference()
    if f is None:
        raise RuntimeError("reference oracle not available")
    mt = which != "restatement" and threads is not None and threads > 1 and int(spec.kind) <= 1
    if mt:
        lib = _LIBS[REF_SO][0]
        fmt = lib.ref_compute_neighbor_mt
        fmt.restype = C.c_longlong
    n = len(ids)
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    cap = max(1, n * nbr_per_cell)
    while True:
        res = {k: np.zeros(n, np.int32) for k in ("id", "ijk", "q", "exists", "nfaces")}
        res["vol"] = np.zeros(n)
        res["cent"] = np.zeros((n, 3))
        res["pos"] = np.zeros((n, 3))
        res["nbr_off"] = np.zeros(n + 1, np.int64)
        res["nbr"] = np.zeros(cap, np.int32)
        o = Out()
        for k in ("id", "ijk", "q", "exists", "nfaces"):
            setattr(o, k, _ptr(res[k], C.c_int))
        o.vol = _ptr(res["vol"], C.c_double)
        o.cent = _ptr(res["cent"], C.c_double)
        o.pos = _ptr(res["pos"], C.c_double)
        o.nbr_off = _ptr(res["nbr_off"], C.c_longlong)
        o.nbr = _ptr(res["nbr"], C.c_int)
        o.nbr_cap = cap
        st = None
        if stats and which == "restatement":
            sres = {k: np.zeros(n, np.int32) for k in ("n_planes", "wl_exit", "phase", "max_p", "max_order")}
            st = Stats()
            for k, v in sres.items():
                setattr(st, k, _ptr(v, C.c_int))
        args = [C.byref(spec), C.c_int(n), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
                None if rr is None else _ptr(rr, C.c_double), C.byref(o)]
        if which == "restatement":
            args.append(C.byref(st) if st is not None else None)
        elif mt:
            args += [C.c_int(threads), None]
        else:
            secs = C.c_double(0)
            args.append(C.byref(secs))
        m = (fmt if mt else f)(*args)
        if m < 0:
            raise RuntimeError("oracle reported an internal error (%d)" % m)
        if res["nbr_off"][m] > cap:
            cap = int(res["nbr_off"][m]) + 16
            continue
        break
    m = int(m)
    out = {k: v[:m] for k, v in res.items() if k not in ("nbr_off", "nbr")}
    out["nbr_off"] = res["nbr_off"][:m + 1]
    out["nbr"] = res["nbr"][:int(res["nbr_off"][m])]
    if st is not None:
        for k, v in sres.items():
            out["st_" + k] = v[:m]
    return out


def time_reference(spec, ids, xyz, r=None, loop="compute_all_cells"):
    """Seconds the reference takes for compute_all_cells() (or the
    voronoicell_neighbor output loop), steady_clock around the call only."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return None
    fn = lib.ref_time_compute_all_cells if loop == "compute_all_cells" else lib.ref_time_neighbor_loop
    fn.restype = C.c_double
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    return fn(C.byref(spec), C.c_int(len(ids)), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
              None if rr is None else _ptr(rr, C.c_double))


def time_reference_mt(spec, ids, xyz, r=None, threads=None, loop="compute_all_cells"):
    """Wall seconds of the reference on `threads` host threads (z-slabs of
    blocks, one container copy each, copies built before the clock starts):
    the compute_all_cells() body, or the voronoicell_neighbor output loop."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return None
    fn = lib.ref_time_mt
    fn.restype = C.c_double
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    t = threads or default_threads()
    return fn(C.byref(spec), C.c_int(len(ids)), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
              None if rr is None else _ptr(rr, C.c_double), C.c_int(t), C.c_int(1 if loop == "compute_all_cells" else 2))


class Geo(C.Structure):
    _fields_ = [("vert_off", C.POINTER(C.c_longlong)), ("vertices", C.POINTER(C.c_double)),
                ("orders", C.POINTER(C.c_int)), ("pts", C.POINTER(C.c_double)), ("vcap", C.c_longlong),
                ("edge_off", C.POINTER(C.c_longlong)), ("edges", C.POINTER(C.c_int)),
                ("edge_rel", C.POINTER(C.c_int)), ("ecap", C.c_longlong),
                ("face_off", C.POINTER(C.c_longlong)), ("face_area", C.POINTER(C.c_double)),
                ("face_perim", C.POINTER(C.c_double)), ("normal", C.POINTER(C.c_double)),
                ("face_order", C.POINTER(C.c_int)), ("fcap", C.c_longlong),
                ("fv_off", C.POINTER(C.c_longlong)), ("face_verts", C.POINTER(C.c_int)), ("fvcap", C.c_longlong),
                ("surface", C.POINTER(C.c_double)), ("edge_dist", C.POINTER(C.c_double)),
                ("n_edges", C.POINTER(C.c_int)), ("maxr2", C.POINTER(C.c_double)), ("tol", C.c_double)]


def geometry_reference(spec, ids, xyz, r=None):
    """Every cell's polyhedron statistics from the reference library's own
    voronoicell_neighbor functions (loop order), with the raw tables: a dict
    of numpy arrays (None where the library is unavailable)."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return None
    fn = lib.ref_cell_geometry
    fn.restype = C.c_longlong
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    n = len(ids)
    caps = [max(1, 40 * n), max(1, 120 * n), max(1, 20 * n), max(1, 140 * n)]
    while True:
        vc, ec, fc, fvc = caps
        o = {"vert_off": np.zeros(n + 1, np.int64), "vertices": np.zeros(3 * vc), "orders": np.zeros(vc, np.int32),
             "pts": np.zeros(3 * vc), "edge_off": np.zeros(n + 1, np.int64), "edges": np.zeros(ec, np.int32),
             "edge_rel": np.zeros(ec, np.int32), "face_off": np.zeros(n + 1, np.int64),
             "face_area": np.zeros(fc), "face_perim": np.zeros(fc), "normal": np.zeros(3 * fc),
             "face_order": np.zeros(fc, np.int32), "fv_off": np.zeros(n + 1, np.int64),
             "face_verts": np.zeros(fvc, np.int32), "surface": np.zeros(n), "edge_dist": np.zeros(n),
             "n_edges": np.zeros(n, np.int32), "maxr2": np.zeros(n)}
        g = Geo()
        for k, v in o.items():
            setattr(g, k, _ptr(v, C.c_longlong if v.dtype == np.int64 else
                             (C.c_int if v.dtype == np.int32 else C.c_double)))
        g.vcap, g.ecap, g.fcap, g.fvcap = vc, ec, fc, fvc
        m = fn(C.byref(spec), C.c_int(n), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
               None if rr is None else _ptr(rr, C.c_double), C.byref(g))
        need = [int(o["vert_off"][m]), int(o["edge_off"][m]), int(o["face_off"][m]), int(o["fv_off"][m])]
        if all(a <= b for a, b in zip(need, caps)):
            break
        caps = [max(a, b) + 16 for a, b in zip(need, caps)]
    m = int(m)
    for k in ("vert_off", "edge_off", "face_off", "fv_off"):
        o[k] = o[k][:m + 1]
    nv, ne, nf, nfv = need
    o["vertices"] = o["vertices"][:3 * nv].reshape(-1, 3)
    o["pts"] = o["pts"][:3 * nv].reshape(-1, 3)
    o["orders"] = o["orders"][:nv]
    o["edges"], o["edge_rel"] = o["edges"][:ne], o["edge_rel"][:ne]
    for k in ("face_area", "face_perim", "face_order"):
        o[k] = o[k][:nf]
    o["normal"] = o["normal"][:3 * nf].reshape(-1, 3)
    o["face_verts"] = o["face_verts"][:nfv]
    for k in ("surface", "edge_dist", "n_edges", "maxr2"):
        o[k] = o[k][:m]
    o["tol"] = g.tol
    return o


def tolerance(max_len_sq):
    """The cells' tolerance, tolerance * max_len_sq (src/config.hh:80,
    src/cell.cc:26)."""
    return 10.0 * 2.220446049250313e-16 * max_len_sq


def find_reference(spec, ids, xyz, r, q, timed=False):
    """find_voronoi_cell of the reference library for every row of q:
    (found, pid, pos[, seconds of the query loop]); None where the library is
    unavailable."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return None
    fn = lib.ref_find_voronoi_cells
    fn.restype = C.c_longlong
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    found = np.zeros(len(q), np.int32)
    pid = np.zeros(len(q), np.int32)
    pos = np.zeros((len(q), 3))
    secs = C.c_double(0)
    m = fn(C.byref(spec), C.c_int(len(ids)), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
           None if rr is None else _ptr(rr, C.c_double), C.c_int(len(q)), _ptr(q, C.c_double),
           _ptr(found, C.c_int), _ptr(pid, C.c_int), _ptr(pos, C.c_double), C.byref(secs))
    if m < 0:
        raise RuntimeError("reference find_voronoi_cell: rectangular containers only")
    out = (found.astype(bool), pid, pos)
    return out + (secs.value,) if timed else out


def ghost_reference(spec, ids, xyz, r, q, qr=None, max_nbr=128):
    """compute_ghost_cell of the reference library for every row of q:
    dict of exists, vol, cent, nfaces, nbr (max_nbr per point); None where
    the library is unavailable."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return None
    fn = lib.ref_compute_ghost_cells
    fn.restype = C.c_longlong
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    m = len(q)
    qq = None if qr is None else np.ascontiguousarray(qr, dtype=np.float64)
    o = {"exists": np.zeros(m, np.int32), "vol": np.zeros(m), "cent": np.zeros((m, 3)),
         "nfaces": np.zeros(m, np.int32), "nbr": np.zeros((m, max_nbr), np.int32)}
    k = fn(C.byref(spec), C.c_int(len(ids)), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
           None if rr is None else _ptr(rr, C.c_double), C.c_int(m), _ptr(q, C.c_double),
           None if qq is None else _ptr(qq, C.c_double), _ptr(o["exists"], C.c_int), _ptr(o["vol"], C.c_double),
           _ptr(o["cent"], C.c_double), _ptr(o["nfaces"], C.c_int), _ptr(o["nbr"], C.c_int), C.c_int(max_nbr))
    if k < 0:
        raise RuntimeError("reference compute_ghost_cell: rectangular containers only")
    return o


def prd_queries_reference(spec, ids, xyz, r, q, qr=None, pre=False, ghost=False, max_nbr=128):
    """The periodic classes' find_voronoi_cell (ghost False: (found, pid,
    pos)) or compute_ghost_cell (ghost True: dict as ghost_reference) of the
    reference library for every row of q, made one after another on one
    container; pre: compute_all_cells() first.  None where the library is
    unavailable."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return None
    fn = lib.ref_prd_queries
    fn.restype = C.c_longlong
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    m = len(q)
    qq = None if qr is None else np.ascontiguousarray(qr, dtype=np.float64)
    found = np.zeros(m, np.int32)
    pid = np.zeros(m, np.int32)
    pos = np.zeros((m, 3))
    o = {"vol": np.zeros(m), "cent": np.zeros((m, 3)), "nfaces": np.zeros(m, np.int32),
         "nbr": np.zeros((m, max_nbr), np.int32)}
    k = fn(C.byref(spec), C.c_int(len(ids)), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
           None if rr is None else _ptr(rr, C.c_double), C.c_int(1 if pre else 0), C.c_int(1 if ghost else 0),
           C.c_int(m), _ptr(q, C.c_double), None if qq is None else _ptr(qq, C.c_double), _ptr(found, C.c_int),
           _ptr(pid, C.c_int), _ptr(pos, C.c_double), _ptr(o["vol"], C.c_double), _ptr(o["cent"], C.c_double),
           _ptr(o["nfaces"], C.c_int), _ptr(o["nbr"], C.c_int), C.c_int(max_nbr))
    if k < 0:
        raise RuntimeError("reference periodic queries: periodic containers only")
    if ghost:
        o["exists"] = found
        return o
    return found.astype(bool), pid, pos


def print_custom_reference(spec, ids, xyz, r, fmt, filename):
    """The reference library's print_custom(format, filename) for this
    container; False where the library is unavailable."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return False
    fn = lib.ref_print_custom
    fn.restype = C.c_longlong
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    m = fn(C.byref(spec), C.c_int(len(ids)), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
           None if rr is None else _ptr(rr, C.c_double), fmt.encode(), filename.encode())
    return m == 0


def loop_subset_reference(spec, ids, xyz, r, mode, v=None, iv=None, bounds_test=True):
    """The reference's c_loop_subset steps: an (m, 6) array of particle id,
    i, j, k, ijk, q (None where the library is unavailable)."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return None
    fn = lib.ref_loop_subset
    fn.restype = C.c_longlong
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    vv = np.ascontiguousarray(v if v is not None else np.zeros(6), dtype=np.float64)
    ii = np.ascontiguousarray(iv if iv is not None else np.zeros(6), dtype=np.int32)
    cap = max(1, len(ids) * 8)
    while True:
        out = np.zeros((cap, 6), np.int32)
        m = fn(C.byref(spec), C.c_int(len(ids)), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
               None if rr is None else _ptr(rr, C.c_double), C.c_int(mode), _ptr(vv, C.c_double), _ptr(ii, C.c_int),
               C.c_int(1 if bounds_test else 0), _ptr(out, C.c_int), C.c_longlong(cap))
        if m <= cap:
            return out[:m]
        cap = int(m)


def point_inside_reference(spec, q):
    """container::point_inside of the reference for every row of q (the
    spec's box and walls); None where the library is unavailable."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return None
    fn = lib.ref_point_inside
    fn.restype = C.c_longlong
    q = np.ascontiguousarray(q, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(len(q), np.int32)
    fn(C.byref(spec), C.c_int(len(q)), _ptr(q, C.c_double), _ptr(out, C.c_int))
    return out.astype(bool)


def draw_reference(spec, ids, xyz, r, what, filename):
    """The reference library's drawing into filename: what is
    "particles", "particles_pov", "cells_gnuplot" or "cells_pov"; False
    where the library is unavailable."""
    lib = _load(REF_SO, "ref_compute_neighbor") and _LIBS[REF_SO][0]
    if not lib:
        return False
    code = {"particles": 1, "particles_pov": 2, "cells_gnuplot": 3, "cells_pov": 4}[what]
    fn = lib.ref_draw
    fn.restype = C.c_longlong
    ids = np.ascontiguousarray(ids, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz, dtype=np.float64)
    rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
    return fn(C.byref(spec), C.c_int(len(ids)), _ptr(ids, C.c_int), _ptr(xyz, C.c_double),
              None if rr is None else _ptr(rr, C.c_double), C.c_int(code), filename.encode()) == 0


def compare(a, b, vol_rtol=1e-12):
    """Returns a list of human-readable mismatches between two result dicts
    (empty when identical to the stated tolerance)."""
    errs = []
    for k in ("id", "ijk", "q", "exists", "nfaces"):
        if len(a[k]) != len(b[k]):
            return ["%s length %d vs %d" % (k, len(a[k]), len(b[k]))]
        bad = np.nonzero(a[k] != b[k])[0]
        if len(bad):
            errs.append("%s differs at %d cells (first loop index %d)" % (k, len(bad), bad[0]))
    if not np.array_equal(a["nbr_off"], b["nbr_off"]) or not np.array_equal(a["nbr"], b["nbr"]):
        na = np.diff(a["nbr_off"])
        nb = np.diff(b["nbr_off"])
        bad = np.nonzero(na != nb)[0]
        if len(bad) == 0:
            for i in range(len(na)):
                s, e = a["nbr_off"][i], a["nbr_off"][i + 1]
                if not np.array_equal(a["nbr"][s:e], b["nbr"][s:e]):
                    bad = [i]
                    break
        errs.append("neighbour lists differ (first loop index %s)" % (bad[0] if len(bad) else "?"))
    va, vb = a["vol"], b["vol"]
    rel = np.abs(va - vb) / np.maximum(np.abs(vb), 1e-300)
    rel[(va == 0) & (vb == 0)] = 0
    if np.any(rel > vol_rtol):
        errs.append("volume rel err %.3g at loop index %d" % (rel.max(), int(np.argmax(rel))))
    ca, cb = a["cent"], b["cent"]
    scale = np.maximum(np.abs(cb).max(axis=1), 1e-300)
    crel = np.abs(ca - cb).max(axis=1) / scale
    crel[np.all(ca == cb, axis=1)] = 0
    if np.any(crel > 1e-9):
        errs.append("centroid rel err %.3g at loop index %d" % (crel.max(), int(np.argmax(crel))))
    return errs
