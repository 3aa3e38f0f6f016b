"""voro_amd -- MI355X (gfx950) engine for Voro++-style per-particle Voronoi cells.

Python mirror of the reference's container interface for the hot path
(src/container.hh:288-728, src/container_prd.hh:213-651,
src/c_loops.hh:165-209, 325-393, src/cell.hh:432-549),
over the C ABI of include/voro_gpu.h (libvoro_gpu.so, built in-tree by
voro_amd/build.py).  There is no CPU fallback: if the library is missing or
no GPU is visible, construction raises.

    con = container(0, 1, 0, 1, 0, 1, 6, 6, 6, False, False, False, 8)
    for i, (x, y, z) in enumerate(pts):
        con.put(i, x, y, z)            # or con.put_array(ids, xyz)
    con.compute_all_cells()
    vl = c_loop_all(con)
    c = voronoicell_neighbor(con)
    if vl.start():
        while True:
            if con.compute_cell(c, vl):
                c.volume(); c.centroid(); c.number_of_faces(); c.neighbors()
            if not vl.inc():
                break
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libvoro_gpu.so")

OK = 0
FILE_ERROR = 1
MEMORY_ERROR = 2
INTERNAL_ERROR = 3
INVALID_ARGUMENT = 5
HIP_ERROR = 6
DUPLICATE = 7
NO_DEVICE = 8

class TuneTrial(C.Structure):
    """vgpu_tune_trial: one configuration the autotuner timed."""
    _fields_ = [("first_tier", C.c_int), ("occupancy", C.c_int), ("waves", C.c_int),
                ("depth_offset", C.c_float), ("ms", C.c_double)]


F_NEIGHBORS = 1
F_STATS = 2
F_LEARNED = 4
F_CSR = 8
F_GRAPH = 16
F_GEOMETRY = 32

KIND_CONTAINER = 0
KIND_POLY = 1
KIND_PERIODIC = 2
KIND_PERIODIC_POLY = 3

_lib = None


class VoroGPUError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = _lib.vgpu_status_string(status).decode() if _lib is not None else str(status)
        super().__init__("%s: %s (status %d)" % (what, msg, status))


def lib():
    """The loaded engine library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("voro_amd: %s is missing; build it with voro_amd.build.build()" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    p = C.c_void_p
    i32p = C.POINTER(C.c_int)
    f64p = C.POINTER(C.c_double)
    i64p = C.POINTER(C.c_longlong)
    sig = {
        "vgpu_create": (C.c_int, [C.c_int, f64p, i32p, i32p, C.c_int, C.POINTER(p)]),
        "vgpu_destroy": (C.c_int, [p]),
        "vgpu_add_wall_sphere": (C.c_int, [p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
        "vgpu_add_wall": (C.c_int, [p, C.c_int, f64p, C.c_int]),
        "vgpu_point_inside": (C.c_int, [p, C.c_double, C.c_double, C.c_double, i32p]),
        "vgpu_loop_subset": (C.c_int, [p, C.c_int, f64p, i32p, C.c_int, i64p, i32p, C.c_longlong, i64p]),
        "vgpu_put": (C.c_int, [p, C.c_int, i32p, f64p, f64p]),
        "vgpu_put_device": (C.c_int, [p, C.c_int, p, p, p]),
        "vgpu_clear": (C.c_int, [p]),
        "vgpu_compute_all": (C.c_int, [p, C.c_uint, p]),
        "vgpu_num_cells": (C.c_int, [p, i64p]),
        "vgpu_find_voronoi_cells": (C.c_int, [p, C.c_int, f64p, i32p, i32p, f64p, C.c_int]),
        "vgpu_get_put_index": (C.c_int, [p, i32p]),
        "vgpu_compute_ghost_cells": (C.c_int, [p, C.c_int, f64p, f64p, i32p, f64p, f64p, i32p, i32p, C.c_int]),
        "vgpu_get_results": (C.c_int, [p, i32p, i32p, i32p, i32p, f64p, f64p, f64p, i32p, i64p, i32p,
                                       C.c_longlong, i64p]),
        "vgpu_get_stats": (C.c_int, [p, i32p, i32p, i32p]),
        "vgpu_sum_volumes": (C.c_int, [p, f64p]),
        "vgpu_set_launch": (C.c_int, [p, C.c_int, C.c_int]),
        "vgpu_set_first_tier": (C.c_int, [p, C.c_int, C.c_int, C.c_int]),
        "vgpu_last_timings": (C.c_int, [p, f64p]),
        "vgpu_status_string": (C.c_char_p, [C.c_int]),
        "vgpu_predictor_create": (C.c_int, [C.c_int, C.POINTER(p)]),
        "vgpu_predictor_destroy": (C.c_int, [p]),
        "vgpu_predictor_add_samples": (C.c_int, [p, p]),
        "vgpu_predictor_clear_samples": (C.c_int, [p]),
        "vgpu_predictor_num_samples": (C.c_int, [p, i64p]),
        "vgpu_predictor_train": (C.c_int, [p, C.c_int, C.c_float, C.c_uint, C.c_float, C.POINTER(C.c_float)]),
        "vgpu_predictor_params": (C.c_int, [p, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int]),
        "vgpu_predictor_margin": (C.c_int, [p, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "vgpu_predictor_route": (C.c_int, [p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "vgpu_set_predictor": (C.c_int, [p, p]),
        "vgpu_learned_stats": (C.c_int, [p, i64p, i64p, f64p, i64p]),
        "vgpu_batch_create": (C.c_int, [C.POINTER(p), C.c_int, C.POINTER(p)]),
        "vgpu_batch_destroy": (C.c_int, [p]),
        "vgpu_batch_compute": (C.c_int, [p, C.c_uint, p]),
        "vgpu_batch_set_predictor": (C.c_int, [p, p]),
        "vgpu_batch_learned_stats": (C.c_int, [p, i64p, i64p, i64p, i64p]),
        "vgpu_batch_list_capacity": (C.c_int, [p, i64p, i64p]),
        "vgpu_predictor_add_batch_samples": (C.c_int, [p, p]),
        "vgpu_batch_set_positions": (C.c_int, [p, p, C.c_int]),
        "vgpu_batch_num_cells": (C.c_int, [p, C.c_int, i64p]),
        "vgpu_batch_get_results": (C.c_int, [p, C.c_int, i32p, i32p, i32p, i32p, f64p, f64p, f64p, i32p, i64p, i32p,
                                             C.c_longlong, i64p]),
        "vgpu_batch_sum_volumes": (C.c_int, [p, f64p]),
        "vgpu_batch_config": (C.c_int, [p, i32p, i32p]),
        "vgpu_batch_autotune": (C.c_int, [p, C.c_uint, C.c_int, i32p, f64p]),
        "vgpu_autotune": (C.c_int, [p, C.c_uint, C.c_int, C.POINTER(TuneTrial), C.c_int, i32p]),
        "vgpu_batch_last_timings": (C.c_int, [p, f64p]),
        "vgpu_mlp_num_params": (C.c_int, []),
        "vgpu_mlp_loss_grad": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int,
                                         C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "vgpu_mlp_forward": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)]),
        "vgpu_geometry_sizes": (C.c_int, [p, i64p, i64p, i64p, i64p]),
        "vgpu_get_geometry": (C.c_int, [p, i64p, f64p, i32p, i64p, i32p, i32p, f64p, f64p, f64p, i32p, i64p, i32p,
                                        f64p, f64p, i32p, f64p]),
        "vgpu_kind": (C.c_int, [p, i32p]),
        "vgpu_bin": (C.c_int, [p]),
        "vgpu_get_particles": (C.c_int, [p, i64p, i32p, f64p, f64p]),
        "vgpu_print_custom": (C.c_int, [p, C.c_char_p, C.c_char_p]),
        "vgpu_import": (C.c_int, [p, C.c_char_p, C.c_int, i64p]),
        "vgpu_draw_particles": (C.c_int, [p, C.c_char_p, C.c_int]),
        "vgpu_draw_cells": (C.c_int, [p, C.c_char_p, C.c_int]),
        "vgpu_mlp_loss_grad_noalloc": (C.c_int, [C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float),
                                                 C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_float), p]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _check(st, what):
    if st != OK:
        raise VoroGPUError(st, what)


class _Base:
    """Common part of the container mirrors: a handle on the C ABI."""

    kind = KIND_CONTAINER

    def __init__(self, geom, grid, periodic, init_mem):
        L = lib()
        g = np.asarray(geom, dtype=np.float64)
        n = np.asarray(grid, dtype=np.int32)
        pr = np.asarray([1 if b else 0 for b in periodic], dtype=np.int32)
        h = C.c_void_p()
        _check(L.vgpu_create(self.kind, _p(g, C.c_double), _p(n, C.c_int), _p(pr, C.c_int), int(init_mem),
                             C.byref(h)), "vgpu_create")
        self._h = h
        self.geom = tuple(float(v) for v in g)
        self.nx, self.ny, self.nz = (int(v) for v in n)
        self.periodic = tuple(bool(v) for v in pr)
        self._res = None
        self._flags = F_NEIGHBORS
        self._walls = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.vgpu_destroy(h)
            self._h = None

    # -- particles ------------------------------------------------------
    def put_array(self, ids, xyz, r=None, order=None):
        """Appends particles in order, as repeated put() calls would; with a
        particle_order, as put(po, ...) calls (src/container.hh:294-296)."""
        ids = np.ascontiguousarray(ids, dtype=np.int32)
        xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
        _check(lib().vgpu_put(self._h, len(ids), _p(ids, C.c_int), _p(xyz, C.c_double), _p(rr, C.c_double)),
               "vgpu_put")
        first = getattr(self, "_nput", 0)
        self._nput = first + len(ids)
        if order is not None:
            order.add(self, first, len(ids))
        self._res = None

    def put_device(self, n, ids_ptr, xyz_ptr, r_ptr=None):
        """Replaces the particles by n particles already in device memory
        (raw device pointers, e.g. torch tensor data_ptr())."""
        _check(lib().vgpu_put_device(self._h, int(n), C.c_void_p(ids_ptr), C.c_void_p(xyz_ptr),
                                     C.c_void_p(r_ptr) if r_ptr else None), "vgpu_put_device")
        self._nput = int(n)
        self._res = None

    def clear(self):
        _check(lib().vgpu_clear(self._h), "vgpu_clear")
        self._nput = 0
        self._res = None

    def add_wall(self, wall):
        """add_wall (src/container.hh:63-69): a wall_sphere, wall_plane,
        wall_cylinder or wall_cone; walls cut each cell in the order added."""
        a = np.asarray(wall._args, dtype=np.float64)
        _check(lib().vgpu_add_wall(self._h, wall._type, _p(a, C.c_double), wall.w_id), "vgpu_add_wall")
        self._walls.append(wall)
        self._res = None

    def point_inside(self, x, y, z):
        """container_base::point_inside (src/container.cc:484-487): inside the
        box and every wall."""
        r = C.c_int()
        _check(lib().vgpu_point_inside(self._h, float(x), float(y), float(z), C.byref(r)), "vgpu_point_inside")
        return bool(r.value)

    # -- compute ---------------------------------------------------------
    def compute_all_cells(self, flags=F_NEIGHBORS, stream=None):
        """Bins the particles and computes every cell (src/container.cc:438-443),
        keeping the results for c_loop_all reads."""
        _check(lib().vgpu_compute_all(self._h, int(flags), stream), "vgpu_compute_all")
        self._flags = flags | (F_NEIGHBORS | F_CSR if flags & F_GEOMETRY else 0)
        self._res = None
        self._geo = None

    def sum_cell_volumes(self):
        if self._res is None and not self._computed():
            self.compute_all_cells()
        out = C.c_double()
        _check(lib().vgpu_sum_volumes(self._h, C.byref(out)), "vgpu_sum_volumes")
        return out.value

    def _computed(self):
        n = C.c_longlong()
        return lib().vgpu_num_cells(self._h, C.byref(n)) == OK

    def set_predictor(self, pred):
        """Uses a trained predictor for compute_all_cells(F_LEARNED | ...)."""
        _check(lib().vgpu_set_predictor(self._h, pred._h if pred is not None else None), "vgpu_set_predictor")
        self._pred = pred

    def learned_stats(self):
        """fallbacks: cells whose list ended first; routed: cells given no
        list (block search predicted); list_entries; gather_ms."""
        fb, ent, gm, ro = C.c_longlong(), C.c_longlong(), C.c_double(), C.c_longlong()
        _check(lib().vgpu_learned_stats(self._h, C.byref(fb), C.byref(ent), C.byref(gm), C.byref(ro)),
               "vgpu_learned_stats")
        return {"fallbacks": fb.value, "list_entries": ent.value, "gather_ms": gm.value, "routed": ro.value}

    def set_launch(self, fast_waves=0, fast_occupancy=0):
        _check(lib().vgpu_set_launch(self._h, int(fast_waves), int(fast_occupancy)), "vgpu_set_launch")

    def set_first_tier(self, tier=0, waves=0, occupancy=0):
        """0: one cell per wavefront (default); 1: one order-3 cell per lane.
        Results do not depend on the choice."""
        _check(lib().vgpu_set_first_tier(self._h, int(tier), int(waves), int(occupancy)), "vgpu_set_first_tier")

    def timings(self):
        a = np.zeros(4)
        _check(lib().vgpu_last_timings(self._h, _p(a, C.c_double)), "vgpu_last_timings")
        return {"bin_ms": a[0], "fast_ms": a[1], "retry_ms": a[2], "out_ms": a[3]}

    def autotune(self, flags=F_NEIGHBORS | F_CSR, reps=2, path=None, config=None):
        """The project's autotuner for this container (vgpu_autotune): times
        each launch configuration -- the first tier's cells per wavefront
        and register budget and, with F_LEARNED, the candidate-list length
        (depth margin) -- on the particles put, and keeps the fastest.
        Returns {"config", "best", "trials"}; with `path`, also writes that
        table as JSON (`config` names the workload in it)."""
        tr = (TuneTrial * 16)()
        n = C.c_int()
        _check(lib().vgpu_autotune(self._h, int(flags), int(reps), tr, 16, C.byref(n)), "vgpu_autotune")
        rows = [{"first_tier": "wavefront" if t.first_tier == 0 else "lane", "cells_per_wavefront":
                 1 if t.first_tier == 0 else 64, "occupancy": t.occupancy, "waves": t.waves,
                 "depth_offset": t.depth_offset, "ms": t.ms} for t in tr[:min(n.value, 16)]]
        out = {"config": config, "flags": int(flags), "reps": int(reps), "best": rows[0], "trials": rows}
        if path is not None:
            import json
            with open(path, "w") as f:
                json.dump(out, f, indent=1)
        return out

    def results(self):
        """All stored results as numpy arrays in c_loop_all order."""
        if self._res is not None:
            return self._res
        L = lib()
        n = C.c_longlong()
        _check(L.vgpu_num_cells(self._h, C.byref(n)), "vgpu_num_cells")
        n = n.value
        r = {k: np.zeros(n, np.int32) for k in ("id", "ijk", "q", "exists", "nfaces")}
        r["pos"] = np.zeros((n, 3))
        r["vol"] = np.zeros(n)
        r["cent"] = np.zeros((n, 3))
        r["nbr_off"] = np.zeros(n + 1, np.int64)
        tot = C.c_longlong()
        _check(L.vgpu_get_results(self._h, _p(r["id"], C.c_int), _p(r["ijk"], C.c_int), _p(r["q"], C.c_int),
                                  _p(r["exists"], C.c_int), _p(r["pos"], C.c_double), _p(r["vol"], C.c_double),
                                  _p(r["cent"], C.c_double), _p(r["nfaces"], C.c_int),
                                  _p(r["nbr_off"], C.c_longlong), None, 0, C.byref(tot)), "vgpu_get_results")
        nb = np.zeros(max(tot.value, 1), np.int32)
        if self._flags & F_NEIGHBORS:
            _check(L.vgpu_get_results(self._h, None, None, None, None, None, None, None, None, None,
                                      _p(nb, C.c_int), tot.value, C.byref(tot)), "vgpu_get_results")
        r["nbr"] = nb[:tot.value]
        self._res = r
        return r

    def stats(self):
        n = C.c_longlong()
        _check(lib().vgpu_num_cells(self._h, C.byref(n)), "vgpu_num_cells")
        a = np.zeros(n.value, np.int32)
        b = np.zeros(n.value, np.int32)
        t = np.zeros(n.value, np.int32)
        _check(lib().vgpu_get_stats(self._h, _p(a, C.c_int), _p(b, C.c_int), _p(t, C.c_int)), "vgpu_get_stats")
        return {"n_planes": a, "wl_exit": b, "tier": t}

    # -- polyhedron statistics (VGPU_F_GEOMETRY) --------------------------------
    def geometry(self):
        """Every cell's polyhedron statistics after a compute with F_GEOMETRY
        (computed here if the stored results lack them), in loop order, each
        the reference's voronoicell_neighbor function's value
        (src/cell.cc): per-vertex arrays (vertices, orders) by vert_off,
        per-edge-slot arrays (edges, edge_rel) by edge_off, per-face arrays
        (face_area, face_perim, normal, face_order) by the results' nbr_off,
        face_verts by fv_off; per-cell surface, edge_dist, n_edges, maxr2."""
        if getattr(self, "_geo", None) is not None:
            return self._geo
        L = lib()
        sz = [C.c_longlong() for _ in range(4)]
        if L.vgpu_geometry_sizes(self._h, *[C.byref(t) for t in sz]) != OK:
            self.compute_all_cells(F_NEIGHBORS | F_CSR | F_GEOMETRY)
            _check(L.vgpu_geometry_sizes(self._h, *[C.byref(t) for t in sz]), "vgpu_geometry_sizes")
        nv, ne, nf, nfv = (t.value for t in sz)
        n = len(self.results()["id"])
        g = {"vert_off": np.zeros(n + 1, np.int64), "vertices": np.zeros((max(nv, 1), 3)),
             "orders": np.zeros(max(nv, 1), np.int32), "edge_off": np.zeros(n + 1, np.int64),
             "edges": np.zeros(max(ne, 1), np.int32), "edge_rel": np.zeros(max(ne, 1), np.int32),
             "face_area": np.zeros(max(nf, 1)), "face_perim": np.zeros(max(nf, 1)), "normal": np.zeros((max(nf, 1), 3)),
             "face_order": np.zeros(max(nf, 1), np.int32), "fv_off": np.zeros(n + 1, np.int64),
             "face_verts": np.zeros(max(nfv, 1), np.int32), "surface": np.zeros(n), "edge_dist": np.zeros(n),
             "n_edges": np.zeros(n, np.int32), "maxr2": np.zeros(n)}
        f64, i32, i64 = C.c_double, C.c_int, C.c_longlong
        _check(L.vgpu_get_geometry(self._h, _p(g["vert_off"], i64), _p(g["vertices"], f64), _p(g["orders"], i32),
                                   _p(g["edge_off"], i64), _p(g["edges"], i32), _p(g["edge_rel"], i32),
                                   _p(g["face_area"], f64), _p(g["face_perim"], f64), _p(g["normal"], f64),
                                   _p(g["face_order"], i32), _p(g["fv_off"], i64), _p(g["face_verts"], i32),
                                   _p(g["surface"], f64), _p(g["edge_dist"], f64), _p(g["n_edges"], i32),
                                   _p(g["maxr2"], f64)), "vgpu_get_geometry")
        g["vertices"], g["orders"] = g["vertices"][:nv], g["orders"][:nv]
        g["edges"], g["edge_rel"] = g["edges"][:ne], g["edge_rel"][:ne]
        for k in ("face_area", "face_perim", "normal", "face_order"):
            g[k] = g[k][:nf]
        g["face_verts"] = g["face_verts"][:nfv]
        self._geo = g
        return g

    # -- text formats (src/container.cc:333-432, src/container.hh:325-450) -------
    def _text(self, fn, *args):
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "out.txt")
            _check(fn(self._h, path.encode(), *args), fn.__name__)
            with open(path) as f:
                return f.read()

    def print_custom(self, format, filename=None):
        """container::print_custom(format, filename) (src/container.cc:401-432):
        one line per computed cell in loop order, every code of
        voronoicell_base::output_custom (src/cell.cc:2485-2558), written by
        the engine library with the reference's printf conversions.  Returns
        the text, or writes it to filename."""
        L = lib()
        self._res = None
        self._geo = None
        if filename is not None:
            _check(L.vgpu_print_custom(self._h, format.encode(), str(filename).encode()), "vgpu_print_custom")
            self._flags = F_NEIGHBORS | F_CSR
            return None
        import tempfile
        with tempfile.TemporaryDirectory() as d:
            path = os.path.join(d, "custom.txt")
            _check(L.vgpu_print_custom(self._h, format.encode(), path.encode()), "vgpu_print_custom")
            self._flags = F_NEIGHBORS | F_CSR
            with open(path) as f:
                return f.read()

    def import_(self, filename):
        """container::import (src/container.cc:333-376): "id x y z" records
        ("id x y z r" for radical containers) until the end of the file, each
        put; a record that does not parse raises VoroGPUError(FILE_ERROR)
        after the records before it were put (the reference exits)."""
        poly = self.kind in (KIND_POLY, KIND_PERIODIC_POLY)
        n = C.c_longlong()
        st = lib().vgpu_import(self._h, str(filename).encode(), 1 if poly else 0, C.byref(n))
        self._nput = getattr(self, "_nput", 0) + n.value
        self._res = None
        _check(st, "import")

    def draw_particles(self, filename=None):
        """draw_particles (src/container.hh:329-347; radii for container_poly)."""
        return self._draw(lib().vgpu_draw_particles, 0, filename)

    def draw_particles_pov(self, filename=None):
        return self._draw(lib().vgpu_draw_particles, 1, filename)

    def draw_cells_gnuplot(self, filename=None):
        """draw_cells_gnuplot (src/container.hh:379-397)."""
        return self._draw(lib().vgpu_draw_cells, 0, filename)

    def draw_cells_pov(self, filename=None):
        """draw_cells_pov (src/container.hh:405-424)."""
        return self._draw(lib().vgpu_draw_cells, 1, filename)

    def _draw(self, fn, pov, filename):
        self._res = None
        self._geo = None
        if filename is not None:
            _check(fn(self._h, str(filename).encode(), pov), "draw")
            return None
        return self._text(fn, pov)

    # -- point location, as container::find_voronoi_cell ----------------------
    def find_voronoi_cells(self, xyz):
        """find_voronoi_cell (src/container.cc:237-297; the periodic classes'
        src/container_prd.cc:303-360) for every row of xyz: returns (found,
        pid, pos) arrays; found[i] is False where the point lies outside a
        non-periodic direction.  pos is the particle's position shifted by
        whole periods to the image nearest the point."""
        q = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        n = len(q)
        found = np.zeros(n, np.int32)
        pid = np.zeros(n, np.int32)
        pos = np.zeros((n, 3))
        _check(lib().vgpu_find_voronoi_cells(self._h, n, _p(q, C.c_double), _p(found, C.c_int), _p(pid, C.c_int),
                                             _p(pos, C.c_double), 0), "vgpu_find_voronoi_cells")
        return found.astype(bool), pid, pos

    def find_voronoi_cell(self, x, y, z):
        """The reference's signature in Python form: (found, (rx, ry, rz), pid)."""
        f, pid, pos = self.find_voronoi_cells(np.array([[x, y, z]], np.float64))
        return bool(f[0]), tuple(float(t) for t in pos[0]), int(pid[0])

    def compute_ghost_cells(self, xyz, r=None, max_nbr=128):
        """compute_ghost_cell (src/container.hh:488-498,710-721; the periodic
        classes' src/container_prd.hh:414-420,635-645) for every row of xyz
        (r: radii for radical containers): dict of exists, vol, cent
        (relative to the point), nfaces and nbr (max_nbr ids per point).
        exists[i] is -1 for a point this engine does not compute: a cell over
        the largest cell's capacity or max_nbr faces, a search that comes back
        to the point's own block through a periodic image, or (periodic
        classes) one that needs an image block the reference would create
        during the call."""
        q = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        m = len(q)
        rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
        o = {"exists": np.zeros(m, np.int32), "vol": np.zeros(m), "cent": np.zeros((m, 3)),
             "nfaces": np.zeros(m, np.int32), "nbr": np.zeros((m, max_nbr), np.int32)}
        st = lib().vgpu_compute_ghost_cells(self._h, m, _p(q, C.c_double), None if rr is None else _p(rr, C.c_double),
                                            _p(o["exists"], C.c_int), _p(o["vol"], C.c_double),
                                            _p(o["cent"], C.c_double), _p(o["nfaces"], C.c_int),
                                            _p(o["nbr"], C.c_int), int(max_nbr))
        if st != MEMORY_ERROR:      # per-point failures are marked exists = -1
            _check(st, "vgpu_compute_ghost_cells")
        return o

    # -- per-cell access, as voro_compute::compute_cell(c, vl) ----------------
    def compute_cell(self, c, vl):
        """Fills c with the stored cell at the loop position; returns the
        reference's compute_cell() value (False: the cell was removed)."""
        r = self.results()
        m = vl._m
        c._set(r, m, self._geo if getattr(self, "_geo", None) is not None else None)
        return bool(r["exists"][m])


class container(_Base):
    """container(ax,bx,ay,by,az,bz,nx,ny,nz,xp,yp,zp,init_mem), src/container.hh:290-291."""
    kind = KIND_CONTAINER

    def __init__(self, ax, bx, ay, by, az, bz, nx, ny, nz, xp, yp, zp, init_mem=8):
        super().__init__((ax, bx, ay, by, az, bz), (nx, ny, nz), (xp, yp, zp), init_mem)
        self._pend = []

    def put(self, n, x, y, z):
        self.put_array(np.array([n]), np.array([[x, y, z]]))


class container_poly(_Base):
    """container_poly(...), src/container.hh:510-515; put(n,x,y,z,r)."""
    kind = KIND_POLY

    def __init__(self, ax, bx, ay, by, az, bz, nx, ny, nz, xp, yp, zp, init_mem=8):
        super().__init__((ax, bx, ay, by, az, bz), (nx, ny, nz), (xp, yp, zp), init_mem)

    def put(self, n, x, y, z, r):
        self.put_array(np.array([n]), np.array([[x, y, z]]), np.array([r]))


class container_periodic(_Base):
    """container_periodic(bx,bxy,by,bxz,byz,bz,nx,ny,nz,init_mem),
    src/container_prd.hh:213-217: a triclinic lattice periodic in all three
    directions; put() remaps a particle into the primary domain
    (src/container_prd.cc:189-217) and rejects one within 1e-5 of another in
    its block, which surfaces as VoroGPUError(DUPLICATE) at compute time."""
    kind = KIND_PERIODIC

    def __init__(self, bx, bxy, by, bxz, byz, bz, nx, ny, nz, init_mem=8):
        super().__init__((bx, bxy, by, bxz, byz, bz), (nx, ny, nz), (True, True, True), init_mem)

    def put(self, n, x, y, z):
        self.put_array(np.array([n]), np.array([[x, y, z]]))


class container_periodic_poly(_Base):
    """container_periodic_poly(...), src/container_prd.hh:434-441; put(n,x,y,z,r)."""
    kind = KIND_PERIODIC_POLY

    def __init__(self, bx, bxy, by, bxz, byz, bz, nx, ny, nz, init_mem=8):
        super().__init__((bx, bxy, by, bxz, byz, bz), (nx, ny, nz), (True, True, True), init_mem)

    def put(self, n, x, y, z, r):
        self.put_array(np.array([n]), np.array([[x, y, z]]), np.array([r]))


class wall_sphere:
    """wall_sphere(xc,yc,zc,rc,w_id=-99), src/wall.hh:27-30."""
    _type = 0

    def __init__(self, xc, yc, zc, rc, w_id=-99):
        self.xc, self.yc, self.zc, self.rc, self.w_id = float(xc), float(yc), float(zc), float(rc), int(w_id)
        self._args = (self.xc, self.yc, self.zc, self.rc)


class wall_plane:
    """wall_plane(xc,yc,zc,ac,w_id=-99), src/wall.hh:44-47: the half-space
    x xc + y yc + z zc < ac."""
    _type = 1

    def __init__(self, xc, yc, zc, ac, w_id=-99):
        self._args = (float(xc), float(yc), float(zc), float(ac))
        self.w_id = int(w_id)


class wall_cylinder:
    """wall_cylinder(xc,yc,zc,xa,ya,za,rc,w_id=-99), src/wall.hh:65-68."""
    _type = 2

    def __init__(self, xc, yc, zc, xa, ya, za, rc, w_id=-99):
        self._args = tuple(float(t) for t in (xc, yc, zc, xa, ya, za, rc))
        self.w_id = int(w_id)


class wall_cone:
    """wall_cone(xc,yc,zc,xa,ya,za,ang,w_id=-99), src/wall.hh:97-101."""
    _type = 3

    def __init__(self, xc, yc, zc, xa, ya, za, ang, w_id=-99):
        self._args = tuple(float(t) for t in (xc, yc, zc, xa, ya, za, ang))
        self.w_id = int(w_id)


class particle_order:
    """particle_order (src/c_loops.hh:35-75): the particles put through
    container.put_array(..., order=po), in put order."""

    def __init__(self):
        self._con = None
        self._seq = []          # put-sequence numbers

    def add(self, con, first, count):
        if self._con is not None and self._con is not con:
            raise ValueError("a particle_order records the puts of one container")
        self._con = con
        self._seq.extend(range(first, first + count))


class c_loop_order:
    """c_loop_order (src/c_loops.hh:271-317): the stored particles a
    particle_order recorded, in record order (particles the container
    dropped are not stored, so not visited)."""

    def __init__(self, con, vo):
        self.con = con
        res = con.results()
        m = len(res["id"])
        put = np.zeros(max(m, 1), np.int32)
        _check(lib().vgpu_get_put_index(con._h, _p(put, C.c_int)), "vgpu_get_put_index")
        where = {int(put[k]): k for k in range(m)}
        self._idx = [where[s] for s in vo._seq if s in where]
        self._k = 0
        self._m = 0

    def start(self):
        self._k = 0
        if not self._idx:
            return False
        self._m = self._idx[0]
        return True

    def inc(self):
        self._k += 1
        if self._k >= len(self._idx):
            return False
        self._m = self._idx[self._k]
        return True

    def pid(self):
        return int(self.con.results()["id"][self._m])

    def pos(self):
        return tuple(self.con.results()["pos"][self._m])

    @property
    def ijk(self):
        return int(self.con.results()["ijk"][self._m])

    @property
    def q(self):
        return int(self.con.results()["q"][self._m])


class pre_container:
    """pre_container (src/pre_container.hh, src/pre_container.cc): keeps
    particles before the container exists, guesses its block grid and puts
    them into it.  radii=True for pre_container_poly."""
    optimal_particles = 5.6     # src/config.hh:97

    def __init__(self, ax, bx, ay, by, az, bz, xp, yp, zp, radii=False):
        self.ax, self.bx, self.ay, self.by, self.az, self.bz = ax, bx, ay, by, az, bz
        self.xp, self.yp, self.zp = bool(xp), bool(yp), bool(zp)
        self.radii = radii
        self.ids, self.xyz, self.r = [], [], []

    def put(self, n, x, y, z, r=None):
        # src/pre_container.cc:71-79 (out-of-bounds particles are not kept)
        if ((self.xp or (self.ax <= x <= self.bx)) and (self.yp or (self.ay <= y <= self.by)) and
                (self.zp or (self.az <= z <= self.bz))):
            self.ids.append(int(n)); self.xyz.append((x, y, z))
            if self.radii: self.r.append(r)

    def total_particles(self):
        return len(self.ids)

    def guess_optimal(self):
        """(nx, ny, nz), src/pre_container.cc:57-63."""
        dx, dy, dz = self.bx - self.ax, self.by - self.ay, self.bz - self.az
        ilscale = (self.total_particles() / (self.optimal_particles * dx * dy * dz)) ** (1 / 3.0)
        return int(dx * ilscale + 1), int(dy * ilscale + 1), int(dz * ilscale + 1)

    def setup(self, con, order=None):
        if self.ids:
            con.put_array(np.array(self.ids, np.int32), np.array(self.xyz, np.float64),
                          np.array(self.r, np.float64) if self.radii else None, order=order)


class c_loop_all:
    """Loop over all stored particles: blocks by ascending index, then
    storage order (src/c_loops.hh:165-209)."""

    def __init__(self, con):
        self.con = con
        self._m = 0

    def start(self):
        self._m = 0
        return len(self.con.results()["id"]) > 0

    def inc(self):
        self._m += 1
        return self._m < len(self.con.results()["id"])

    def pid(self):
        return int(self.con.results()["id"][self._m])

    def pos(self):
        return tuple(self.con.results()["pos"][self._m])

    @property
    def ijk(self):
        return int(self.con.results()["ijk"][self._m])

    @property
    def q(self):
        return int(self.con.results()["q"][self._m])


class c_loop_all_periodic(c_loop_all):
    """Loop over the particles of a periodic container: primary-domain
    blocks by ascending index, then storage order (src/c_loops.hh:325-393)."""


class c_loop_subset:
    """c_loop_subset (src/c_loops.hh:218-258, src/c_loops.cc:24-150) over a
    rectangular container: setup_sphere / setup_box / setup_intbox choose the
    blocks (and, with bounds_test, the particles) exactly as the reference's
    loop does; start() / inc() then step through them in its order.
    compute_cell(c, vl) gives the cell c_loop_all computes for the particle
    (the reference's own compute_cell over a subset loop that wraps a
    periodic direction passes the loop's unwrapped block index to frac_pos,
    src/container.hh:227-232, and is not reproduced)."""

    def __init__(self, con):
        self.con = con
        self._idx = np.zeros(0, np.int64)
        self._blk = np.zeros((0, 3), np.int32)
        self._t = 0

    def _setup(self, mode, v=None, iv=None, bounds_test=True):
        L = lib()
        vv = np.ascontiguousarray(v if v is not None else np.zeros(6), dtype=np.float64)
        ii = np.ascontiguousarray(iv if iv is not None else np.zeros(6), dtype=np.int32)
        n = C.c_longlong()
        _check(L.vgpu_loop_subset(self.con._h, mode, _p(vv, C.c_double), _p(ii, C.c_int), 1 if bounds_test else 0,
                                  None, None, 0, C.byref(n)), "vgpu_loop_subset")
        idx = np.zeros(max(n.value, 1), np.int64)
        blk = np.zeros((max(n.value, 1), 3), np.int32)
        _check(L.vgpu_loop_subset(self.con._h, mode, _p(vv, C.c_double), _p(ii, C.c_int), 1 if bounds_test else 0,
                                  _p(idx, C.c_longlong), _p(blk, C.c_int), n.value, C.byref(n)), "vgpu_loop_subset")
        self._idx, self._blk = idx[:n.value], blk[:n.value]
        self._t = 0

    def setup_sphere(self, vx, vy, vz, r, bounds_test=True):
        self._setup(0, v=[vx, vy, vz, r], bounds_test=bounds_test)

    def setup_box(self, xmin, xmax, ymin, ymax, zmin, zmax, bounds_test=True):
        self._setup(1, v=[xmin, xmax, ymin, ymax, zmin, zmax], bounds_test=bounds_test)

    def setup_intbox(self, ai, bi, aj, bj, ak, bk):
        self._setup(2, iv=[ai, bi, aj, bj, ak, bk])

    def start(self):
        self._t = 0
        return len(self._idx) > 0

    def inc(self):
        self._t += 1
        return self._t < len(self._idx)

    @property
    def _m(self):
        return int(self._idx[self._t])

    def _particles(self):
        if getattr(self, "_pp", None) is None:
            L = lib()
            n = C.c_longlong()
            _check(L.vgpu_get_particles(self.con._h, C.byref(n), None, None, None), "vgpu_get_particles")
            ids, pos, r = np.zeros(max(n.value, 1), np.int32), np.zeros((max(n.value, 1), 3)), np.zeros(max(n.value, 1))
            _check(L.vgpu_get_particles(self.con._h, C.byref(n), _p(ids, C.c_int), _p(pos, C.c_double),
                                        _p(r, C.c_double)), "vgpu_get_particles")
            self._pp = (ids, pos, r)
        return self._pp

    def pid(self):
        return int(self._particles()[0][self._m])

    def pos(self):
        return tuple(float(t) for t in self._particles()[1][self._m])

    @property
    def i(self):
        return int(self._blk[self._t, 0])

    @property
    def j(self):
        return int(self._blk[self._t, 1])

    @property
    def k(self):
        return int(self._blk[self._t, 2])


class voronoicell_neighbor:
    """Read-only view of one computed cell (src/cell.hh:432-549 outputs)."""

    def __init__(self, con=None):
        self._r = None
        self._m = 0
        self._g = None

    def _set(self, r, m, g=None):
        self._r = r
        self._m = m
        self._g = g

    # The polyhedron statistics (src/cell.cc) need a compute with
    # F_GEOMETRY (container.geometry() does one).
    def _geo(self):
        if self._g is None:
            raise RuntimeError("the container was computed without F_GEOMETRY: call its geometry() first")
        return self._g

    def _faces(self):
        s, e = self._r["nbr_off"][self._m], self._r["nbr_off"][self._m + 1]
        return s, e

    def _verts(self):
        g = self._geo()
        return g["vert_off"][self._m], g["vert_off"][self._m + 1]

    def vertices(self, x=None, y=None, z=None):
        """vertices() (local) or vertices(x,y,z) (global), flattened as the reference's vector."""
        s, e = self._verts()
        v = self._geo()["vertices"][s:e]
        if x is not None:
            v = v + np.array([x, y, z])
        return [float(t) for t in v.ravel()]

    def vertex_orders(self):
        s, e = self._verts()
        return [int(t) for t in self._geo()["orders"][s:e]]

    def face_areas(self):
        s, e = self._faces()
        return [float(t) for t in self._geo()["face_area"][s:e]]

    def face_perimeters(self):
        s, e = self._faces()
        return [float(t) for t in self._geo()["face_perim"][s:e]]

    def normals(self):
        s, e = self._faces()
        return [float(t) for t in self._geo()["normal"][s:e].ravel()]

    def face_orders(self):
        s, e = self._faces()
        return [int(t) for t in self._geo()["face_order"][s:e]]

    def face_freq_table(self):
        v = []
        for q in self.face_orders():
            if q >= len(v):
                v += [0] * (q + 1 - len(v))
            v[q] += 1
        return v

    def face_vertices(self):
        g = self._geo()
        return [int(t) for t in g["face_verts"][g["fv_off"][self._m]:g["fv_off"][self._m + 1]]]

    def surface_area(self):
        return float(self._geo()["surface"][self._m])

    def total_edge_distance(self):
        return float(self._geo()["edge_dist"][self._m])

    def number_of_edges(self):
        return int(self._geo()["n_edges"][self._m])

    def max_radius_squared(self):
        return float(self._geo()["maxr2"][self._m])

    @property
    def p(self):
        s, e = self._verts()
        return int(e - s)

    def volume(self):
        return float(self._r["vol"][self._m])

    def centroid(self):
        return tuple(float(v) for v in self._r["cent"][self._m])

    def number_of_faces(self):
        return int(self._r["nfaces"][self._m])

    def neighbors(self):
        s, e = self._r["nbr_off"][self._m], self._r["nbr_off"][self._m + 1]
        return [int(v) for v in self._r["nbr"][s:e]]


class container_batch:
    """Many independent rectangular containers computed together (no
    counterpart in the reference, which computes one container at a time):
    one binning pass and one persistent work-queue cell kernel over all of
    them.  Built from containers set up as usual (geometry, walls, put());
    results(i) gives container i's results exactly as its own
    compute_all_cells() would."""

    def __init__(self, containers):
        arr = (C.c_void_p * len(containers))(*[c._h for c in containers])
        h = C.c_void_p()
        _check(lib().vgpu_batch_create(arr, len(containers), C.byref(h)), "vgpu_batch_create")
        self._h = h
        self.count = len(containers)
        self._flags = F_NEIGHBORS

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.vgpu_batch_destroy(h)
            self._h = None

    def compute(self, flags=F_NEIGHBORS | F_CSR, graph=False, stream=None):
        """Computes every container; graph=True captures the device sequence
        as a hipGraph on the first call and replays it afterwards."""
        f = int(flags) | (F_GRAPH if graph else 0)
        _check(lib().vgpu_batch_compute(self._h, f, stream), "vgpu_batch_compute")
        self._flags = f

    def set_predictor(self, pred):
        """Uses a trained block predictor for compute(F_LEARNED | ...)."""
        _check(lib().vgpu_batch_set_predictor(self._h, pred._h if pred is not None else None),
               "vgpu_batch_set_predictor")
        self._pred = pred

    def learned_stats(self):
        """fallbacks: cells whose list ended first; routed: cells given no
        list; list_entries written; list_need: entries every list needed."""
        a = [C.c_longlong() for _ in range(4)]
        _check(lib().vgpu_batch_learned_stats(self._h, *[C.byref(x) for x in a]), "vgpu_batch_learned_stats")
        return dict(zip(("fallbacks", "routed", "list_entries", "list_need"), [x.value for x in a]))

    @property
    def list_capacity(self):
        v = C.c_longlong()
        _check(lib().vgpu_batch_list_capacity(self._h, C.byref(v), None), "vgpu_batch_list_capacity")
        return v.value

    @list_capacity.setter
    def list_capacity(self, entries):
        v = C.c_longlong(int(entries))
        _check(lib().vgpu_batch_list_capacity(self._h, None, C.byref(v)), "vgpu_batch_list_capacity")

    def set_positions(self, xyz):
        """New positions for every particle (same count and order)."""
        xyz = np.ascontiguousarray(xyz, dtype=np.float64)
        _check(lib().vgpu_batch_set_positions(self._h, xyz.ctypes.data_as(C.c_void_p), 0), "vgpu_batch_set_positions")

    def results(self, i):
        L = lib()
        n = C.c_longlong()
        _check(L.vgpu_batch_num_cells(self._h, int(i), C.byref(n)), "vgpu_batch_num_cells")
        n = n.value
        r = {k: np.zeros(n, np.int32) for k in ("id", "ijk", "q", "exists", "nfaces")}
        r["pos"] = np.zeros((n, 3))
        r["vol"] = np.zeros(n)
        r["cent"] = np.zeros((n, 3))
        r["nbr_off"] = np.zeros(n + 1, np.int64)
        tot = C.c_longlong()
        _check(L.vgpu_batch_get_results(self._h, int(i), _p(r["id"], C.c_int), _p(r["ijk"], C.c_int),
                                        _p(r["q"], C.c_int), _p(r["exists"], C.c_int), _p(r["pos"], C.c_double),
                                        _p(r["vol"], C.c_double), _p(r["cent"], C.c_double),
                                        _p(r["nfaces"], C.c_int), _p(r["nbr_off"], C.c_longlong), None, 0,
                                        C.byref(tot)), "vgpu_batch_get_results")
        nb = np.zeros(max(tot.value, 1), np.int32)
        if self._flags & F_NEIGHBORS:
            _check(L.vgpu_batch_get_results(self._h, int(i), None, None, None, None, None, None, None, None, None,
                                            _p(nb, C.c_int), tot.value, C.byref(tot)), "vgpu_batch_get_results")
        r["nbr"] = nb[:tot.value]
        return r

    def sum_cell_volumes(self):
        out = np.zeros(self.count)
        _check(lib().vgpu_batch_sum_volumes(self._h, _p(out, C.c_double)), "vgpu_batch_sum_volumes")
        return out

    @property
    def config(self):
        a = np.zeros(3, np.int32)
        _check(lib().vgpu_batch_config(self._h, _p(a, C.c_int), None), "vgpu_batch_config")
        return {"waves": int(a[0]), "occupancy": int(a[1]), "chunk": int(a[2])}

    def set_config(self, waves=0, occupancy=2, chunk=64):
        """occupancy 2..4: one cell per lane at that many waves per SIMD
        (the default first tier); 0, 5, 6: one cell per wavefront at 4, 5 or
        6 waves per SIMD.  waves: the first tier's persistent wavefronts
        (0 = as many as fit); chunk: cells per queue entry."""
        a = np.array([waves, occupancy, chunk], np.int32)
        _check(lib().vgpu_batch_config(self._h, None, _p(a, C.c_int)), "vgpu_batch_config")

    def autotune(self, flags=F_NEIGHBORS | F_CSR, reps=1):
        """The project's autotuner: times each launch configuration on this
        batch and keeps the fastest; returns it with its time (ms)."""
        a = np.zeros(3, np.int32)
        ms = C.c_double()
        _check(lib().vgpu_batch_autotune(self._h, int(flags), int(reps), _p(a, C.c_int), C.byref(ms)),
               "vgpu_batch_autotune")
        return {"waves": int(a[0]), "occupancy": int(a[1]), "chunk": int(a[2]), "ms": ms.value}

    def timings(self):
        a = np.zeros(3)
        _check(lib().vgpu_batch_last_timings(self._h, _p(a, C.c_double)), "vgpu_batch_last_timings")
        return {"setup_ms": a[0], "first_ms": a[1], "retry_ms": a[2]}


class predictor:
    """The learned candidate-depth predictor (no counterpart in the
    reference): a 28-32-32-2 perceptron trained on the GPU on statistics the
    plain path records (compute_all_cells with F_STATS)."""

    def __init__(self, with_radius=False):
        h = C.c_void_p()
        _check(lib().vgpu_predictor_create(1 if with_radius else 0, C.byref(h)), "vgpu_predictor_create")
        self._h = h
        self.with_radius = bool(with_radius)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.vgpu_predictor_destroy(h)
            self._h = None

    def add_samples(self, con):
        """Samples from a container's, or a batch's, last compute with F_STATS."""
        if isinstance(con, container_batch):
            _check(lib().vgpu_predictor_add_batch_samples(self._h, con._h), "vgpu_predictor_add_batch_samples")
        else:
            _check(lib().vgpu_predictor_add_samples(self._h, con._h), "vgpu_predictor_add_samples")

    def clear_samples(self):
        _check(lib().vgpu_predictor_clear_samples(self._h), "vgpu_predictor_clear_samples")

    def num_samples(self):
        n = C.c_longlong()
        _check(lib().vgpu_predictor_num_samples(self._h, C.byref(n)), "vgpu_predictor_num_samples")
        return n.value

    def train(self, epochs=400, lr=3e-3, seed=1, fallback=0.02):
        """Full-batch Adam on the GPU; returns the final training loss."""
        lo = C.c_float()
        _check(lib().vgpu_predictor_train(self._h, int(epochs), float(lr), int(seed), float(fallback), C.byref(lo)),
               "vgpu_predictor_train")
        return lo.value

    def params(self):
        n = lib().vgpu_mlp_num_params()
        a = np.zeros(n, np.float32)
        _check(lib().vgpu_predictor_params(self._h, _p(a, C.c_float), None, n), "vgpu_predictor_params")
        return a

    def set_params(self, a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        _check(lib().vgpu_predictor_params(self._h, None, _p(a, C.c_float), len(a)), "vgpu_predictor_params")

    @property
    def margin(self):
        m = C.c_float()
        _check(lib().vgpu_predictor_margin(self._h, C.byref(m), None), "vgpu_predictor_margin")
        return m.value

    @margin.setter
    def margin(self, v):
        m = C.c_float(v)
        _check(lib().vgpu_predictor_margin(self._h, None, C.byref(m)), "vgpu_predictor_margin")

    @property
    def route(self):
        """Predicted depths from this one on give a cell no list (the block
        search is predicted; the cell takes the plain path at once).  64 by
        default; 65 or more gives every cell a list."""
        m = C.c_int()
        _check(lib().vgpu_predictor_route(self._h, C.byref(m), None), "vgpu_predictor_route")
        return m.value

    @route.setter
    def route(self, v):
        m = C.c_int(int(v))
        _check(lib().vgpu_predictor_route(self._h, None, C.byref(m)), "vgpu_predictor_route")


def mlp_num_params():
    return lib().vgpu_mlp_num_params()


def mlp_loss_grad(params, X, T):
    """Loss and gradient of the predictor network on the GPU (its training
    kernels) for host arrays X (n x 28), T (n x 2)."""
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    g = np.zeros(len(params), np.float32)
    lo = C.c_float()
    _check(lib().vgpu_mlp_loss_grad(_p(params, C.c_float), _p(X, C.c_float), _p(T, C.c_float), len(X), C.byref(lo),
                                    _p(g, C.c_float)), "vgpu_mlp_loss_grad")
    return lo.value, g


def mlp_forward(params, X):
    params = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    Y = np.zeros((len(X), 2), np.float32)
    _check(lib().vgpu_mlp_forward(_p(params, C.c_float), _p(X, C.c_float), len(X), _p(Y, C.c_float)),
           "vgpu_mlp_forward")
    return Y
