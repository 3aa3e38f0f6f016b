"""Walls, the reference's own example packings, subset loops and edge-case
inputs on the GPU, against the reference library:

* plane, cylinder and cone walls (src/wall.cc:41-120), the set-ups of
  examples/walls/{cylinder,frustum,tetrahedron}.cc with point_inside;
* the packings examples/basic/pack_ten_cube, examples/custom/pack_six_cube
  and pack_six_cube_poly, examples/walls/pack_cylinder (copied under
  tests/golden/packs/), read through import and drawn as the examples do;
* c_loop_subset (setup_sphere / setup_box / setup_intbox, with and without
  periodic wrap) step for step;
* a simple-cubic lattice (degenerate vertices), NaN and infinite
  coordinates (dropped by put, as the reference's put_locate_block drops
  them), an empty container.

Neighbours, face counts and existence bit-exact, volumes within 1e-12
relative (bit-identical), text byte for byte."""
import os

import numpy as np
import pytest

import voro_amd
from voro_amd import inputs as I

pytestmark = pytest.mark.gpu

PACKS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "packs")


def _flush(msg):
    print(msg, flush=True)


def _ref(O):
    if O.reference() is None:
        pytest.skip("reference library not built")


def _wall(t):
    kind, wid, *a = t
    cls = {0: voro_amd.wall_sphere, 1: voro_amd.wall_plane, 2: voro_amd.wall_cylinder, 3: voro_amd.wall_cone}[kind]
    return cls(*a, w_id=wid)


def _container(kind, geom, grid, walls=(), per=(0, 0, 0)):
    cls = voro_amd.container_poly if kind == 1 else voro_amd.container
    con = cls(*geom, *grid, *[bool(p) for p in per], 8)
    for w in walls:
        con.add_wall(_wall(w))
    return con


def _read_pack(name, poly=False):
    a = np.loadtxt(os.path.join(PACKS, name))
    return a[:, 0].astype(np.int32), np.ascontiguousarray(a[:, 1:4]), (a[:, 4].copy() if poly else None)


CASES = {
    "cylinder_pack": lambda: ((-6.5, 6.5, -6.5, 6.5, 0, 18.5), (7, 7, 14), [(2, -99, 0, 0, 0, 0, 0, 1, 6)],
                              "pack_cylinder"),
    "tetrahedron": lambda: ((-2, 2, -2, 2, -2, 2), (7, 7, 7),
                            [(1, -99, 1, 1, 1, 1), (1, -99, -1, -1, 1, 1), (1, -99, 1, -1, -1, 1),
                             (1, -99, -1, 1, -1, 1)], I.uniform_box(81, 800, (-2, -2, -2), (2, 2, 2))),
    "frustum": lambda: ((-1.2, 1.2, -1.2, 1.2, 0, 1), (14, 14, 7), [(3, -99, 0, 0, 2, 0, 0, -1, np.arctan(0.5))],
                        np.array([[x, y, z] for z in np.arange(0.1, 1, 0.2) for y in np.arange(-0.85, 1, 0.2)
                                  for x in np.arange(-0.95, 1, 0.2)])),
    "mixed": lambda: ((0, 10, 0, 10, 0, 10), (6, 6, 6),
                      [(0, -5, 5, 5, 5, 4.8), (2, -7, 5, 5, 5, 1, 0.5, 0.2, 4.2), (3, -8, 5, 5, -2, 0.1, 0.2, 1, 0.5),
                       (1, -9, 0.3, 0.1, 1, 8.5)], I.uniform_box(82, 6000, (0, 0, 0), (10, 10, 10))),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_walls_match_reference(oracle_mod, case):
    O = oracle_mod
    _ref(O)
    geom, grid, walls, pts = CASES[case]()
    spec = O.make_spec(0, geom, grid, walls=walls)
    con = _container(0, geom, grid, walls)
    if isinstance(pts, str):
        ids, xyz, _ = _read_pack(pts)
        con.import_(os.path.join(PACKS, pts))
    else:
        inside = O.point_inside_reference(spec, pts)
        got_inside = np.array([con.point_inside(*p) for p in pts])
        assert np.array_equal(inside, got_inside)
        xyz = pts[inside] if case in ("tetrahedron", "frustum") else pts
        ids = np.arange(len(xyz), dtype=np.int32)
        con.put_array(ids, xyz)
    _flush("walls %s n=%d" % (case, len(ids)))
    con.compute_all_cells(voro_amd.F_NEIGHBORS | voro_amd.F_CSR)
    got = con.results()
    want = O.run("reference", spec, ids, xyz)
    assert O.compare(got, want, vol_rtol=1e-12) == []
    assert np.array_equal(got["vol"], want["vol"])
    if case == "frustum":
        # examples/walls/frustum.cc prints the cells' volume against the
        # frustum's: each cell is cut by the cone's tangent plane, so the sum
        # exceeds the frustum, by about 1% here
        evol = np.pi * 1 * (0.5 * 0.5 + 0.5 * 1 + 1 * 1) / 3
        assert evol < got["vol"].sum() < 1.02 * evol
    if case == "cylinder_pack":
        for what, meth in (("cells_pov", "draw_cells_pov"), ("particles_pov", "draw_particles_pov")):
            path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "vgpu_%s_%d.txt" % (what, os.getpid()))
            assert O.draw_reference(spec, ids, xyz, None, what, path)
            ref_text = open(path).read()
            os.unlink(path)
            assert getattr(con, meth)() == ref_text, what


@pytest.mark.parametrize("name,kind,geom,grid", [
    ("pack_ten_cube", 0, (-5, 5, -5, 5, 0, 10), (6, 6, 6)),      # examples/basic/import.cc
    ("pack_six_cube", 0, (-3, 3, -3, 3, 0, 6), (3, 3, 3)),       # examples/custom/radical.cc
    ("pack_six_cube_poly", 1, (-3, 3, -3, 3, 0, 6), (3, 3, 3)),
])
def test_gpu_reference_packings(oracle_mod, name, kind, geom, grid):
    O = oracle_mod
    _ref(O)
    ids, xyz, r = _read_pack(name, poly=kind == 1)
    con = _container(kind, geom, grid)
    con.import_(os.path.join(PACKS, name))
    _flush("packing %s n=%d" % (name, len(ids)))
    con.compute_all_cells(voro_amd.F_NEIGHBORS | voro_amd.F_CSR)
    got = con.results()
    spec = O.make_spec(kind, geom, grid)
    want = O.run("reference", spec, ids, xyz, r)
    assert O.compare(got, want, vol_rtol=1e-12) == []
    assert np.array_equal(got["vol"], want["vol"])
    path = os.path.join(os.environ.get("TMPDIR", "/tmp"), "vgpu_gnu_%d.txt" % os.getpid())
    assert O.draw_reference(spec, ids, xyz, r, "cells_gnuplot", path)
    ref_text = open(path).read()
    os.unlink(path)
    assert con.draw_cells_gnuplot() == ref_text


SUBSETS = [
    ("sphere", 0, (4.0, 5.5, 6.0, 2.5), None, True),
    ("sphere_blocks", 0, (4.0, 5.5, 6.0, 2.5), None, False),
    ("box", 1, (1.5, 7.2, 0.0, 3.3, 2.0, 9.9), None, True),
    ("box_blocks", 1, (1.5, 7.2, 0.0, 3.3, 2.0, 9.9), None, False),
    ("intbox", 2, None, (1, 3, 0, 5, 2, 2), True),
    ("intbox_clamped", 2, None, (-3, 2, 4, 9, -1, 1), True),
    ("sphere_outside", 0, (-4.0, 5.0, 5.0, 1.0), None, True),
]


@pytest.mark.parametrize("per", [(0, 0, 0), (1, 1, 0)])
@pytest.mark.parametrize("name,mode,v,iv,bt", SUBSETS)
def test_gpu_loop_subset_matches_reference(oracle_mod, name, mode, v, iv, bt, per):
    O = oracle_mod
    _ref(O)
    xyz = I.uniform_box(91, 3000, (0, 0, 0), (10, 10, 10))
    ids = np.arange(len(xyz), dtype=np.int32) + 5
    spec = O.make_spec(0, (0, 10, 0, 10, 0, 10), (6, 6, 6), periodic=per)
    want = O.loop_subset_reference(spec, ids, xyz, None, mode, v, iv, bt)
    con = _container(0, (0, 10, 0, 10, 0, 10), (6, 6, 6), per=per)
    con.put_array(ids, xyz)
    vl = voro_amd.c_loop_subset(con)
    if mode == 0:
        vl.setup_sphere(*v, bounds_test=bt)
    elif mode == 1:
        vl.setup_box(*v, bounds_test=bt)
    else:
        vl.setup_intbox(*iv)
    got = []
    if vl.start():
        while True:
            got.append((vl.pid(), vl.i, vl.j, vl.k))
            if not vl.inc():
                break
    assert got == [tuple(int(t) for t in w[:4]) for w in want]
    if not any(per):
        # the cells at the visited positions are c_loop_all's, which are the reference's
        con.compute_all_cells()
        res = con.results()
        c = voro_amd.voronoicell_neighbor(con)
        full = O.run("reference", spec, ids, xyz)
        if vl.start():
            while True:
                ok = con.compute_cell(c, vl)
                m = vl._m
                assert ok == bool(full["exists"][m]) and c.volume() == full["vol"][m]
                if not vl.inc():
                    break


def test_gpu_simple_cubic_lattice_matches_reference(oracle_mod):
    """A simple-cubic lattice: every vertex shared by eight cells, every plane
    through existing vertices (the marginal cases)."""
    O = oracle_mod
    _ref(O)
    g = np.arange(12) + 0.5
    xyz = np.array([[x, y, z] for z in g for y in g for x in g])
    ids = np.arange(len(xyz), dtype=np.int32)
    con = _container(0, (0, 12, 0, 12, 0, 12), (5, 5, 5))
    con.put_array(ids, xyz)
    con.compute_all_cells()
    got = con.results()
    want = O.run("reference", O.make_spec(0, (0, 12, 0, 12, 0, 12), (5, 5, 5)), ids, xyz)
    assert O.compare(got, want, vol_rtol=1e-12) == []
    assert np.allclose(got["vol"], 1.0, rtol=0, atol=1e-13) and np.all(got["nfaces"] == 6)


def test_gpu_non_finite_puts_are_dropped(oracle_mod):
    """NaN and infinite coordinates: put drops the particle (the reference's
    step_int of a non-finite value lands outside the grid on x86-64)."""
    O = oracle_mod
    _ref(O)
    xyz = I.uniform_box(92, 500, (0, 0, 0), (5, 5, 5))
    bad = np.array([[np.nan, 1, 1], [1, np.inf, 1], [1, 1, -np.inf], [np.inf, np.nan, 2], [2.5, 2.5, 1e300]])
    allp = np.concatenate([xyz[:200], bad, xyz[200:]])
    ids = np.arange(len(allp), dtype=np.int32)
    con = _container(0, (0, 5, 0, 5, 0, 5), (4, 4, 4))
    con.put_array(ids, allp)
    con.compute_all_cells()
    got = con.results()
    want = O.run("reference", O.make_spec(0, (0, 5, 0, 5, 0, 5), (4, 4, 4)), ids, allp)
    assert len(got["id"]) == 500
    assert O.compare(got, want, vol_rtol=1e-12) == []


def test_gpu_empty_container():
    con = _container(0, (0, 1, 0, 1, 0, 1), (100, 100, 100))
    con.compute_all_cells()
    assert len(con.results()["id"]) == 0
    assert con.sum_cell_volumes() == 0.0
    vl = voro_amd.c_loop_all(con)
    assert not vl.start()
