"""Polyhedron statistics (VGPU_F_GEOMETRY) and the text formats on the GPU,
against the reference library: every cell's vertices, vertex orders, edge
table, face areas, perimeters, normals, orders and vertex lists, surface
area, total edge distance, edge count and maximum radius bit for bit, for
the rectangular, radical, walled and periodic kinds (including cells large
enough for the 128/254-vertex wavefront tiers and the general per-lane
tiers); print_custom with every code and the particle and cell drawings
byte for byte."""
import numpy as np
import pytest

import voro_amd
from voro_amd import inputs as I

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def _case(name):
    if name == "uniform":
        return 0, (0, 20, 0, 20, 0, 20), (13, 13, 13), I.uniform_box(21, 12000, (0, 0, 0), (20, 20, 20)), None, None
    if name == "poly_wall":
        return (1, (0, 16, 0, 16, 0, 16), (10, 10, 10), I.uniform_box(22, 6000, (0, 0, 0), (16, 16, 16)),
                I.radii_lognormal(22, 6000), [(8, 8, 8, 7.5, -99)])
    if name == "periodic_flags":
        return 0, (0, 12, 0, 12, 0, 12), (8, 8, 8), I.uniform_box(23, 4000, (0, 0, 0), (12, 12, 12)), None, None
    if name == "large_radical":
        n = 5000
        r = np.full(n, 0.3)
        r[::90] = 2.2
        return 1, (0, 18, 0, 18, 0, 18), (9, 9, 9), I.uniform_box(24, n, (0, 0, 0), (18, 18, 18)), r, None
    if name == "periodic_triclinic":
        return 2, (12, 1.5, 12, 0.7, 1.1, 12), (8, 8, 8), I.uniform_box(25, 5000, (0, 0, 0), (12, 12, 12)), None, None
    if name == "periodic_poly":
        return (3, (12, 1, 12, 0.5, 0.8, 12), (8, 8, 8), I.uniform_box(26, 5000, (0, 0, 0), (12, 12, 12)),
                I.radii_lognormal(26, 5000), None)
    raise KeyError(name)


def _make(kind, geom, grid, xyz, r, walls, per=(0, 0, 0), ids=None):
    if kind >= 2:
        cls = voro_amd.container_periodic_poly if kind == 3 else voro_amd.container_periodic
        con = cls(*geom, *grid, 8)
    else:
        cls = voro_amd.container_poly if kind == 1 else voro_amd.container
        con = cls(*geom, *grid, *[bool(p) for p in per], 8)
    for w in walls or []:
        con.add_wall(voro_amd.wall_sphere(*w))
    con.put_array(np.arange(len(xyz), dtype=np.int32) if ids is None else ids, xyz, r)
    return con


CASES = ["uniform", "poly_wall", "periodic_flags", "large_radical", "periodic_triclinic", "periodic_poly"]


@pytest.mark.parametrize("name", CASES)
def test_gpu_geometry_matches_reference(oracle_mod, name):
    O = oracle_mod
    if O.reference() is None:
        pytest.skip("reference library not built")
    kind, geom, grid, xyz, r, walls = _case(name)
    per = (1, 0, 1) if name == "periodic_flags" else (0, 0, 0)
    _flush("geometry %s n=%d" % (name, len(xyz)))
    con = _make(kind, geom, grid, xyz, r, walls, per)
    con.compute_all_cells(voro_amd.F_NEIGHBORS | voro_amd.F_CSR | voro_amd.F_GEOMETRY)
    got = con.geometry()
    res = con.results()
    spec = O.make_spec(kind, geom, grid, periodic=per, spheres=walls)
    want = O.geometry_reference(spec, np.arange(len(xyz), dtype=np.int32), xyz, r)
    assert np.array_equal(res["nbr_off"], want["face_off"])
    for k in ("vert_off", "edge_off", "fv_off", "vertices", "orders", "edges", "edge_rel", "face_area", "face_perim",
              "normal", "face_order", "face_verts", "surface", "edge_dist", "n_edges", "maxr2"):
        assert np.array_equal(got[k], want[k]), k
    if name == "large_radical":
        st_max = int(np.max(np.diff(got["vert_off"])))
        assert st_max > 128     # cells finished by the wider tiers
    # the per-cell mirror: voronoicell_neighbor's methods at one loop position
    vl = voro_amd.c_loop_all(con)
    c = voro_amd.voronoicell_neighbor(con)
    assert vl.start()
    m = 0
    while not con.compute_cell(c, vl):     # the first cell a wall leaves
        assert vl.inc()
        m += 1
    f0, f1 = want["face_off"][m], want["face_off"][m + 1]
    assert c.face_areas() == want["face_area"][f0:f1].tolist()
    assert c.surface_area() == want["surface"][m]
    assert sum(c.face_orders()) == sum(c.vertex_orders())


@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_gpu_print_custom_every_code_matches_reference(oracle_mod, tmp_path, kind):
    O = oracle_mod
    if O.reference() is None:
        pytest.skip("reference library not built")
    n = 600
    xyz = I.uniform_box(31 + kind, n, (0, 0, 0), (8, 8, 8))
    r = I.radii_lognormal(31, n) * 0.6 if kind in (1, 3) else None
    if kind < 2:
        geom, grid = (0, 8, 0, 8, 0, 8), (5, 5, 5)
    else:
        geom, grid = (8, 0.8, 8, 0.3, 0.5, 8), (5, 5, 5)
    fmt = ("%i|%x %y %z|%q|%r|%w|%p|%P|%o|%m|%g|%E|%e|%s|%F|%A|%a|%f|%t|%l|%n|%v|%c|%C|100%% %u %")
    want = tmp_path / "ref.txt"
    assert O.print_custom_reference(O.make_spec(kind, geom, grid), np.arange(n, dtype=np.int32) + 7, xyz, r, fmt,
                                    str(want))
    con = _make(kind, geom, grid, xyz, r, None, ids=np.arange(n, dtype=np.int32) + 7)
    _flush("print_custom kind %d" % kind)
    assert con.print_custom(fmt) == want.read_text()


@pytest.mark.parametrize("kind", [0, 1, 3])
def test_gpu_drawings_match_reference(oracle_mod, tmp_path, kind):
    O = oracle_mod
    if O.reference() is None:
        pytest.skip("reference library not built")
    n = 300
    xyz = I.uniform_box(41 + kind, n, (0, 0, 0), (6, 6, 6))
    r = I.radii_lognormal(41, n) * 0.6 if kind in (1, 3) else None
    geom = (0, 6, 0, 6, 0, 6) if kind < 2 else (6, 0.4, 6, 0.2, 0.3, 6)
    grid = (4, 4, 4)
    ids = np.arange(n, dtype=np.int32)
    for what, meth in (("particles", "draw_particles"), ("particles_pov", "draw_particles_pov"),
                       ("cells_gnuplot", "draw_cells_gnuplot"), ("cells_pov", "draw_cells_pov")):
        want = tmp_path / ("%s.txt" % what)
        assert O.draw_reference(O.make_spec(kind, geom, grid), ids, xyz, r, what, str(want))
        con = _make(kind, geom, grid, xyz, r, None)
        _flush("draw %s kind %d" % (what, kind))
        assert getattr(con, meth)() == want.read_text(), what
