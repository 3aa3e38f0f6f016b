"""The engine's device code (voro_amd/csrc/compute_dev.h), compiled for the
CPU, against the oracle: the kernel logic is checked on hosts without a GPU,
including tier escalation on capacity overflow."""
import os
import sys

import numpy as np
import pytest

from voro_amd import inputs as I

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_build"))
import harness_py as H  # noqa: E402


@pytest.fixture(scope="module")
def harness():
    H.build()
    return H


CASES = {
    "unit_cube": lambda: (0, (0, 1, 0, 1, 0, 1), (6, 6, 6), None, I.make("unit_cube", 1)[1], None, (0, 0, 0)),
    "uniform": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), None, I.uniform_box(3, 6000, (0, 0, 0), (20, 20, 20)),
                        None, (0, 0, 0)),
    "clusters": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), None, I.clusters(4, 6000, box=20.0, n_clusters=8),
                         None, (0, 0, 0)),
    "poly_periodic": lambda: (1, (0, 20, 0, 20, 0, 20), (10, 10, 10), None,
                              I.uniform_box(5, 6000, (0, 0, 0), (20, 20, 20)), I.radii_lognormal(5, 6000), (1, 1, 1)),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_device_code_on_cpu_matches_oracle(oracle_mod, harness, case):
    O = oracle_mod
    kind, geom, grid, sph, xyz, r, per = CASES[case]()
    ids = np.arange(len(xyz), dtype=np.int32)
    sp = O.make_spec(kind, geom, grid, periodic=per, spheres=sph)
    a = O.run("restatement", sp, ids, xyz, r)
    b = harness.run(O, sp, ids, xyz, r)
    assert O.compare(b, a, vol_rtol=0.0) == []
    assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])


def test_tier_escalation_keeps_results(oracle_mod, harness):
    """Exact FCC creates vertices of order > 4, overflowing the fast tier's
    rows; those cells are recomputed by the wide tier with identical output."""
    O = oracle_mod
    xyz, L = I.fcc_sphere(1, 3000, 1.0, 0.0)
    ids = np.arange(len(xyz), dtype=np.int32)
    sp = O.make_spec(0, (0, L, 0, L, 0, L), (7, 7, 7), spheres=[[L / 2, L / 2, L / 2, 0.45 * L, -99]])
    a = O.run("restatement", sp, ids, xyz)
    b = harness.run(O, sp, ids, xyz)
    assert b["tiers"][1] > 0
    assert O.compare(b, a, vol_rtol=0.0) == []


def test_stats_match_restatement(oracle_mod, harness):
    O = oracle_mod
    xyz = I.uniform_box(9, 6000, (0, 0, 0), (20, 20, 20))
    ids = np.arange(6000, dtype=np.int32)
    sp = O.make_spec(0, (0, 20, 0, 20, 0, 20), (10, 10, 10))
    a = O.run("restatement", sp, ids, xyz, stats=True)
    b = harness.run(O, sp, ids, xyz)
    assert np.array_equal(b["stats"][:, 0], a["st_n_planes"])
    assert np.array_equal(b["stats"][:, 1], a["st_wl_exit"])


# ---- periodic containers: the engine's image-block replay, on the CPU ----

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("name", ["prd_poly_small.npz", "prd_tri_small.npz"])
def test_periodic_device_code_on_cpu_matches_golden(oracle_mod, harness, name):
    O = oracle_mod
    d = dict(np.load(os.path.join(GOLD, name), allow_pickle=False))
    sp = O.make_spec(int(d["spec_kind"]), d["spec_geom"], d["spec_grid"])
    b = harness.run(O, sp, d["ids"], d["xyz"], d.get("r"))
    assert O.compare(b, d, vol_rtol=0.0) == []
    assert np.array_equal(b["vol"], d["vol"]) and np.array_equal(b["cent"], d["cent"])


def _para(geom, n, seed, lo=-0.15, hi=1.15):
    bx, bxy, by, bxz, byz, bz = geom
    u = np.stack([I.uniform(seed, 80 + a, n) for a in range(3)], 1) * (hi - lo) + lo
    return np.stack([u[:, 0] * bx + u[:, 1] * bxy + u[:, 2] * bxz, u[:, 1] * by + u[:, 2] * byz, u[:, 2] * bz], 1)


PRD_CASES = {
    "cubic": (2, (10, 0, 10, 0, 0, 10), (4, 4, 4), 500, False),
    "triclinic_poly": (3, (10, 1.5, 10, 0.7, -1.2, 10), (5, 5, 5), 800, True),
    "skew": (3, (5, 7.5, 5, -6.2, 8.1, 5), (4, 4, 4), 600, True),
    "one_particle": (2, (1, 0.3, 1, 0.2, 0.1, 1), (1, 1, 1), 1, False),
    "fcc_cell": (2, (1, 0.5, np.sqrt(3) / 2, 0.5, np.sqrt(3) / 6, np.sqrt(2.0 / 3.0)), (2, 2, 2), 40, False),
}


@pytest.mark.parametrize("case", sorted(PRD_CASES))
def test_periodic_device_code_on_cpu_matches_reference(oracle_mod, harness, case):
    O = oracle_mod
    if O.reference() is None:
        pytest.skip("reference library not built")
    kind, geom, grid, n, poly = PRD_CASES[case]
    xyz = _para(geom, n, 100 + len(case))
    r = I.radii_lognormal(7, n) * 0.6 if poly else None
    ids = np.arange(n, dtype=np.int32)
    sp = O.make_spec(kind, geom, grid)
    a = O.run("reference", sp, ids, xyz, r)
    b = harness.run(O, sp, ids, xyz, r)
    assert O.compare(b, a, vol_rtol=0.0) == []
    assert np.array_equal(a["pos"], b["pos"])


def test_periodic_image_order_matters(oracle_mod, harness):
    """Image blocks built in create_all_images order (no replay of the
    first-visit order) give different neighbour orders than the reference's
    lazy creation; the replay restores them exactly."""
    import ctypes
    O = oracle_mod
    d = dict(np.load(os.path.join(GOLD, "prd_poly_small.npz"), allow_pickle=False))
    sp = O.make_spec(int(d["spec_kind"]), d["spec_geom"], d["spec_grid"])
    lib = ctypes.CDLL(harness.SO)
    lib.hx_image_passes.restype = ctypes.c_longlong
    lib.hx_set_replay(0)
    try:
        b = harness.run(O, sp, d["ids"], d["xyz"], d.get("r"))
    finally:
        lib.hx_set_replay(1)
    assert O.compare(b, d, vol_rtol=0.0) != []
    b = harness.run(O, sp, d["ids"], d["xyz"], d.get("r"))
    assert O.compare(b, d, vol_rtol=0.0) == []
    assert lib.hx_image_passes() >= 2


# ---- the wavefront-cell tier (wcell_dev.h), 64 lanes emulated on the CPU ----

@pytest.mark.parametrize("case", sorted(CASES))
def test_wave_cell_on_cpu_matches_oracle(oracle_mod, harness, case):
    """Every cell first goes through the wavefront cell (vertex v in lane v);
    the results, bit for bit, and the share it finishes."""
    O = oracle_mod
    kind, geom, grid, sph, xyz, r, per = CASES[case]()
    ids = np.arange(len(xyz), dtype=np.int32)
    sp = O.make_spec(kind, geom, grid, periodic=per, spheres=sph)
    a = O.run("restatement", sp, ids, xyz, r)
    b = harness.run(O, sp, ids, xyz, r, wave=True)
    assert O.compare(b, a, vol_rtol=0.0) == []
    assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])
    assert b["wave_done"] >= 0.99 * len(ids)


def test_wave_cell_walls_and_escalation(oracle_mod, harness):
    """Sphere walls, and exact FCC (vertices of order > 3): the wavefront cell
    hands every such cell on, with identical results."""
    O = oracle_mod
    xyz, L = I.fcc_sphere(1, 3000, 1.0, 0.0)
    ids = np.arange(len(xyz), dtype=np.int32)
    sp = O.make_spec(0, (0, L, 0, L, 0, L), (7, 7, 7), spheres=[[L / 2, L / 2, L / 2, 0.45 * L, -99]])
    a = O.run("restatement", sp, ids, xyz)
    b = harness.run(O, sp, ids, xyz, wave=True)
    assert b["tiers"][1] > 0
    assert O.compare(b, a, vol_rtol=0.0) == []
    xyz, L = I.fcc_sphere(2, 3000, 1.0, 0.08)
    ids = np.arange(len(xyz), dtype=np.int32)
    sp = O.make_spec(0, (0, L, 0, L, 0, L), (7, 7, 7), spheres=[[L / 2, L / 2, L / 2, 0.45 * L, -99]])
    a = O.run("restatement", sp, ids, xyz)
    b = harness.run(O, sp, ids, xyz, wave=True)
    assert b["wave_done"] > 0.9 * len(ids)
    assert O.compare(b, a, vol_rtol=0.0) == []
    assert np.array_equal(a["vol"], b["vol"])


@pytest.mark.parametrize("name", ["prd_poly_small.npz", "prd_tri_small.npz"])
def test_wave_cell_periodic_matches_golden(oracle_mod, harness, name):
    O = oracle_mod
    d = dict(np.load(os.path.join(GOLD, name), allow_pickle=False))
    sp = O.make_spec(int(d["spec_kind"]), d["spec_geom"], d["spec_grid"])
    b = harness.run(O, sp, d["ids"], d["xyz"], d.get("r"), wave=True)
    assert b["wave_done"] > 0
    assert O.compare(b, d, vol_rtol=0.0) == []
    assert np.array_equal(b["vol"], d["vol"]) and np.array_equal(b["cent"], d["cent"])


def test_wave_cell_128_vertices(oracle_mod, harness):
    """Radical cells of a few large particles among small ones grow beyond 64
    vertices: the 128- and 254-vertex wavefront cells (two and four vertices
    per lane) take them, all bit for bit."""
    O = oracle_mod
    n = 6000
    xyz = I.uniform_box(11, n, (0, 0, 0), (20, 20, 20))
    r = np.full(n, 0.3)
    r[::100] = 2.0
    ids = np.arange(n, dtype=np.int32)
    sp = O.make_spec(1, (0, 20, 0, 20, 0, 20), (10, 10, 10))
    a = O.run("restatement", sp, ids, xyz, r)
    b = harness.run(O, sp, ids, xyz, r, wave=True)
    assert b["wave2_done"] > 0 and b["wave4_done"] > 0
    assert O.compare(b, a, vol_rtol=0.0) == []
    assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])


def test_wave_cell_far_block_search(oracle_mod, harness):
    """Isolated points beside a dense cluster, with small blocks: block
    searches reach beyond the +-6-block window of the 64- and 128-vertex
    wavefront cells; the 254-vertex one, whose window reaches +-31 blocks,
    finishes them, so no cell goes on to the per-lane tiers (bit for bit)."""
    O = oracle_mod
    rng = np.random.default_rng(21)
    xyz = np.concatenate([rng.uniform(0, 8, (12000, 3)), rng.uniform(0, 40, (80, 3))])
    ids = np.arange(len(xyz), dtype=np.int32)
    sp = O.make_spec(0, (0, 40, 0, 40, 0, 40), (40, 40, 40))
    a = O.run("restatement", sp, ids, xyz)
    b = harness.run(O, sp, ids, xyz, wave=True)
    assert b["wave4_done"] >= 100
    assert b["wave_done"] + b["wave2_done"] + b["wave4_done"] == len(xyz)
    assert O.compare(b, a, vol_rtol=0.0) == []
    assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])


GEO_CASES = {
    "uniform": lambda O: (O.make_spec(0, (0, 10, 0, 10, 0, 10), (8, 8, 8)), I.uniform_box(3, 3000, (0, 0, 0), (10, 10, 10)), None),
    "poly_wall": lambda O: (O.make_spec(1, (0, 10, 0, 10, 0, 10), (7, 7, 7), spheres=[(5, 5, 5, 4.5, -99)]),
                            I.uniform_box(4, 2000, (0, 0, 0), (10, 10, 10)), I.radii_lognormal(4, 2000)),
    "periodic_triclinic": lambda O: (O.make_spec(2, (10, 1, 10, 0.5, 0.3, 10), (7, 7, 7)),
                                     I.uniform_box(5, 2000, (0, 0, 0), (10, 10, 10)), None),
    "large_radical": lambda O: (O.make_spec(1, (0, 12, 0, 12, 0, 12), (6, 6, 6)),
                                I.uniform_box(6, 1500, (0, 0, 0), (12, 12, 12)),
                                np.where(np.arange(1500) % 97 == 0, 2.2, 0.3)),
}


@pytest.mark.parametrize("case", sorted(GEO_CASES))
def test_geometry_statistics_restatement_matches_reference(oracle_mod, harness, case):
    """The statistics pass (geom_dev.h's cell_geometry_stats, compiled for
    the CPU) over the reference's own cell tables gives what the reference's
    face_areas, face_perimeters, normals, face_orders, face_vertices,
    surface_area, total_edge_distance, number_of_edges, max_radius_squared
    and vertices return, bit for bit, and leaves the tables unchanged."""
    O = oracle_mod
    if O.reference() is None:
        pytest.skip("reference library not built")
    spec, xyz, r = GEO_CASES[case](O)
    ids = np.arange(len(xyz), dtype=np.int32)
    g = O.geometry_reference(spec, ids, xyz, r)
    a = O.run("reference", spec, ids, xyz, r)
    assert np.array_equal(np.diff(g["face_off"]), a["nfaces"])
    h = H.geometry_stats(g, a["exists"], g["tol"])
    for k in ("vertices", "face_area", "face_perim", "normal", "face_order", "face_verts", "surface", "edge_dist",
              "n_edges", "maxr2"):
        assert np.array_equal(h[k], g[k]), k
    if case != "periodic_triclinic":
        # the tolerance is the reference's 10 eps max_len_sq of the box
        L = spec.geom[1] - spec.geom[0]
        assert g["tol"] == O.tolerance(3 * L * L)


def test_geometry_statistics_closed_forms(oracle_mod, harness):
    """One particle in a 2 x 3 x 5 box: six rectangular faces; areas, orders,
    perimeters, unit normals along the axes and the edge totals follow from
    the box (a planted mistake in the restatement, such as a dropped factor
    or a transposed cross product, fails here)."""
    O = oracle_mod
    if O.reference() is None:
        pytest.skip("reference library not built")
    spec = O.make_spec(0, (0, 2, 0, 3, 0, 5), (1, 1, 1))
    g = O.geometry_reference(spec, np.array([0], np.int32), np.array([[0.7, 1.1, 2.9]]))
    h = H.geometry_stats(g, np.array([1]), g["tol"])
    assert sorted(h["face_area"].tolist()) == sorted([6.0, 6.0, 10.0, 10.0, 15.0, 15.0])
    assert h["face_order"].tolist() == [4] * 6
    assert sorted(h["face_perim"].tolist()) == sorted([10.0, 10.0, 14.0, 14.0, 16.0, 16.0])
    assert h["surface"][0] == 62.0 and h["edge_dist"][0] == 40.0 and h["n_edges"][0] == 12
    nrm = h["normal"]
    assert np.allclose(np.abs(nrm).sum(axis=1), 1.0) and np.allclose(np.abs(nrm).max(axis=1), 1.0)
    assert np.allclose(nrm.sum(axis=0), 0.0)
    fv = h["face_verts"]
    assert len(fv) == 6 * 5 and all(fv[5 * t] == 4 for t in range(6))


WALL_CASES = {
    # examples/walls/tetrahedron.cc: four plane walls
    "tetrahedron": lambda: ((-2, 2, -2, 2, -2, 2), (7, 7, 7),
                            [(1, -99, 1, 1, 1, 1), (1, -99, -1, -1, 1, 1), (1, -99, 1, -1, -1, 1), (1, -99, -1, 1, -1, 1)],
                            I.uniform_box(71, 400, (-2, -2, -2), (2, 2, 2))),
    # examples/walls/frustum.cc: a cone
    "frustum": lambda: ((-1.2, 1.2, -1.2, 1.2, 0, 1), (14, 14, 7), [(3, -99, 0, 0, 2, 0, 0, -1, np.arctan(0.5))],
                        np.array([[x, y, z] for z in np.arange(0.1, 1, 0.2) for y in np.arange(-0.85, 1, 0.2)
                                  for x in np.arange(-0.95, 1, 0.2)])),
    # examples/walls/cylinder.cc: a cylinder along z, plus a sphere and a tilted plane
    "cylinder_sphere_plane": lambda: ((-6.5, 6.5, -6.5, 6.5, 0, 18.5), (7, 7, 14),
                                      [(2, 7, 0, 0, 0, 0, 0, 1, 6), (0, 8, 0, 0, 9, 8.0), (1, 9, 0.1, 0.2, 1.0, 15.0)],
                                      I.uniform_box(72, 2500, (-6.5, -6.5, 0), (6.5, 6.5, 18.5))),
    # a tilted cylinder and cone in one container
    "tilted": lambda: ((0, 10, 0, 10, 0, 10), (6, 6, 6),
                       [(2, -7, 5, 5, 5, 1, 0.5, 0.2, 4.2), (3, -8, 5, 5, -2, 0.1, 0.2, 1, 0.5)],
                       I.uniform_box(73, 2000, (0, 0, 0), (10, 10, 10))),
}


@pytest.mark.parametrize("case", sorted(WALL_CASES))
def test_wall_cuts_on_cpu_match_reference(oracle_mod, harness, case):
    """Plane, cylinder and cone walls (src/wall.cc:41-120) cut by the
    device code compiled for the CPU, per-lane and wavefront cells, equal
    the reference library bit for bit; point_inside selects the same
    particles."""
    O = oracle_mod
    if O.reference() is None:
        pytest.skip("reference library not built")
    geom, grid, walls, pts = WALL_CASES[case]()
    spec = O.make_spec(0, geom, grid, walls=walls)
    inside = O.point_inside_reference(spec, pts)
    xyz = pts[inside] if case in ("tetrahedron", "frustum") else pts
    ids = np.arange(len(xyz), dtype=np.int32)
    a = O.run("reference", spec, ids, xyz)
    assert a["exists"].sum() > 0
    for wave in (False, True):
        b = H.run(O, spec, ids, xyz, wave=wave)
        assert O.compare(b, a, vol_rtol=0.0) == [], wave
        assert np.array_equal(a["vol"], b["vol"])
