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
