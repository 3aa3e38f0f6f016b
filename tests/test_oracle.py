"""The oracle, pinned: against the reference's own compiled code (golden
fixtures, and the live reference library where the checkout exists), against
closed forms, and against an independent half-space intersection."""
import hashlib
import os
import re

import numpy as np
import pytest

from voro_amd import inputs as I
from voro_amd import worklist_gen

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
FIXTURES = ["unit_cube_seed1.npz", "unit_cube_seed2.npz", "unit_cube_seed3.npz", "poly_small.npz",
            "wall_small.npz"]


def load(name):
    return dict(np.load(os.path.join(GOLD, name), allow_pickle=False))


def spec_of(O, d, name):
    sph = None
    if name.startswith("wall"):
        L = float(d["spec_geom"][1])
        sph = [[L / 2, L / 2, L / 2, 0.45 * L, -99]]
    return O.make_spec(int(d["spec_kind"]), d["spec_geom"], d["spec_grid"], spheres=sph)


def test_worklist_table_matches_reference_fixture():
    t = worklist_gen.generate()
    h = hashlib.sha256(np.array(t, dtype=np.uint32).tobytes()).hexdigest()
    assert h == open(os.path.join(GOLD, "worklist_sha256.txt")).read().strip()


def test_worklist_header_is_current():
    path = worklist_gen.HEADER
    assert open(path).read() == worklist_gen.header_text(worklist_gen.generate())


@pytest.mark.skipif(not os.path.exists("/root/reference/src/v_base_wl.cc"), reason="reference checkout absent")
def test_worklist_table_equals_reference_source():
    src = open("/root/reference/src/v_base_wl.cc").read()
    body = src[src.index("{") + 1:src.rindex("}")]
    ref = [int(x, 0) for x in re.findall(r"0x[0-9a-f]+|\b\d+\b", body)]
    assert ref == worklist_gen.generate()


def test_inputs_are_reproducible():
    d = load("unit_cube_seed1.npz")
    ids, xyz, _ = I.make("unit_cube", 1)
    assert np.array_equal(xyz, d["xyz"]) and np.array_equal(ids, d["ids"])


@pytest.mark.parametrize("name", FIXTURES)
def test_restatement_matches_golden(oracle_mod, name):
    O = oracle_mod
    d = load(name)
    out = O.run("restatement", spec_of(O, d, name), d["ids"], d["xyz"], d.get("r"))
    assert O.compare(out, d, vol_rtol=0.0) == []
    assert np.array_equal(out["vol"], d["vol"]) and np.array_equal(out["cent"], d["cent"])


def _ref_available(O):
    return O.reference() is not None


@pytest.mark.parametrize("name", ["prd_poly_small.npz", "prd_tri_small.npz"])
def test_periodic_golden_fixtures_match_reference_library(oracle_mod, name):
    """The periodic fixtures are the reference's own outputs (regenerated
    live where the reference library is built)."""
    O = oracle_mod
    if not _ref_available(O):
        pytest.skip("reference library not built here")
    d = load(name)
    out = O.run("reference", O.make_spec(int(d["spec_kind"]), d["spec_geom"], d["spec_grid"]), d["ids"], d["xyz"],
                d.get("r"))
    assert O.compare(out, d, vol_rtol=0.0) == []
    assert np.array_equal(out["vol"], d["vol"]) and np.array_equal(out["cent"], d["cent"])


@pytest.mark.parametrize("case", ["uniform", "clusters", "poly", "periodic_flags", "poly_periodic", "sc_lattice",
                                  "fcc_exact", "dropped"])
def test_restatement_matches_reference_library(oracle_mod, case):
    O = oracle_mod
    if not _ref_available(O):
        pytest.skip("reference library not built here")
    n = 6000
    ids = np.arange(n, dtype=np.int32)
    r = None
    per = (0, 0, 0)
    kind = 0
    geom, grid = (0, 20, 0, 20, 0, 20), (10, 10, 10)
    sph = None
    if case == "uniform":
        xyz = I.uniform_box(3, n, (0, 0, 0), (20, 20, 20))
    elif case == "clusters":
        xyz = I.clusters(4, n, box=20.0, n_clusters=8)
    elif case in ("poly", "poly_periodic"):
        xyz = I.uniform_box(5, n, (0, 0, 0), (20, 20, 20))
        r = I.radii_lognormal(5, n)
        kind = 1
        per = (1, 1, 1) if case == "poly_periodic" else (0, 0, 0)
    elif case == "periodic_flags":
        xyz = I.uniform_box(6, n, (-3, -3, -3), (23, 23, 23))
        per = (1, 0, 1)
    elif case == "sc_lattice":
        a = np.arange(8) + 0.5
        X, Y, Z = np.meshgrid(a, a, a, indexing="ij")
        xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
        ids = np.arange(len(xyz), dtype=np.int32)
        geom, grid = (0, 8, 0, 8, 0, 8), (4, 4, 4)
    elif case == "fcc_exact":
        xyz, L = I.fcc_sphere(1, 3000, 1.0, 0.0)
        ids = np.arange(len(xyz), dtype=np.int32)
        geom, grid = (0, L, 0, L, 0, L), (7, 7, 7)
        sph = [[L / 2, L / 2, L / 2, 0.45 * L, -99]]
    else:  # particles outside the box are dropped
        xyz = I.uniform_box(7, n, (-2, -2, -2), (22, 22, 22))
    sp = O.make_spec(kind, geom, grid, periodic=per, spheres=sph)
    a = O.run("reference", sp, ids, xyz, r)
    b = O.run("restatement", sp, ids, xyz, r)
    assert O.compare(b, a, vol_rtol=0.0) == []
    assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])


def test_single_particle_is_the_box(oracle_mod):
    O = oracle_mod
    out = O.run("restatement", O.make_spec(0, (0, 2, 0, 3, 0, 5), (1, 1, 1)), np.array([7], np.int32),
                np.array([[0.5, 1.0, 4.0]]))
    assert out["vol"][0] == 30.0 and out["nfaces"][0] == 6
    assert sorted(out["nbr"].tolist()) == [-6, -5, -4, -3, -2, -1]
    assert np.allclose(out["cent"][0], [1.0 - 0.5, 1.5 - 1.0, 2.5 - 4.0], rtol=0, atol=1e-14)


def test_simple_cubic_lattice_gives_unit_cubes(oracle_mod):
    O = oracle_mod
    a = np.arange(6) + 0.5
    X, Y, Z = np.meshgrid(a, a, a, indexing="ij")
    xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    out = O.run("restatement", O.make_spec(0, (0, 6, 0, 6, 0, 6), (3, 3, 3)), np.arange(216, dtype=np.int32), xyz)
    assert np.allclose(out["vol"], 1.0, rtol=1e-14, atol=0)
    assert np.all(out["nfaces"] == 6)
    assert np.allclose(out["cent"], 0.0, atol=1e-14)


def test_fcc_interior_cells_are_rhombic_dodecahedra(oracle_mod):
    O = oracle_mod
    m = 6
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) + 0.25
    g = np.stack(np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij"), -1).reshape(-1, 3)
    pts = (g[:, None, :] + base[None]).reshape(-1, 3)
    out = O.run("restatement", O.make_spec(0, (0, m, 0, m, 0, m), (4, 4, 4), periodic=(1, 1, 1)),
                np.arange(len(pts), dtype=np.int32), pts)
    # periodic in every direction: every cell is interior
    assert np.allclose(out["vol"], 0.25, rtol=1e-13, atol=0)
    assert np.all(out["nfaces"] == 12)
    assert abs(out["vol"].sum() - m ** 3) < 1e-11


def _halfspace_cells(xyz, box):
    """Independent check: each cell as the intersection of the box with the
    bisector half-spaces of every other particle (scipy / Qhull)."""
    from scipy.spatial import ConvexHull, HalfspaceIntersection
    n = len(xyz)
    vols, nsets = [], []
    (ax, bx, ay, by, az, bz) = box
    for i in range(n):
        p = xyz[i]
        hs = [[-1, 0, 0, ax], [1, 0, 0, -bx], [0, -1, 0, ay], [0, 1, 0, -by], [0, 0, -1, az], [0, 0, 1, -bz]]
        ids = [-1, -2, -3, -4, -5, -6]
        for j in range(n):
            if j == i:
                continue
            d = xyz[j] - p
            mid = 0.5 * (xyz[j] + p)
            hs.append([d[0], d[1], d[2], -float(d @ mid)])
            ids.append(j)
        hs = np.array(hs, dtype=float)
        h = HalfspaceIntersection(hs, p)
        hull = ConvexHull(h.intersections)
        vols.append(hull.volume)
        # a half-space is a face if the hull has a facet on its plane
        eq = hull.equations
        faces = set()
        for t, row in enumerate(hs):
            nrm = row[:3] / np.linalg.norm(row[:3])
            off = row[3] / np.linalg.norm(row[:3])
            if np.any((np.abs(eq[:, :3] @ nrm - 1) < 1e-9) & (np.abs(eq[:, 3] - off) < 1e-9)):
                faces.add(ids[t])
        nsets.append(faces)
    return np.array(vols), nsets


def test_restatement_matches_halfspace_intersection(oracle_mod):
    O = oracle_mod
    n = 120
    xyz = I.uniform_box(21, n, (0, 0, 0), (1, 1, 1))
    out = O.run("restatement", O.make_spec(0, (0, 1, 0, 1, 0, 1), (3, 3, 3)), np.arange(n, dtype=np.int32), xyz)
    vols, nsets = _halfspace_cells(xyz, (0, 1, 0, 1, 0, 1))
    order = out["id"]
    assert np.allclose(out["vol"], vols[order], rtol=1e-9, atol=0)
    for m, pid in enumerate(order):
        s, e = out["nbr_off"][m], out["nbr_off"][m + 1]
        assert set(out["nbr"][s:e].tolist()) == nsets[pid]
    assert abs(out["vol"].sum() - 1.0) < 1e-13


def test_volume_closure_and_neighbour_symmetry(oracle_mod):
    O = oracle_mod
    n = 5000
    xyz = I.uniform_box(8, n, (0, 0, 0), (10, 10, 10))
    out = O.run("restatement", O.make_spec(0, (0, 10, 0, 10, 0, 10), (10, 10, 10)), np.arange(n, dtype=np.int32),
                xyz)
    assert abs(out["vol"].sum() - 1000.0) / 1000.0 < 1e-12
    pairs = set()
    for m, pid in enumerate(out["id"]):
        for q in out["nbr"][out["nbr_off"][m]:out["nbr_off"][m + 1]]:
            if q >= 0:
                pairs.add((int(pid), int(q)))
    assert all((b, a) in pairs for (a, b) in pairs)


def test_planted_mistake_is_caught(oracle_mod):
    """The closure and golden checks fail on a perturbed result (checks that
    the comparison is not vacuous)."""
    O = oracle_mod
    d = load("unit_cube_seed1.npz")
    bad = dict(d)
    bad["vol"] = d["vol"].copy()
    bad["vol"][5] *= 1 + 1e-11
    assert O.compare(bad, d, vol_rtol=1e-12) != []
    bad = dict(d)
    bad["nbr"] = d["nbr"].copy()
    bad["nbr"][[3, 4]] = bad["nbr"][[4, 3]]
    assert O.compare(bad, d) != []


@pytest.mark.parametrize("case", ["uniform", "clusters_wall", "poly_periodic_flags"])
def test_threaded_reference_equals_single_thread(oracle_mod, case):
    """The threaded reference run (one container copy per thread, z-slabs of
    blocks through c_loop_subset::setup_intbox) gives the single-threaded
    c_loop_all outputs exactly, in the same order."""
    O = oracle_mod
    if not _ref_available(O):
        pytest.skip("reference library not built")
    r = None
    if case == "uniform":
        xyz = I.uniform_box(5, 30000, (0, 0, 0), (30, 30, 30))
        spec = O.make_spec(0, (0, 30, 0, 30, 0, 30), (18, 18, 18))
    elif case == "clusters_wall":
        xyz = I.clusters(3, 20000, box=30, n_clusters=8)
        spec = O.make_spec(0, (0, 30, 0, 30, 0, 30), (16, 16, 16), spheres=[(15, 15, 15, 13, -99)])
    else:
        xyz = I.uniform_box(6, 20000, (0, 0, 0), (25, 25, 25))
        r = I.radii_lognormal(6, len(xyz))
        spec = O.make_spec(1, (0, 25, 0, 25, 0, 25), (15, 15, 15), periodic=(1, 0, 1))
    ids = np.arange(len(xyz), dtype=np.int32)
    a = O.run("reference", spec, ids, xyz, r)
    for t in (3, 8):
        b = O.run("reference", spec, ids, xyz, r, threads=t)
        for k in a:
            assert np.array_equal(a[k], b[k]), (t, k)
