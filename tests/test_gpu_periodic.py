"""GPU parity of the periodic containers (container_periodic,
container_periodic_poly): the HIP engine against the golden fixtures and
the reference library itself (oracle/_ref), bit for bit on neighbour lists,
face counts and existence and to 1e-12 relative on volumes; at the full
2,000,000-particle size against size-independent properties (volume closure
to the lattice cell volume, neighbour symmetry)."""
import os

import numpy as np
import pytest

import voro_amd
from voro_amd import inputs as I

from gpu_util import engine_run

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
VOL_RTOL = 1e-12


def _flush(msg):
    print(msg, flush=True)


def para(geom, n, seed, lo=-0.15, hi=1.15):
    """n points in a slightly enlarged lattice cell (put() remaps them)."""
    bx, bxy, by, bxz, byz, bz = geom
    u = np.stack([I.uniform(seed, 80 + a, n) for a in range(3)], 1) * (hi - lo) + lo
    return np.stack([u[:, 0] * bx + u[:, 1] * bxy + u[:, 2] * bxz, u[:, 1] * by + u[:, 2] * byz, u[:, 2] * bz], 1)


def _reference(O):
    if O.reference() is None:
        pytest.skip("reference library not built")


@pytest.mark.parametrize("name", ["prd_poly_small.npz", "prd_tri_small.npz"])
def test_gpu_periodic_matches_golden(oracle_mod, name):
    O = oracle_mod
    d = dict(np.load(os.path.join(GOLD, name), allow_pickle=False))
    _flush("golden %s" % name)
    out = engine_run(int(d["spec_kind"]), d["spec_geom"], d["spec_grid"], d["ids"], d["xyz"], d.get("r"))
    assert O.compare(out, d, vol_rtol=VOL_RTOL) == []
    assert np.array_equal(out["vol"], d["vol"])


CASES = {
    "cubic": (2, (10, 0, 10, 0, 0, 10), (4, 4, 4), 500, False),
    "triclinic": (2, (10, 1.5, 10, 0.7, -1.2, 10), (4, 4, 4), 800, False),
    "triclinic_poly": (3, (10, 1.5, 10, 0.7, -1.2, 10), (5, 5, 5), 800, True),
    "skew": (3, (5, 7.5, 5, -6.2, 8.1, 5), (4, 4, 4), 600, True),
    "lattice_poly_20k": (3, (60, 5, 60, 3, 4, 60), (16, 16, 16), 20000, True),
    "lattice_one_block": (2, (1, 0.3, 1, 0.2, 0.1, 1), (1, 1, 1), 7, False),
    "fcc_cell": (2, (1, 0.5, np.sqrt(3) / 2, 0.5, np.sqrt(3) / 6, np.sqrt(2.0 / 3.0)), (2, 2, 2), 40, False),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_periodic_matches_reference(oracle_mod, case):
    O = oracle_mod
    _reference(O)
    kind, geom, grid, n, poly = CASES[case]
    xyz = para(geom, n, 100 + len(case))
    r = I.radii_lognormal(7, n) * 0.6 if poly else None
    ids = np.arange(n, dtype=np.int32)
    _flush("periodic case %s n=%d" % (case, n))
    a = O.run("reference", O.make_spec(kind, geom, grid), ids, xyz, r)
    b = engine_run(kind, geom, grid, ids, xyz, r)
    assert O.compare(b, a, vol_rtol=VOL_RTOL) == []
    assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["pos"], b["pos"])


def test_gpu_periodic_clusters_at_boundary(oracle_mod):
    """Dense clusters around a lattice corner: most cells read image blocks."""
    O = oracle_mod
    _reference(O)
    geom = (10, 1.5, 10, 0.7, -1.2, 10)
    c = np.concatenate([0.6 * np.stack([I.normal(3, 90 + 2 * a, 1500) for a in range(3)], 1),
                        0.6 * np.stack([I.normal(4, 90 + 2 * a, 1500) for a in range(3)], 1) + [9.8, 9.7, 0.1]])
    ids = np.arange(len(c), dtype=np.int32)
    _flush("periodic clusters")
    a = O.run("reference", O.make_spec(2, geom, (5, 5, 5)), ids, c)
    b = engine_run(2, geom, (5, 5, 5), ids, c)
    assert O.compare(b, a, vol_rtol=VOL_RTOL) == []


def test_gpu_periodic_duplicate_rejected():
    """put() of a particle within 1e-5 of another in its block ends the
    reference (src/common.cc:14-20); the engine reports DUPLICATE."""
    _flush("periodic duplicate")
    con = voro_amd.container_periodic(10, 0, 10, 0, 0, 10, 3, 3, 3)
    con.put_array(np.array([0, 1], np.int32), np.array([[1.0, 2.0, 3.0], [11.0, 2.0, 3.0 + 1e-7]]))
    with pytest.raises(voro_amd.VoroGPUError) as e:
        con.compute_all_cells(voro_amd.F_NEIGHBORS)
    assert e.value.status == voro_amd.DUPLICATE


def test_gpu_periodic_mid_size_matches_reference(oracle_mod):
    """250,000 particles of the triclinic radical workload's distribution in
    its lattice cell (37^3 blocks), against the reference library."""
    O = oracle_mod
    _reference(O)
    ids, xyz, r = I.make("periodic_poly", 7, 250000)
    _flush("periodic mid size 250k")
    sp = O.make_spec(3, (60, 5, 60, 3, 4, 60), (37, 37, 37))
    a = O.run("reference", sp, ids, xyz, r)
    b = engine_run(3, (60, 5, 60, 3, 4, 60), (37, 37, 37), ids, xyz, r)
    assert O.compare(b, a, vol_rtol=VOL_RTOL) == []
    assert np.array_equal(a["vol"], b["vol"])


def test_gpu_periodic_full_size_properties():
    """The 2,000,000-particle triclinic workload (container_periodic_poly
    (60,5,60,3,4,60), 74^3 blocks): the cells (radical cells: a small
    particle's cell may vanish) tile the lattice cell, and every neighbour
    relation is mutual (with multiplicity)."""
    ids, xyz, r = I.make("periodic_poly", 42)
    _flush("periodic full size 2M")
    b = engine_run(3, (60, 5, 60, 3, 4, 60), (74, 74, 74), ids, xyz, r)
    assert b["exists"].mean() > 0.5
    assert abs(b["vol"].sum() - 216000.0) / 216000.0 < 1e-10
    n = len(b["id"])
    owner = np.repeat(b["id"], np.diff(b["nbr_off"]))
    nb = b["nbr"]
    fwd = np.sort(owner.astype(np.int64) * n + nb)
    bwd = np.sort(nb.astype(np.int64) * n + owner)
    assert np.array_equal(fwd, bwd)
