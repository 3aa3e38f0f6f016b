"""GPU parity: the HIP engine against the golden fixtures (reference
outputs) and the oracle restatement, bit for bit on neighbour lists, face
counts and existence, and to 1e-12 relative on volumes (they come out
bit-identical; the tolerance is the contract)."""
import os

import numpy as np
import pytest

from voro_amd import inputs as I

from gpu_util import engine_run

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
VOL_RTOL = 1e-12


def _flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("name", ["unit_cube_seed1.npz", "unit_cube_seed2.npz", "unit_cube_seed3.npz",
                                  "poly_small.npz", "wall_small.npz"])
def test_gpu_matches_golden(oracle_mod, name):
    O = oracle_mod
    d = dict(np.load(os.path.join(GOLD, name), allow_pickle=False))
    _flush("golden %s" % name)
    geom = d["spec_geom"]
    sph = [[geom[1] / 2, geom[1] / 2, geom[1] / 2, 0.45 * geom[1], -99]] if name.startswith("wall") else None
    out = engine_run(int(d["spec_kind"]), geom, d["spec_grid"], d["ids"], d["xyz"], d.get("r"), spheres=sph)
    assert O.compare(out, d, vol_rtol=VOL_RTOL) == []
    assert np.array_equal(out["vol"], d["vol"])


CASES = {
    "uniform": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), I.uniform_box(3, 6000, (0, 0, 0), (20, 20, 20)),
                        None, (0, 0, 0), None),
    "clusters": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), I.clusters(4, 6000, box=20.0, n_clusters=8), None,
                         (0, 0, 0), None),
    "poly": lambda: (1, (0, 20, 0, 20, 0, 20), (10, 10, 10), I.uniform_box(5, 6000, (0, 0, 0), (20, 20, 20)),
                     I.radii_lognormal(5, 6000), (0, 0, 0), None),
    "poly_periodic": lambda: (1, (0, 20, 0, 20, 0, 20), (10, 10, 10), I.uniform_box(5, 6000, (0, 0, 0), (20, 20, 20)),
                              I.radii_lognormal(5, 6000), (1, 1, 1), None),
    "periodic_flags": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10),
                               I.uniform_box(6, 6000, (-3, -3, -3), (23, 23, 23)), None, (1, 0, 1), None),
    "dropped": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), I.uniform_box(7, 6000, (-2, -2, -2), (22, 22, 22)),
                        None, (0, 0, 0), None),
    # ~1,000 particles per block: the wavefront cells check a block's
    # particles 64 at a time, so these take many rounds per block
    "dense_blocks": lambda: (0, (0, 10, 0, 10, 0, 10), (2, 2, 2), I.uniform_box(8, 8000, (0, 0, 0), (10, 10, 10)),
                             None, (0, 0, 0), None),
    "dense_poly": lambda: (1, (0, 10, 0, 10, 0, 10), (3, 3, 3), I.uniform_box(9, 8000, (0, 0, 0), (10, 10, 10)),
                           I.radii_lognormal(9, 8000) * 0.5, (0, 0, 0), None),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_matches_oracle(oracle_mod, case):
    O = oracle_mod
    kind, geom, grid, xyz, r, per, sph = CASES[case]()
    ids = np.arange(len(xyz), dtype=np.int32)
    _flush("case %s n=%d" % (case, len(ids)))
    a = O.run("restatement", O.make_spec(kind, geom, grid, periodic=per, spheres=sph), ids, xyz, r)
    b = engine_run(kind, geom, grid, ids, xyz, r, periodic=per, spheres=sph)
    assert O.compare(b, a, vol_rtol=VOL_RTOL) == []
    assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])


def test_gpu_tier_escalation(oracle_mod):
    O = oracle_mod
    xyz, L = I.fcc_sphere(1, 3000, 1.0, 0.0)
    ids = np.arange(len(xyz), dtype=np.int32)
    sph = [[L / 2, L / 2, L / 2, 0.45 * L, -99]]
    _flush("tier escalation n=%d" % len(ids))
    a = O.run("restatement", O.make_spec(0, (0, L, 0, L, 0, L), (7, 7, 7), spheres=sph), ids, xyz)
    b = engine_run(0, (0, L, 0, L, 0, L), (7, 7, 7), ids, xyz, spheres=sph, stats=True)
    assert (b["st_tier"] > 0).any()
    assert O.compare(b, a, vol_rtol=VOL_RTOL) == []


def test_gpu_stats_match_restatement(oracle_mod):
    O = oracle_mod
    xyz = I.uniform_box(9, 6000, (0, 0, 0), (20, 20, 20))
    ids = np.arange(6000, dtype=np.int32)
    _flush("stats")
    a = O.run("restatement", O.make_spec(0, (0, 20, 0, 20, 0, 20), (10, 10, 10)), ids, xyz, stats=True)
    b = engine_run(0, (0, 20, 0, 20, 0, 20), (10, 10, 10), ids, xyz, stats=True)
    assert np.array_equal(a["st_n_planes"], b["st_n_planes"])
    assert np.array_equal(a["st_wl_exit"], b["st_wl_exit"])


def test_gpu_edge_inputs(oracle_mod):
    O = oracle_mod
    _flush("edge inputs")
    # empty container
    out = engine_run(0, (0, 1, 0, 1, 0, 1), (3, 3, 3), np.zeros(0, np.int32), np.zeros((0, 3)))
    assert len(out["id"]) == 0
    # one particle; a particle exactly on the upper bound is dropped
    xyz = np.array([[0.25, 0.5, 0.75], [1.0, 0.5, 0.5], [0.5, 0.5, 1.0]])
    out = engine_run(0, (0, 1, 0, 1, 0, 1), (3, 3, 3), np.array([4, 5, 6], np.int32), xyz)
    assert out["id"].tolist() == [4] and out["vol"][0] == 1.0 and out["nfaces"][0] == 6
    # many particles in one block (ragged occupancy)
    xyz = np.concatenate([I.uniform_box(2, 500, (0, 0, 0), (0.1, 0.1, 0.1)), I.uniform_box(3, 50, (0, 0, 0), (1, 1, 1))])
    ids = np.arange(len(xyz), dtype=np.int32)
    a = O.run("restatement", O.make_spec(0, (0, 1, 0, 1, 0, 1), (4, 4, 4)), ids, xyz)
    b = engine_run(0, (0, 1, 0, 1, 0, 1), (4, 4, 4), ids, xyz)
    assert O.compare(b, a, vol_rtol=VOL_RTOL) == []


def test_gpu_repeated_compute_is_identical():
    """Reusing a container (and its workspace stamps) gives identical results."""
    xyz = I.uniform_box(12, 20000, (0, 0, 0), (30, 30, 30))
    ids = np.arange(len(xyz), dtype=np.int32)
    _flush("repeat")
    r1 = engine_run(0, (0, 30, 0, 30, 0, 30), (16, 16, 16), ids, xyz)
    con = r1["_con"]
    con.compute_all_cells()
    r2 = con.results()
    for k in ("id", "exists", "nfaces", "vol", "cent", "nbr", "nbr_off"):
        assert np.array_equal(r1[k], r2[k]), k


def test_gpu_full_size_uniform(oracle_mod):
    """The benchmark workload (1,000,000 uniform particles, [0,100]^3, 58^3
    blocks), bit for bit against the reference library itself (run on the
    host's cores over z-slabs, outputs identical to its single-threaded
    loop), plus volume closure."""
    O = oracle_mod
    ids, xyz, _ = I.make("uniform", 42)
    _flush("full size uniform 1M")
    b = engine_run(0, (0, 100, 0, 100, 0, 100), (58, 58, 58), ids, xyz)
    assert abs(b["vol"].sum() - 1e6) / 1e6 < 1e-12
    _flush("full size uniform 1M: reference on %d threads" % O.default_threads())
    a = O.run("reference", O.make_spec(0, (0, 100, 0, 100, 0, 100), (58, 58, 58)), ids, xyz,
              threads=O.default_threads())
    assert O.compare(b, a, vol_rtol=VOL_RTOL) == []
    assert np.array_equal(a["vol"], b["vol"])


@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_first_tiers_agree(oracle_mod, case):
    """One cell per wavefront (default) and one order-3 cell per lane as the
    first tier: both bit-exact against the restatement."""
    O = oracle_mod
    kind, geom, grid, xyz, r, per, sph = CASES[case]()
    ids = np.arange(len(xyz), dtype=np.int32)
    _flush("first tiers %s n=%d" % (case, len(ids)))
    a = O.run("restatement", O.make_spec(kind, geom, grid, periodic=per, spheres=sph), ids, xyz, r)
    for tier in (0, 1):
        b = engine_run(kind, geom, grid, ids, xyz, r, periodic=per, spheres=sph, stats=True, first_tier=tier)
        assert O.compare(b, a, vol_rtol=VOL_RTOL) == []
        assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])


def test_gpu_far_block_searches_take_the_wide_window_wavefront_tier(oracle_mod):
    """Isolated points beside a dense cluster, with small blocks: their cells'
    block searches reach well beyond the +-6-block window of the 64- and
    128-vertex wavefront tiers.  The 254-vertex tier, whose window reaches
    +-31 blocks, finishes such a cell (stats tier 1), bit-exact; none reaches
    the per-lane tiers."""
    O = oracle_mod
    rng = np.random.default_rng(21)
    dense = rng.uniform(0, 8, (12000, 3))
    far = rng.uniform(0, 40, (80, 3))
    xyz = np.concatenate([dense, far])
    ids = np.arange(len(xyz), dtype=np.int32)
    sp = (0, (0, 40, 0, 40, 0, 40), (40, 40, 40))
    _flush("far block searches n=%d" % len(xyz))
    a = O.run("restatement", O.make_spec(*sp), ids, xyz, stats=True)
    b = engine_run(*sp, ids, xyz, stats=True)
    assert O.compare(b, a, vol_rtol=VOL_RTOL) == []
    assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])
    assert (b["st_wl_exit"] >= 64).sum() >= 100 and (b["st_tier"] == 1).sum() >= 100
    assert (b["st_tier"] >= 2).sum() == 0


def test_gpu_large_cells_take_the_wide_wavefront_tiers(oracle_mod):
    """Radical cells of a few large particles among small ones have more
    than 64 (and more than 128) vertices: the 128- and 254-vertex wavefront
    tiers (stats tier 1) finish them, bit-exact."""
    O = oracle_mod
    n = 6000
    xyz = I.uniform_box(11, n, (0, 0, 0), (20, 20, 20))
    r = np.full(n, 0.3)
    r[::100] = 2.0
    ids = np.arange(n, dtype=np.int32)
    _flush("large radical cells n=%d" % n)
    a = O.run("restatement", O.make_spec(1, (0, 20, 0, 20, 0, 20), (10, 10, 10)), ids, xyz, r, stats=True)
    b = engine_run(1, (0, 20, 0, 20, 0, 20), (10, 10, 10), ids, xyz, r, stats=True)
    assert (a["st_max_p"] > 128).any() and (b["st_tier"] == 1).sum() >= (a["st_max_p"] > 64).sum()
    assert O.compare(b, a, vol_rtol=VOL_RTOL) == []
    assert np.array_equal(a["vol"], b["vol"])
