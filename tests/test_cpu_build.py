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
