"""GPU checks of the single-container autotuner (vgpu_autotune): every
configuration it tries gives the same cells (checked against the
restatement), it keeps the fastest on the handle, and the table it returns
(and writes as JSON) lists every configuration tried with its time."""
import json

import numpy as np
import pytest

import voro_amd
from voro_amd import inputs as I

pytestmark = pytest.mark.gpu

F = voro_amd.F_NEIGHBORS | voro_amd.F_CSR


def _container(seed, n=60000, L=39.1, G=23):
    xyz = I.uniform_box(seed, n, (0, 0, 0), (L, L, L))
    con = voro_amd.container(0, L, 0, L, 0, L, G, G, G, False, False, False, 8)
    con.put_array(np.arange(n, dtype=np.int32), xyz)
    return con, xyz, (0, L, 0, L, 0, L), (G, G, G)


def test_gpu_autotune_plain_keeps_results_and_writes_table(oracle_mod, tmp_path):
    O = oracle_mod
    con, xyz, geom, grid = _container(81)
    print("autotune plain n=%d" % len(xyz), flush=True)
    path = tmp_path / "tune.json"
    t = con.autotune(F, reps=1, path=str(path), config="uniform 60k")
    kinds = {(r["first_tier"], r["occupancy"]) for r in t["trials"]}
    assert kinds == {("wavefront", 3), ("wavefront", 4), ("wavefront", 5), ("wavefront", 6),
                     ("lane", 2), ("lane", 3), ("lane", 4)}
    assert t["best"]["ms"] <= 1.01 * min(r["ms"] for r in t["trials"])
    assert json.load(open(path))["best"] == t["best"]
    # the kept configuration computes the oracle's cells
    con.compute_all_cells(F)
    want = O.run("restatement", O.make_spec(0, geom, grid), np.arange(len(xyz), dtype=np.int32), xyz)
    got = con.results()
    assert O.compare(got, want, vol_rtol=1e-12) == []
    assert np.array_equal(got["vol"], want["vol"])


def test_gpu_autotune_learned_depth_margin(oracle_mod):
    O = oracle_mod
    pred = voro_amd.predictor()
    for sd in (82, 83):
        c, _, _, _ = _container(sd)
        c.compute_all_cells(F | voro_amd.F_STATS)
        pred.add_samples(c)
    pred.train(epochs=200, lr=3e-3, seed=1)
    con, xyz, geom, grid = _container(84)
    con.set_predictor(pred)
    print("autotune learned n=%d" % len(xyz), flush=True)
    t = con.autotune(F | voro_amd.F_LEARNED, reps=1)
    offs = sorted(r["depth_offset"] for r in t["trials"] if r["first_tier"] == "wavefront")
    assert len(t["trials"]) == 9 and max(offs) == 4.0
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    want = O.run("restatement", O.make_spec(0, geom, grid), np.arange(len(xyz), dtype=np.int32), xyz)
    got = con.results()
    assert O.compare(got, want, vol_rtol=1e-12) == []
    assert np.array_equal(got["vol"], want["vol"])
