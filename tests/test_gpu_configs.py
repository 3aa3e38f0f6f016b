"""The benchmark configurations at their full sizes, against the reference
library itself (oracle/_ref, run on the host's cores over z-slabs of blocks,
outputs identical to its single-threaded loop) or the restatement:

* the clustered workload with the predictor on, as configured: 1,000,000
  points (64 Gaussian clusters + 30% uniform background) in [0,100]^3 with
  58^3 blocks, predictor trained on the plain path's statistics for seeds
  1..10, evaluated on seed 101;
* the 4,096-container FCC-jitter batch computed by one captured-graph launch,
  32 containers sampled across sizes checked against the restatement, and
  every container's volume closure checked against the wall sphere's cut.

Neighbour lists, face counts and existence bit-exact, volumes within 1e-12
relative (they come out bit-identical)."""
import numpy as np
import pytest

import voro_amd
from voro_amd import inputs as I

pytestmark = pytest.mark.gpu

F = voro_amd.F_NEIGHBORS | voro_amd.F_CSR
VOL_RTOL = 1e-12


def _flush(msg):
    print(msg, flush=True)


def test_gpu_clusters_learned_full_size_matches_reference(oracle_mod):
    O = oracle_mod

    def make(seed):
        ids, xyz, _ = I.make("clusters", seed)
        con = voro_amd.container(0, 100, 0, 100, 0, 100, 58, 58, 58, False, False, False, 8)
        con.put_array(ids, xyz)
        return con, ids, xyz

    pred = voro_amd.predictor()
    for seed in range(1, 11):
        _flush("clusters: training statistics, seed %d" % seed)
        c, _, _ = make(seed)
        c.compute_all_cells(F | voro_amd.F_STATS)
        pred.add_samples(c)
        del c
    loss = pred.train(epochs=400, lr=3e-3, seed=1, fallback=0.02)
    _flush("clusters: trained, loss %.4g" % loss)
    con, ids, xyz = make(101)
    con.set_predictor(pred)
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    got = con.results()
    n = len(got["id"])
    fb = con.learned_stats()["fallbacks"]
    _flush("clusters: learned path done, fallback rate %.4f; reference on %d threads" % (fb / n, O.default_threads()))
    want = O.run("reference", O.make_spec(0, (0, 100, 0, 100, 0, 100), (58, 58, 58)), ids, xyz,
                 threads=O.default_threads())
    assert O.compare(got, want, vol_rtol=VOL_RTOL) == []
    assert np.array_equal(got["vol"], want["vol"])
    assert abs(float(got["vol"].sum()) - 1e6) < 1e-12 * 1e6
    assert 0 < fb < n


def test_gpu_batch_4096_graph_sample_matches_oracle(oracle_mod):
    O = oracle_mod
    _flush("batch 4096: building")
    specs = I.fcc_batch(11, 4096)
    cons = []
    for L, g, pts in specs:
        c = voro_amd.container(0, L, 0, L, 0, L, g, g, g, False, False, False, 8)
        c.add_wall(voro_amd.wall_sphere(L / 2, L / 2, L / 2, 0.45 * L, -99))
        c.put_array(np.arange(len(pts), dtype=np.int32), pts)
        cons.append(c)
    b = voro_amd.container_batch(cons)
    del cons
    b.compute(F, graph=True)
    b.compute(F, graph=True)      # a replay of the captured graph
    _flush("batch 4096: computed (%d particles)" % sum(len(s[2]) for s in specs))
    sizes = np.array([len(s[2]) for s in specs])
    order = np.argsort(sizes, kind="stable")
    sample = sorted(set(int(order[int(t)]) for t in np.linspace(0, len(specs) - 1, 32)))
    for i in sample:
        L, g, pts = specs[i]
        sp = O.make_spec(0, (0, L, 0, L, 0, L), (g, g, g), spheres=[[L / 2, L / 2, L / 2, 0.45 * L, -99]])
        want = O.run("restatement", sp, np.arange(len(pts), dtype=np.int32), pts)
        got = b.results(i)
        assert O.compare(got, want, vol_rtol=VOL_RTOL) == [], i
        assert np.array_equal(got["vol"], want["vol"]), i
    # every container: the cells fill the box cut by the wall's tangent
    # planes, so the volume sum lies between the sphere's and the box's
    sums = b.sum_cell_volumes()
    for i, (L, g, pts) in enumerate(specs):
        if len(pts):
            assert 4 / 3 * np.pi * (0.45 * L) ** 3 * (1 - 1e-12) <= sums[i] <= L ** 3, i
