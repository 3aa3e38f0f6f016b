"""GPU checks of the learned path: the predictor's kernels against the numpy
restatement (oracle/mlp.py), training on statistics the plain path records,
results with the predictor on equal to the oracle's (the restatement, or the
reference library itself) bit for bit, forced fallback, and what the trained
network saves over gathering every cell's lists to full depth."""
import numpy as np
import pytest

import voro_amd
from oracle import mlp as M
from voro_amd import inputs as I

pytestmark = pytest.mark.gpu

F = voro_amd.F_NEIGHBORS | voro_amd.F_CSR


def _flush(msg):
    print(msg, flush=True)


def test_gpu_mlp_gradient_matches_numpy():
    rng = np.random.default_rng(7)
    p = rng.normal(0, 0.3, M.NP).astype(np.float32)
    X = rng.uniform(0, 2, (3000, M.IN)).astype(np.float32)
    T = rng.uniform(0, 1, (3000, M.OUT)).astype(np.float32)
    _flush("mlp gradient n=3000")
    lo, g = voro_amd.mlp_loss_grad(p, X, T)
    lo_ref, g_ref = M.loss_grad(p.astype(np.float64), X.astype(np.float64), T.astype(np.float64))
    assert abs(lo - lo_ref) <= 1e-5 * abs(lo_ref)
    scale = np.max(np.abs(g_ref))
    assert np.max(np.abs(g - g_ref)) <= 2e-5 * scale
    Y = voro_amd.mlp_forward(p, X)
    assert np.allclose(Y, M.forward(p.astype(np.float64), X.astype(np.float64))[2], rtol=1e-5, atol=1e-5)


def _uniform(seed, n, L, G):
    xyz = I.uniform_box(seed, n, (0, 0, 0), (L, L, L))
    con = voro_amd.container(0, L, 0, L, 0, L, G, G, G, False, False, False, 8)
    con.put_array(np.arange(n, dtype=np.int32), xyz)
    con.oracle_inputs = (0, (0, L, 0, L, 0, L), (G, G, G), xyz, None)
    return con


def _train(make, seeds, epochs=300):
    pred = voro_amd.predictor()
    for sd in seeds:
        c = make(sd)
        c.compute_all_cells(F | voro_amd.F_STATS)
        pred.add_samples(c)
    loss0 = pred.train(epochs=1, lr=3e-3, seed=1)
    loss = pred.train(epochs=epochs, lr=3e-3, seed=1)
    return pred, loss0, loss


def _same(a, b):
    for k in ("id", "exists", "nfaces", "nbr_off", "nbr"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])


def _vs_oracle(O, got, con, which="restatement"):
    """The learned path's results against the oracle run on the container's
    own inputs: neighbours, face counts and existence bit-exact, volumes
    within 1e-12 relative (they come out bit-identical)."""
    kind, geom, grid, xyz, r = con.oracle_inputs
    want = O.run(which, O.make_spec(kind, geom, grid), np.arange(len(xyz), dtype=np.int32), xyz, r,
                 threads=O.default_threads() if which == "reference" else None)
    assert O.compare(got, want, vol_rtol=1e-12) == []
    assert np.array_equal(got["vol"], want["vol"])


def test_gpu_learned_uniform_is_exact_and_saves_work(oracle_mod):
    n, L, G = 100000, 46.4, 27
    make = lambda sd: _uniform(sd, n, L, G)
    _flush("train uniform n=%d" % n)
    pred, loss0, loss = _train(make, [11, 12])
    assert loss < 0.5 * loss0
    con = make(13)
    con.compute_all_cells(F)
    plain = dict(con.results())
    con.set_predictor(pred)
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    con._res = None
    got = con.results()
    _vs_oracle(oracle_mod, got, con)
    _same(plain, got)
    st = con.learned_stats()
    _flush("fallbacks %d of %d, entries %d" % (st["fallbacks"], n, st["list_entries"]))
    assert st["fallbacks"] < 0.1 * n
    # without the network: every list gathered to the full worklist
    pred.margin = 1000.0
    pred.route = 65
    con.set_predictor(pred)
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    con._res = None
    got = con.results()
    _vs_oracle(oracle_mod, got, con)
    _same(plain, got)
    full = con.learned_stats()
    _flush("entries trained %d, full depth %d" % (st["list_entries"], full["list_entries"]))
    assert st["list_entries"] < full["list_entries"]


def test_gpu_learned_per_lane_lists_are_exact(oracle_mod):
    """With the per-lane first tier the lists are interleaved per batch of 64
    cells and cut by the per-lane list kernel: the same results."""
    n, L, G = 50000, 36.8, 21
    make = lambda sd: _uniform(sd, n, L, G)
    pred, _, _ = _train(make, [31], epochs=100)
    con = make(32)
    con.compute_all_cells(F)
    plain = dict(con.results())
    con.set_first_tier(1)
    con.set_predictor(pred)
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    con._res = None
    got = con.results()
    _vs_oracle(oracle_mod, got, con)
    _same(plain, got)


def test_gpu_learned_dense_blocks_are_exact(oracle_mod):
    """About 250 particles per block: a checked block's list entries are
    radius-checked 64 at a time in several rounds; the results equal the
    plain path's."""
    n, L, G = 32000, 20.0, 5
    make = lambda sd: _uniform(sd, n, L, G)
    pred, _, _ = _train(make, [41], epochs=100)
    con = make(42)
    con.compute_all_cells(F)
    plain = dict(con.results())
    con.set_predictor(pred)
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    con._res = None
    got = con.results()
    _vs_oracle(oracle_mod, got, con)
    _same(plain, got)


def test_gpu_learned_forced_fallback_is_exact(oracle_mod):
    n, L, G = 20000, 27.1, 16
    make = lambda sd: _uniform(sd, n, L, G)
    pred, _, _ = _train(make, [21], epochs=50)
    pred.margin = -1000.0     # depth 1 for every block: lists end at once
    con = make(22)
    con.compute_all_cells(F)
    plain = dict(con.results())
    con.set_predictor(pred)
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    con._res = None
    got = con.results()
    _vs_oracle(oracle_mod, got, con)
    _same(plain, got)
    assert con.learned_stats()["fallbacks"] > 0.9 * n


def test_gpu_learned_clusters_held_out_seed(oracle_mod):
    n, box, G = 100000, 46.4, 27

    def make(sd):
        xyz = I.clusters(sd, n, box=box, n_clusters=64)
        con = voro_amd.container(0, box, 0, box, 0, box, G, G, G, False, False, False, 8)
        con.put_array(np.arange(len(xyz), dtype=np.int32), xyz)
        con.oracle_inputs = (0, (0, box, 0, box, 0, box), (G, G, G), xyz, None)
        return con
    _flush("train clusters n=%d" % n)
    pred, loss0, loss = _train(make, [1, 2, 3])
    con = make(101)
    con.compute_all_cells(F)
    plain = dict(con.results())
    con.set_predictor(pred)
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    con._res = None
    got = con.results()
    _vs_oracle(oracle_mod, got, con, which="reference")
    _same(plain, got)
    _flush("clusters fallback rate %.4f" % (con.learned_stats()["fallbacks"] / n))


def test_gpu_learned_radical_per_particle(oracle_mod):
    n, L, G = 30000, 31.0, 18

    def make(sd):
        xyz = I.uniform_box(sd, n, (0, 0, 0), (L, L, L))
        r = I.radii_lognormal(sd, n)
        con = voro_amd.container_poly(0, L, 0, L, 0, L, G, G, G, False, False, False, 8)
        con.put_array(np.arange(n, dtype=np.int32), xyz, r)
        con.oracle_inputs = (1, (0, L, 0, L, 0, L), (G, G, G), xyz, r)
        return con
    pred = voro_amd.predictor(with_radius=True)
    for sd in (31, 32):
        c = make(sd)
        c.compute_all_cells(F | voro_amd.F_STATS)
        pred.add_samples(c)
    pred.train(epochs=200)
    con = make(33)
    con.compute_all_cells(F)
    plain = dict(con.results())
    con.set_predictor(pred)
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    con._res = None
    got = con.results()
    _vs_oracle(oracle_mod, got, con)
    _same(plain, got)


# ---- periodic containers: the list kernel records the image blocks' first
# visits with the keys the plain path gives them ----

PRD_BOX = (30.0, 2.5, 30.0, 1.5, 2.0, 30.0)


def _prd_poly(sd, g=37):
    """The radical triclinic workload at 1/8 the volume and particle count
    (same density): container_periodic_poly(30,2.5,30,1.5,2,30), 37^3 blocks,
    250,000 particles."""
    ids, xyz, r = I.make("periodic_poly", sd, 250000)
    xyz = xyz * 0.5
    con = voro_amd.container_periodic_poly(*PRD_BOX, g, g, g)
    con.put_array(ids, xyz, r)
    con.oracle_inputs = (3, PRD_BOX, (g, g, g), xyz, r)
    return con


def test_gpu_learned_periodic_poly_vs_reference(oracle_mod):
    """The radical periodic workload with the radius-input predictor: a
    250,000-particle sample checked against the reference library cell for
    cell, and against the plain path bit for bit."""
    if oracle_mod.reference() is None:
        pytest.skip("reference library not built")
    pred = voro_amd.predictor(with_radius=True)
    for sd in (51, 52):
        _flush("periodic train seed %d" % sd)
        c = _prd_poly(sd)
        c.compute_all_cells(F | voro_amd.F_STATS)
        pred.add_samples(c)
    pred.train(epochs=200)
    con = _prd_poly(53)
    con.compute_all_cells(F)
    plain = dict(con.results())
    con.set_predictor(pred)
    _flush("periodic learned compute")
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    con._res = None
    got = con.results()
    _same(plain, got)
    st = con.learned_stats()
    n = len(got["id"])
    _flush("periodic fallbacks %d, given no list %d, of %d; entries %d"
           % (st["fallbacks"], st["routed"], n, st["list_entries"]))
    # about 60% of these cells reach the block search, which no list covers:
    # the predictor sends many of them to the plain path at once
    assert st["routed"] > 0.25 * n and st["fallbacks"] < 0.4 * n
    kind, geom, grid, xyz, r = con.oracle_inputs
    want = oracle_mod.run("reference", oracle_mod.make_spec(kind, geom, grid), np.arange(len(xyz), dtype=np.int32),
                          xyz, r)
    assert oracle_mod.compare(got, want, vol_rtol=1e-12) == []
    assert np.array_equal(got["vol"], want["vol"])


@pytest.mark.parametrize("margin", [-1000.0, None, 1000.0], ids=["depth1", "trained", "full"])
def test_gpu_learned_periodic_depths_match_reference(oracle_mod, margin):
    """A small sheared periodic container at list depth 1 (every cell falls
    back), as trained, and at full depth: the same cells and neighbour orders
    as the reference."""
    if oracle_mod.reference() is None:
        pytest.skip("reference library not built")
    box = (8.0, 1.3, 8.0, -0.9, 2.1, 8.0)
    n, g = 3000, 7

    def make(sd):
        u = np.stack([I.uniform(sd, 90 + a, n) for a in range(3)], 1) * 1.3 - 0.15
        xyz = np.stack([u[:, 0] * box[0] + u[:, 1] * box[1] + u[:, 2] * box[3], u[:, 1] * box[2] + u[:, 2] * box[4],
                        u[:, 2] * box[5]], 1)
        con = voro_amd.container_periodic(*box, g, g, g)
        con.put_array(np.arange(n, dtype=np.int32), xyz)
        con.oracle_inputs = (2, box, (g, g, g), xyz, None)
        return con
    pred, _, _ = _train(make, [61], epochs=100)
    if margin is not None:
        pred.margin = margin
        pred.route = 65
    con = make(62)
    con.compute_all_cells(F)
    plain = dict(con.results())
    con.set_predictor(pred)
    con.compute_all_cells(F | voro_amd.F_LEARNED)
    con._res = None
    got = con.results()
    _same(plain, got)
    kind, geom, grid, xyz, r = con.oracle_inputs
    want = oracle_mod.run("reference", oracle_mod.make_spec(kind, geom, grid), np.arange(n, dtype=np.int32), xyz, r)
    assert oracle_mod.compare(got, want, vol_rtol=1e-12) == []
    if margin is not None and margin < 0:
        assert con.learned_stats()["fallbacks"] > 0.9 * n
