"""GPU checks of the learned path on batches of containers: a block
predictor trained on a batch's own statistics, then every container's
results with it equal the plain batch's bit for bit (and the oracle's for a
sample of containers), eagerly and through graph replays (also after new
positions); lists that do not fit the list capacity, forced fallback and
full-depth lists change nothing but the statistics."""
import numpy as np
import pytest

import voro_amd
from voro_amd import inputs as I

pytestmark = pytest.mark.gpu

F = voro_amd.F_NEIGHBORS | voro_amd.F_CSR
VOL_RTOL = 1e-12


def _flush(msg):
    print(msg, flush=True)


def _fcc(seed, count, n_hi=6000):
    specs = I.fcc_batch(seed, count, n_lo=100, n_hi=n_hi)
    cons = []
    for L, g, xyz in specs:
        con = voro_amd.container(0, L, 0, L, 0, L, g, g, g, False, False, False, 8)
        con.add_wall(voro_amd.wall_sphere(L / 2, L / 2, L / 2, 0.45 * L, -99))
        con.put_array(np.arange(len(xyz), dtype=np.int32), xyz)
        cons.append(con)
    return specs, voro_amd.container_batch(cons)


def _trained(seed=5, count=96):
    _, tb = _fcc(seed, count)
    tb.compute(F | voro_amd.F_STATS)
    pred = voro_amd.predictor()
    pred.add_samples(tb)
    assert pred.num_samples() > 0
    pred.train(epochs=200)
    return pred


def _all(b, count):
    return [b.results(i) for i in range(count)]


def _same(a, b):
    for x, y in zip(a, b):
        for k in ("id", "exists", "nfaces", "nbr_off", "nbr"):
            assert np.array_equal(x[k], y[k]), k
        assert np.array_equal(x["vol"], y["vol"]) and np.array_equal(x["cent"], y["cent"])


def test_gpu_batch_learned_matches_plain_and_oracle(oracle_mod):
    O = oracle_mod
    pred = _trained()
    specs, b = _fcc(6, 64)
    n = sum(len(s[2]) for s in specs)
    _flush("batch learned: 64 containers, %d particles" % n)
    b.compute(F)
    plain = _all(b, len(specs))
    b.set_predictor(pred)
    b.compute(F | voro_amd.F_LEARNED)
    got = _all(b, len(specs))
    _same(plain, got)
    st = b.learned_stats()
    _flush("batch learned stats %s" % st)
    assert st["fallbacks"] < 0.2 * n and st["list_entries"] > 0 and st["list_need"] <= b.list_capacity
    # the list tier at 5 and 6 waves per SIMD
    for occ in (5, 6):
        _flush("batch learned: list tier at %d waves per SIMD" % occ)
        b.set_config(0, occ, 16)
        b.compute(F | voro_amd.F_LEARNED)
        _same(plain, _all(b, len(specs)))
    for i in (0, 7, 31, 63):
        L, g, xyz = specs[i]
        sp = O.make_spec(0, (0, L, 0, L, 0, L), (g, g, g), spheres=[[L / 2, L / 2, L / 2, 0.45 * L, -99]])
        want = O.run("restatement", sp, np.arange(len(xyz), dtype=np.int32), xyz)
        assert O.compare(got[i], want, vol_rtol=VOL_RTOL) == [], i
        assert np.array_equal(got[i]["vol"], want["vol"])


def test_gpu_batch_learned_graph_replays_and_new_positions():
    pred = _trained(7, 48)
    specs, b = _fcc(8, 48)
    b.set_predictor(pred)
    b.compute(F)
    plain = _all(b, len(specs))
    for rep in range(3):
        _flush("batch learned graph replay %d" % rep)
        b.compute(F | voro_amd.F_LEARNED, graph=True)
        _same(plain, _all(b, len(specs)))
        b.compute(F)          # eager work between the replays
    xyz = np.concatenate([s[2] for s in specs])
    rng = np.random.default_rng(3)
    moved = xyz + rng.normal(0, 0.01, xyz.shape)
    lo = 0
    for L, g, x in specs:
        m = len(x)
        moved[lo:lo + m] = np.clip(moved[lo:lo + m], 1e-9, L * (1 - 1e-12))
        lo += m
    b.set_positions(moved)
    b.compute(F)
    plain2 = _all(b, len(specs))
    b.compute(F | voro_amd.F_LEARNED, graph=True)
    _same(plain2, _all(b, len(specs)))


def test_gpu_batch_learned_capacity_depth_extremes():
    """A list capacity far too small (most lists not written: their cells take
    the plain path; the next compute is given room for all), depth 1 for
    every block (every list ends at once) and full-depth lists: the same
    results each time."""
    pred = _trained(9, 48)
    specs, b = _fcc(10, 32)
    n = sum(len(s[2]) for s in specs)
    b.compute(F)
    plain = _all(b, len(specs))
    b.set_predictor(pred)
    b.list_capacity = 1000
    b.compute(F | voro_amd.F_LEARNED)
    _same(plain, _all(b, len(specs)))
    st = b.learned_stats()
    _flush("small capacity: %s" % st)
    assert st["list_need"] > 1000 and st["routed"] > 0.5 * n
    assert b.list_capacity >= st["list_need"]
    b.compute(F | voro_amd.F_LEARNED)
    _same(plain, _all(b, len(specs)))
    assert b.learned_stats()["list_entries"] == b.learned_stats()["list_need"]
    for margin in (-1000.0, 1000.0):
        pred.margin = margin
        pred.route = 65
        b.set_predictor(pred)
        b.compute(F | voro_amd.F_LEARNED)
        _same(plain, _all(b, len(specs)))
        st = b.learned_stats()
        _flush("margin %g: %s" % (margin, st))
        if margin < 0:
            assert st["fallbacks"] > 0.9 * n
