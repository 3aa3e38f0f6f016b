"""CPU checks of the learned path's pieces: the numpy restatement of the
predictor network against finite differences (it is the checker of the GPU
kernels), and the list-driven cutting routine (compiled for the CPU) against
the oracle at several list depths, including depths short enough that most
cells fall back to the plain path."""
import os
import sys

import numpy as np
import pytest

from oracle import mlp as M
from voro_amd import inputs as I

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpu_build"))
import harness_py as H  # noqa: E402


def _net(seed, n):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.3, M.NP)
    X = rng.uniform(0, 2, (n, M.IN))
    T = rng.uniform(0, 1, (n, M.OUT))
    return p, X, T


def test_numpy_gradient_matches_finite_differences():
    p, X, T = _net(1, 50)
    lo, g = M.loss_grad(p, X, T)
    idx = list(range(0, M.NP, 37)) + [M.oB1 + 3, M.oB2 + 5, M.oW3 + 7, M.oB3, M.oB3 + 1]
    fd = M.finite_difference(p, X, T, idx)
    assert np.allclose(g[idx], fd, rtol=1e-6, atol=1e-9)
    assert abs(lo - M.loss(p, X, T)) < 1e-12


def test_numpy_gradient_catches_a_planted_mistake():
    """A dropped ReLU mask in the second layer must fail the same check."""
    p, X, T = _net(2, 40)
    W1, b1, W2, b2, W3, b3 = M.unpack(p)
    h1, h2, y = M.forward(p, X)
    dy = 2.0 * (y - T) / len(X)
    dz2_wrong = dy @ W3                        # mask (h2 > 0) dropped
    g_wrong = (dz2_wrong.T @ h1).ravel()
    idx = np.arange(M.oW2, M.oB2, 29)
    fd = M.finite_difference(p, X, T, idx)
    assert not np.allclose(g_wrong[idx - M.oW2], fd, rtol=1e-3, atol=1e-6)


@pytest.fixture(scope="module")
def harness():
    H.build()
    return H


CASES = {
    "uniform": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), None, I.uniform_box(3, 4000, (0, 0, 0), (20, 20, 20)),
                        None, (0, 0, 0)),
    "clusters": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), None, I.clusters(4, 4000, box=20.0, n_clusters=8),
                         None, (0, 0, 0)),
    "poly_periodic": lambda: (1, (0, 20, 0, 20, 0, 20), (10, 10, 10), None,
                              I.uniform_box(5, 4000, (0, 0, 0), (20, 20, 20)), I.radii_lognormal(5, 4000), (1, 1, 1)),
    "wall": lambda: (0, (0, 10, 0, 10, 0, 10), (5, 5, 5), [[5, 5, 5, 4.5, -99]],
                     I.uniform_box(7, 1500, (0, 0, 0), (10, 10, 10)), None, (0, 0, 0)),
}


@pytest.mark.parametrize("wave", [False, True], ids=["lane", "wave"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_list_cutting_on_cpu_matches_oracle(oracle_mod, harness, case, wave):
    """The list-driven cell loop, on the order-3 cell (one per lane) and on
    the wavefront cell (64 lanes emulated)."""
    O = oracle_mod
    kind, geom, grid, sph, xyz, r, per = CASES[case]()
    ids = np.arange(len(xyz), dtype=np.int32)
    sp = O.make_spec(kind, geom, grid, periodic=per, spheres=sph)
    a = O.run("restatement", sp, ids, xyz, r)
    last_fb = None
    for G in (3, 30, 63):
        b = harness.run(O, sp, ids, xyz, r, list_depth=G, wave=wave)
        assert O.compare(b, a, vol_rtol=0.0) == []
        assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])
        if last_fb is not None:
            assert b["list_fallbacks"] <= last_fb
        last_fb = b["list_fallbacks"]
    # depth 63 falls back only for cells that reach the block search
    st = harness.run(O, sp, ids, xyz, r)["stats"]
    assert last_fb == int(np.sum(st[:, 1] >= 64))


@pytest.mark.parametrize("case", sorted(CASES))
def test_list_resume_matches_the_plain_cell(oracle_mod, harness, case):
    """A wavefront cell whose list ends is continued from the list's last
    step instead of being computed again: the continued cell makes the same
    nplane calls as the plain cell (plane count and worklist exit equal), is
    in the same phase where the plain cell passed the list depth, and gives
    the same results as a cell computed again."""
    O = oracle_mod
    kind, geom, grid, sph, xyz, r, per = CASES[case]()
    ids = np.arange(len(xyz), dtype=np.int32)
    sp = O.make_spec(kind, geom, grid, periodic=per, spheres=sph)
    plain = harness.run(O, sp, ids, xyz, r, wave=True)
    for G in (1, 3, 8, 30, 63):
        a = harness.run(O, sp, ids, xyz, r, list_depth=G, wave=True, resume=False)
        b = harness.run(O, sp, ids, xyz, r, list_depth=G, wave=True, resume=True)
        assert b["list_fallbacks"] == a["list_fallbacks"] > 0 or G == 63
        assert O.compare(b, a, vol_rtol=0.0) == [], G
        assert np.array_equal(a["vol"], b["vol"]) and np.array_equal(a["cent"], b["cent"])
        assert np.array_equal(b["stats"][:, :2], plain["stats"][:, :2]), G
        deep = plain["stats"][:, 1] > G
        assert np.array_equal(b["stats"][deep, 2], plain["stats"][deep, 2]), G


def _para(geom, n, seed, lo=-0.15, hi=1.15):
    bx, bxy, by, bxz, byz, bz = geom
    u = np.stack([I.uniform(seed, 80 + a, n) for a in range(3)], 1) * (hi - lo) + lo
    return np.stack([u[:, 0] * bx + u[:, 1] * bxy + u[:, 2] * bxz, u[:, 1] * by + u[:, 2] * byz, u[:, 2] * bz], 1)


GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("wave", [False, True], ids=["lane", "wave"])
def test_list_cutting_periodic_matches_golden(oracle_mod, harness, wave):
    """Periodic containers through the lists: the list cell records each
    image block's first visit with the key the plain path would give it, so
    the replayed creation order, and with it every neighbour order, is the
    reference's (golden vectors written by the reference library)."""
    O = oracle_mod
    d = dict(np.load(os.path.join(GOLD, "prd_poly_small.npz"), allow_pickle=False))
    sp = O.make_spec(int(d["spec_kind"]), d["spec_geom"], d["spec_grid"])
    for G in (2, 12, 40, 63):
        b = harness.run(O, sp, d["ids"], d["xyz"], d.get("r"), list_depth=G, wave=wave)
        assert O.compare(b, d, vol_rtol=0.0) == [], G
        assert np.array_equal(b["vol"], d["vol"]) and np.array_equal(b["cent"], d["cent"])


PRD_CASES = {
    "triclinic_poly": (3, (10, 1.5, 10, 0.7, -1.2, 10), (5, 5, 5), 800, True),
    "skew": (3, (5, 7.5, 5, -6.2, 8.1, 5), (4, 4, 4), 600, True),
    "cubic": (2, (10, 0, 10, 0, 0, 10), (4, 4, 4), 500, False),
}


@pytest.mark.parametrize("wave", [False, True], ids=["lane", "wave"])
@pytest.mark.parametrize("case", sorted(PRD_CASES))
def test_list_cutting_periodic_matches_reference(oracle_mod, harness, case, wave):
    O = oracle_mod
    if O.reference() is None:
        pytest.skip("reference library not built")
    kind, geom, grid, n, poly = PRD_CASES[case]
    xyz = _para(geom, n, 300 + len(case))
    r = I.radii_lognormal(9, n) * 0.6 if poly else None
    ids = np.arange(n, dtype=np.int32)
    sp = O.make_spec(kind, geom, grid)
    a = O.run("reference", sp, ids, xyz, r)
    fb = []
    for G in (3, 20, 63):
        b = harness.run(O, sp, ids, xyz, r, list_depth=G, wave=wave)
        assert O.compare(b, a, vol_rtol=0.0) == [], G
        assert np.array_equal(a["pos"], b["pos"])
        fb.append(b["list_fallbacks"])
    assert fb[0] > 0 and fb == sorted(fb, reverse=True)
