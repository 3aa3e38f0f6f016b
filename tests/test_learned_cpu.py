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
