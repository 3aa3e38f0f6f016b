"""GPU parity of batches of containers (vgpu_batch_*): every container's
results from one batched compute equal the oracle's for that container
alone, bit for bit (neighbours, face counts, existence, volumes), for the
plain and radical kinds, with sphere walls, empty containers and dropped
particles; graph replays (also after new positions) equal eager computes;
both first tiers (one cell per wavefront, one cell per lane) and the
autotuner's choice do not change results."""
import numpy as np
import pytest

import voro_amd
from voro_amd import inputs as I

pytestmark = pytest.mark.gpu

VOL_RTOL = 1e-12


def _flush(msg):
    print(msg, flush=True)


def _fcc_containers(seed, count, n_hi=3000):
    specs = I.fcc_batch(seed, count, n_lo=100, n_hi=n_hi)
    cons = []
    for L, g, xyz in specs:
        con = voro_amd.container(0, L, 0, L, 0, L, g, g, g, False, False, False, 8)
        con.add_wall(voro_amd.wall_sphere(L / 2, L / 2, L / 2, 0.45 * L, -99))
        con.put_array(np.arange(len(xyz), dtype=np.int32), xyz)
        cons.append(con)
    return specs, cons


def _oracle_fcc(O, L, g, xyz):
    sp = O.make_spec(0, (0, L, 0, L, 0, L), (g, g, g), spheres=[[L / 2, L / 2, L / 2, 0.45 * L, -99]])
    return O.run("restatement", sp, np.arange(len(xyz), dtype=np.int32), xyz)


def test_gpu_batch_fcc_matches_oracle(oracle_mod):
    O = oracle_mod
    specs, cons = _fcc_containers(3, 24)
    _flush("batch fcc 24 containers, %d particles" % sum(len(s[2]) for s in specs))
    b = voro_amd.container_batch(cons)
    b.compute()
    for i, (L, g, xyz) in enumerate(specs):
        got = b.results(i)
        want = _oracle_fcc(O, L, g, xyz)
        assert O.compare(got, want, vol_rtol=VOL_RTOL) == [], i
        assert np.array_equal(got["vol"], want["vol"])
    sums = b.sum_cell_volumes()
    for i, (L, g, xyz) in enumerate(specs):
        assert abs(sums[i] - float(np.sum(b.results(i)["vol"]))) < 1e-9 * sums[i]


def test_gpu_batch_poly_uneven_matches_oracle(oracle_mod):
    """Radical containers of very different sizes and boxes, an empty one,
    and one whose particles are all outside its box (dropped)."""
    O = oracle_mod
    cases = []
    for c, n in enumerate([5, 3000, 0, 40, 800, 12000, 1]):
        box = 5.0 + 3 * c
        g = max(1, int(round((max(n, 1) / 5.0) ** (1 / 3))))
        xyz = I.uniform_box(50 + c, n, (0, 0, 0), (box, box, box))
        r = I.radii_lognormal(60 + c, n) * box / 20
        cases.append((box, g, xyz, r))
    box, g = 4.0, 2
    cases.append((box, g, I.uniform_box(77, 30, (5, 5, 5), (9, 9, 9)), I.radii_lognormal(77, 30)))
    cons = []
    for box, g, xyz, r in cases:
        con = voro_amd.container_poly(0, box, 0, box, 0, box, g, g, g, False, False, False, 8)
        con.put_array(np.arange(len(xyz), dtype=np.int32), xyz, r)
        cons.append(con)
    _flush("batch poly %d containers" % len(cons))
    b = voro_amd.container_batch(cons)
    b.compute()
    for i, (box, g, xyz, r) in enumerate(cases):
        got = b.results(i)
        want = O.run("restatement", O.make_spec(1, (0, box, 0, box, 0, box), (g, g, g)),
                     np.arange(len(xyz), dtype=np.int32), xyz, r)
        assert O.compare(got, want, vol_rtol=VOL_RTOL) == [], i


def test_gpu_batch_first_tiers_and_leftovers_match_oracle(oracle_mod):
    """Radical containers with a few large particles among small ones: their
    cells exceed the 64 vertices of the wavefront-cell first tier (more than
    34 faces), so they pass to the order-3 per-lane tier and, beyond 128
    vertices, to the wide tier.  Both first tiers give the oracle's results
    bit for bit."""
    O = oracle_mod
    cases = []
    for c in range(3):
        n = 3000 + 1000 * c
        xyz = I.uniform_box(21 + c, n, (0, 0, 0), (16, 16, 16))
        r = np.full(n, 0.3)
        r[c::90] = 2.0
        cases.append((xyz, r))
    cons = []
    want = []
    for xyz, r in cases:
        con = voro_amd.container_poly(0, 16, 0, 16, 0, 16, 8, 8, 8, False, False, False, 8)
        con.put_array(np.arange(len(xyz), dtype=np.int32), xyz, r)
        cons.append(con)
        want.append(O.run("restatement", O.make_spec(1, (0, 16, 0, 16, 0, 16), (8, 8, 8)),
                          np.arange(len(xyz), dtype=np.int32), xyz, r))
    assert max(int(w["nfaces"].max()) for w in want) > 66     # cells over 128 vertices
    b = voro_amd.container_batch(cons)
    for cfg in ((0, 0, 16), (0, 0, 1), (0, 5, 16), (0, 6, 4), (0, 2, 64)):
        _flush("batch first tier config %s" % (cfg,))
        b.set_config(*cfg)
        b.compute()
        for i in range(len(cases)):
            got = b.results(i)
            assert O.compare(got, want[i], vol_rtol=0.0) == [], (cfg, i)
            assert np.array_equal(got["vol"], want[i]["vol"]), (cfg, i)


def test_gpu_batch_equals_single_compute():
    specs, cons = _fcc_containers(4, 12)
    _flush("batch vs single")
    b = voro_amd.container_batch(cons)
    b.compute()
    for i, con in enumerate(cons):
        con.compute_all_cells(voro_amd.F_NEIGHBORS | voro_amd.F_CSR)
        a = con.results()
        got = b.results(i)
        for k in ("id", "ijk", "q", "exists", "nfaces", "vol", "cent", "pos", "nbr_off", "nbr"):
            assert np.array_equal(a[k], got[k]), (i, k)


def test_gpu_batch_graph_replays(oracle_mod):
    """A captured graph replayed (twice, then after new positions) gives the
    eager results."""
    O = oracle_mod
    specs, cons = _fcc_containers(5, 16)
    _flush("batch graph")
    b = voro_amd.container_batch(cons)
    b.compute()
    eager = [b.results(i) for i in range(len(specs))]
    for rep in range(2):
        b.compute(graph=True)
        for i in range(len(specs)):
            got = b.results(i)
            for k in ("id", "exists", "nfaces", "vol", "cent", "nbr"):
                assert np.array_equal(eager[i][k], got[k]), (rep, i, k)
    # new positions: a small shift inside every container
    allxyz = np.concatenate([s[2] for s in specs])
    moved = allxyz + 0.01 * np.stack([I.normal(9, 500 + 2 * a, len(allxyz)) for a in range(3)], 1)
    b.set_positions(moved)
    b.compute(graph=True)
    replay = [b.results(i) for i in range(len(specs))]
    b.compute()
    off = 0
    for i, (L, g, xyz) in enumerate(specs):
        got = b.results(i)
        for k in ("id", "exists", "nfaces", "vol", "cent", "nbr"):
            assert np.array_equal(replay[i][k], got[k]), (i, k)
        if i < 3:
            mv = moved[off:off + len(xyz)]
            sp = O.make_spec(0, (0, L, 0, L, 0, L), (g, g, g), spheres=[[L / 2, L / 2, L / 2, 0.45 * L, -99]])
            want = O.run("restatement", sp, np.arange(len(mv), dtype=np.int32), mv)
            assert O.compare(replay[i], want, vol_rtol=VOL_RTOL) == [], i
        off += len(xyz)


def test_gpu_batch_autotune_keeps_results():
    specs, cons = _fcc_containers(6, 10)
    _flush("batch autotune")
    b = voro_amd.container_batch(cons)
    b.compute()
    before = [b.results(i) for i in range(len(specs))]
    best = b.autotune(reps=1)
    assert best["occupancy"] in (0, 2, 3, 4) and best["chunk"] in (16, 32, 64) and best["ms"] > 0
    b.compute()
    for i in range(len(specs)):
        got = b.results(i)
        for k in ("id", "exists", "nfaces", "vol", "nbr"):
            assert np.array_equal(before[i][k], got[k]), (i, k)
