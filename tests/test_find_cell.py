"""find_voronoi_cell (src/container.cc:237-297): the reference library's
answers pinned by brute force (nearest particle; least power distance for
radical containers), and the HIP engine's batched queries against the
reference, bit for bit (found flag, particle id, returned position)."""
import numpy as np
import pytest

import voro_amd
from voro_amd import inputs as I


def _ref_or_skip(O):
    if O.reference() is None:
        pytest.skip("reference library not built")


def test_reference_find_matches_brute_force(oracle_mod):
    O = oracle_mod
    _ref_or_skip(O)
    n = 3000
    xyz = I.uniform_box(13, n, (0, 0, 0), (12, 12, 12))
    r = I.radii_lognormal(13, n) * 0.6
    q = I.uniform_box(14, 400, (-1, -1, -1), (13, 13, 13))
    inside = np.all((q >= 0) & (q < 12), axis=1)
    d = ((q[:, None, :] - xyz[None, :, :]) ** 2).sum(-1)
    for kind, rr in ((0, None), (1, r)):
        sp = O.make_spec(kind, (0, 12, 0, 12, 0, 12), (6, 6, 6))
        found, pid, pos = O.find_reference(sp, np.arange(n, dtype=np.int32), xyz, rr, q)
        assert np.array_equal(found, inside)
        best = (d - (rr[None, :] ** 2 if rr is not None else 0)).argmin(axis=1)
        assert np.array_equal(pid[found], best[found])
        assert np.array_equal(pos[found], xyz[best[found]])


CASES = {
    # kind, geometry, grid, periodic flags, particles, radii, queries
    "plain": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), (0, 0, 0),
                      I.uniform_box(21, 6000, (0, 0, 0), (20, 20, 20)), None,
                      I.uniform_box(22, 20000, (-2, -2, -2), (22, 22, 22))),
    "poly": lambda: (1, (0, 20, 0, 20, 0, 20), (10, 10, 10), (0, 0, 0),
                     I.uniform_box(23, 6000, (0, 0, 0), (20, 20, 20)), I.radii_lognormal(23, 6000),
                     I.uniform_box(24, 20000, (0, 0, 0), (20, 20, 20))),
    "periodic": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), (1, 0, 1),
                         I.uniform_box(25, 6000, (0, 0, 0), (20, 20, 20)), None,
                         I.uniform_box(26, 20000, (-25, -2, -25), (45, 22, 45))),
    "clusters": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), (0, 0, 0),
                         I.clusters(27, 6000, box=20.0, n_clusters=6), None,
                         I.uniform_box(28, 20000, (0, 0, 0), (20, 20, 20))),
    # a few particles in many blocks: long block searches, some leaving the
    # first search window
    "sparse": lambda: (0, (0, 40, 0, 40, 0, 40), (40, 40, 40), (0, 0, 0),
                       I.uniform_box(29, 12, (0, 0, 0), (40, 40, 40)), None,
                       I.uniform_box(30, 5000, (0, 0, 0), (40, 40, 40))),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_find_matches_reference(oracle_mod, case):
    O = oracle_mod
    _ref_or_skip(O)
    kind, geom, grid, per, xyz, r, q = CASES[case]()
    n = len(xyz)
    ids = np.arange(n, dtype=np.int32) * 3 + 7
    print("find case %s n=%d queries=%d" % (case, n, len(q)), flush=True)
    want = O.find_reference(O.make_spec(kind, geom, grid, periodic=per), ids, xyz, r, q)
    cls = voro_amd.container_poly if kind == 1 else voro_amd.container
    con = cls(*geom, *grid, *[bool(p) for p in per], 8)
    con.put_array(ids, xyz, r)
    got = con.find_voronoi_cells(q)
    assert np.array_equal(got[0], want[0])
    f = want[0]
    assert np.array_equal(got[1][f], want[1][f])
    assert np.array_equal(got[2][f], want[2][f])
    # after a compute the same answers (the binned arrays are reused)
    con.compute_all_cells()
    again = con.find_voronoi_cells(q)
    assert np.array_equal(again[1][f], want[1][f]) and np.array_equal(again[2][f], want[2][f])


@pytest.mark.gpu
def test_gpu_find_single_point_and_periodic_class():
    con = voro_amd.container(0, 1, 0, 1, 0, 1, 2, 2, 2, False, False, False, 8)
    con.put_array(np.array([5, 6], np.int32), np.array([[0.25, 0.25, 0.25], [0.75, 0.75, 0.75]]))
    assert con.find_voronoi_cell(0.3, 0.2, 0.1) == (True, (0.25, 0.25, 0.25), 5)
    assert con.find_voronoi_cell(1.5, 0.2, 0.1)[0] is False
    # one particle in a unit cubic lattice: every point finds it, at the
    # lattice image nearest the point
    pc = voro_amd.container_periodic(1, 0, 1, 0, 0, 1, 2, 2, 2, 8)
    pc.put_array(np.array([1], np.int32), np.array([[0.5, 0.5, 0.5]]))
    assert pc.find_voronoi_cell(0.1, 0.2, 0.3) == (True, (0.5, 0.5, 0.5), 1)
    assert pc.find_voronoi_cell(3.2, -1.9, 0.6) == (True, (3.5, -1.5, 0.5), 1)


# ---- the periodic classes' find_voronoi_cell and compute_ghost_cell ---------

def _para(geom, n, seed, lo, hi):
    """n points at lattice coordinates uniform in [lo, hi)^3."""
    bx, bxy, by, bxz, byz, bz = geom
    u = np.stack([I.uniform(seed, 90 + a, n) for a in range(3)], 1) * (hi - lo) + lo
    return np.stack([u[:, 0] * bx + u[:, 1] * bxy + u[:, 2] * bxz, u[:, 1] * by + u[:, 2] * byz, u[:, 2] * bz], 1)


def test_reference_periodic_find_matches_brute_force(oracle_mod):
    """The reference's periodic find_voronoi_cell (the oracle the GPU tests
    use) is the nearest particle image over the lattice: same id, and the
    returned position is that image (to rounding: the reference adds the
    lattice shifts in its own order)."""
    O = oracle_mod
    _ref_or_skip(O)
    geom = (10, 1.5, 10, 0.7, -1.2, 10)
    bx, bxy, by, bxz, byz, bz = geom
    n = 300
    xyz = _para(geom, n, 51, 0.0, 1.0)
    r = I.radii_lognormal(51, n) * 0.5
    q = _para(geom, 300, 52, -2.0, 3.0)
    a = np.array([[bx, 0, 0], [bxy, by, 0], [bxz, byz, bz]])
    sh = np.array([(i, j, k) for i in range(-4, 5) for j in range(-4, 5) for k in range(-4, 5)], float) @ a
    img = (xyz[None, :, :] + sh[:, None, :]).reshape(-1, 3)
    for kind, rr in ((2, None), (3, r)):
        found, pid, pos = O.prd_queries_reference(O.make_spec(kind, geom, (4, 4, 4), init_mem=64),
                                                  np.arange(n, dtype=np.int32), xyz, rr, q)
        assert found.all()
        d = ((q[:, None, :] - img[None, :, :]) ** 2).sum(-1)
        if rr is not None:
            d = d - np.tile(rr, len(sh))[None, :] ** 2
        best = d.argmin(axis=1)
        assert np.array_equal(pid, best % n)
        assert np.allclose(pos, img[best], rtol=0, atol=1e-9)


def test_reference_periodic_ghost_equals_appended_particle(oracle_mod):
    """After compute_all_cells() the reference's periodic ghost cell of a
    point deep inside a large lattice is the cell the point gets when put as
    the last particle of a fresh container."""
    O = oracle_mod
    _ref_or_skip(O)
    geom = (30, 4.5, 30, 2.1, -3.6, 30)
    n = 5000
    xyz = _para(geom, n, 53, 0.0, 1.0)
    q = _para(geom, 6, 54, 0.3, 0.7)
    sp = O.make_spec(2, geom, (8, 8, 8), init_mem=64)
    g = O.prd_queries_reference(sp, np.arange(n, dtype=np.int32), xyz, None, q, pre=True, ghost=True)
    assert g["exists"].all()
    for i in range(len(q)):
        a = O.run("reference", sp, np.arange(n + 1, dtype=np.int32), np.vstack([xyz, q[i:i + 1]]))
        m = int(np.nonzero(a["id"] == n)[0][0])
        assert a["vol"][m] == g["vol"][i] and a["nfaces"][m] == g["nfaces"][i]
        assert np.array_equal(a["nbr"][a["nbr_off"][m]:a["nbr_off"][m + 1]], g["nbr"][i, :g["nfaces"][i]])


PRD_FIND = {
    # kind, lattice, grid, particles, radical
    "cubic": (2, (10, 0, 10, 0, 0, 10), (5, 5, 5), 1500, False),
    "triclinic": (2, (10, 1.5, 10, 0.7, -1.2, 10), (5, 5, 5), 2000, False),
    "triclinic_poly": (3, (10, 1.5, 10, 0.7, -1.2, 10), (5, 5, 5), 2000, True),
    "skew_poly": (3, (5, 7.5, 5, -6.2, 8.1, 5), (4, 4, 4), 600, True),
}


@pytest.mark.gpu
@pytest.mark.parametrize("pre", [False, True], ids=["fresh", "after_compute"])
@pytest.mark.parametrize("case", sorted(PRD_FIND))
def test_gpu_periodic_find_matches_reference(oracle_mod, case, pre):
    """Queries over several periods, as one call each on the reference
    (which creates the image blocks they reach), on a fresh container and
    after compute_all_cells(): found, id and returned position bit for bit."""
    O = oracle_mod
    _ref_or_skip(O)
    kind, geom, grid, n, poly = PRD_FIND[case]
    xyz = _para(geom, n, 61, 0.0, 1.0)
    r = I.radii_lognormal(61, n) * 0.5 if poly else None
    q = _para(geom, 6000, 62, -2.5, 3.5)
    ids = np.arange(n, dtype=np.int32) * 2 + 11
    print("periodic find case %s pre=%s n=%d queries=%d" % (case, pre, n, len(q)), flush=True)
    want = O.prd_queries_reference(O.make_spec(kind, geom, grid, init_mem=64), ids, xyz, r, q, pre=pre)
    cls = voro_amd.container_periodic_poly if poly else voro_amd.container_periodic
    con = cls(*geom, *grid, 8)
    con.put_array(ids, xyz, r)
    if pre:
        con.compute_all_cells()
    got = con.find_voronoi_cells(q)
    assert want[0].all() and got[0].all()
    assert np.array_equal(got[1], want[1])
    assert np.array_equal(got[2], want[2])


@pytest.mark.gpu
@pytest.mark.parametrize("pre", [False, True], ids=["fresh", "after_compute"])
def test_gpu_periodic_find_lattice_ties(oracle_mod, pre):
    """A simple cubic lattice queried at its cell corners, edge and face
    centres: every answer is a tie between 8, 4 or 2 particles, which the
    reference breaks by scan order, i.e. by the order in which the image
    blocks were created; the engine's answers equal the reference's."""
    O = oracle_mod
    _ref_or_skip(O)
    m = 6
    g = (np.arange(m) + 0.5)
    xyz = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    ids = np.arange(len(xyz), dtype=np.int32)
    h = np.arange(-2 * m, 2 * m + 1) * 0.5
    q = np.stack(np.meshgrid(h, h, h, indexing="ij"), -1).reshape(-1, 3)
    q = q[(np.abs(q * 2) % 2 == 0).sum(1) >= 1]          # at least one integer coordinate: a tie
    geom = (m, 0, m, 0, 0, m)
    want = O.prd_queries_reference(O.make_spec(2, geom, (3, 3, 3), init_mem=64), ids, xyz, None, q, pre=pre)
    con = voro_amd.container_periodic(*geom, 3, 3, 3, 8)
    con.put_array(ids, xyz)
    if pre:
        con.compute_all_cells()
    got = con.find_voronoi_cells(q)
    assert want[0].all() and got[0].all()
    assert np.array_equal(got[1], want[1])
    assert np.array_equal(got[2], want[2])


@pytest.mark.gpu
@pytest.mark.parametrize("case", ["triclinic", "triclinic_poly"])
def test_gpu_periodic_ghost_cells_match_reference(oracle_mod, case):
    """Ghost cells after compute_all_cells(): a point is computed against the
    image blocks the compute created, bit for bit as the reference's; points
    whose search would need another image block (one the reference would
    create during the call, copying the point into it) come back with
    exists = -1 and every other point's results complete."""
    O = oracle_mod
    _ref_or_skip(O)
    kind, geom, grid, n, poly = PRD_FIND[case]
    geom = tuple(3 * t for t in geom)
    grid = (12, 12, 12)
    n = 27 * n
    xyz = _para(geom, n, 71, 0.0, 1.0)
    r = I.radii_lognormal(71, n) * 0.5 if poly else None
    q = _para(geom, 2000, 72, -1.0, 2.0)
    qr = I.radii_lognormal(72, 2000) * 0.5 if poly else None
    ids = np.arange(n, dtype=np.int32) * 3 + 1
    print("periodic ghost case %s n=%d points=%d" % (case, n, len(q)), flush=True)
    want = O.prd_queries_reference(O.make_spec(kind, geom, grid, init_mem=64), ids, xyz, r, q, qr, pre=True,
                                   ghost=True)
    cls = voro_amd.container_periodic_poly if poly else voro_amd.container_periodic
    con = cls(*geom, *grid, 8)
    con.put_array(ids, xyz, r)
    con.compute_all_cells()
    got = con.compute_ghost_cells(q, qr)
    ok = got["exists"] >= 0
    print("ghosts computed %d of %d" % (ok.sum(), len(q)), flush=True)
    assert ok.mean() > 0.9
    assert np.array_equal(got["exists"][ok], want["exists"][ok])
    e = ok & (want["exists"] == 1)
    for k in ("vol", "cent", "nfaces"):
        assert np.array_equal(got[k][e], want[k][e]), k
    for i in np.nonzero(e)[0]:
        nf = want["nfaces"][i]
        assert np.array_equal(got["nbr"][i, :nf], want["nbr"][i, :nf]), i


# ---- compute_ghost_cell ----------------------------------------------------

def test_reference_ghost_equals_appended_particle(oracle_mod):
    """The reference's ghost cell is the cell the point gets when put as the
    last particle (what src/container.hh:488-498 does in place)."""
    O = oracle_mod
    _ref_or_skip(O)
    n = 3000
    xyz = I.uniform_box(31, n, (0, 0, 0), (12, 12, 12))
    q = I.uniform_box(32, 8, (0, 0, 0), (12, 12, 12))
    sp = O.make_spec(0, (0, 12, 0, 12, 0, 12), (6, 6, 6))
    g = O.ghost_reference(sp, np.arange(n, dtype=np.int32), xyz, None, q)
    for i in range(len(q)):
        a = O.run("reference", sp, np.arange(n + 1, dtype=np.int32), np.vstack([xyz, q[i:i + 1]]))
        m = int(np.nonzero(a["id"] == n)[0][0])
        assert a["vol"][m] == g["vol"][i] and a["nfaces"][m] == g["nfaces"][i]
        assert np.array_equal(a["nbr"][a["nbr_off"][m]:a["nbr_off"][m + 1]], g["nbr"][i, :g["nfaces"][i]])


GHOSTS = {
    "plain": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), (0, 0, 0),
                      I.uniform_box(41, 6000, (0, 0, 0), (20, 20, 20)), None,
                      I.uniform_box(42, 3000, (-1, -1, -1), (21, 21, 21)), None),
    "poly": lambda: (1, (0, 20, 0, 20, 0, 20), (10, 10, 10), (0, 0, 0),
                     I.uniform_box(43, 6000, (0, 0, 0), (20, 20, 20)), I.radii_lognormal(43, 6000),
                     I.uniform_box(44, 3000, (0, 0, 0), (20, 20, 20)), I.radii_lognormal(44, 3000) * 1.5),
    "periodic": lambda: (0, (0, 20, 0, 20, 0, 20), (10, 10, 10), (1, 1, 0),
                         I.uniform_box(45, 6000, (0, 0, 0), (20, 20, 20)), None,
                         I.uniform_box(46, 3000, (-30, -30, 0), (50, 50, 20)), None),
}


@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(GHOSTS))
def test_gpu_ghost_cells_match_reference(oracle_mod, case):
    O = oracle_mod
    _ref_or_skip(O)
    kind, geom, grid, per, xyz, r, q, qr = GHOSTS[case]()
    ids = np.arange(len(xyz), dtype=np.int32) * 5 + 2
    print("ghost case %s n=%d points=%d" % (case, len(xyz), len(q)), flush=True)
    want = O.ghost_reference(O.make_spec(kind, geom, grid, periodic=per), ids, xyz, r, q, qr)
    cls = voro_amd.container_poly if kind == 1 else voro_amd.container
    con = cls(*geom, *grid, *[bool(p) for p in per], 8)
    con.put_array(ids, xyz, r)
    got = con.compute_ghost_cells(q, qr)
    assert np.array_equal(got["exists"], want["exists"])
    e = want["exists"] == 1
    for k in ("vol", "cent", "nfaces"):
        assert np.array_equal(got[k][e], want[k][e]), k
    for i in np.nonzero(e)[0]:
        nf = want["nfaces"][i]
        assert np.array_equal(got["nbr"][i, :nf], want["nbr"][i, :nf]), i


@pytest.mark.gpu
def test_gpu_ghost_cells_through_the_huge_tier(oracle_mod):
    """Ghost points with radii far above the particles': cells of over 1,024
    vertices, which the wide cell hands on to the huge one, equal the
    reference's; points outside the box come back with zeroed outputs."""
    O = oracle_mod
    _ref_or_skip(O)
    n = 40000
    xyz = I.uniform_box(131, n, (0, 0, 0), (34, 34, 34))
    r = np.full(n, 0.3)
    ids = np.arange(n, dtype=np.int32)
    q = np.array([[17.0, 17.0, 17.0], [10.2, 20.5, 15.1], [-3.0, 5.0, 5.0], [12.0, 12.0, 23.0]])
    qr = np.array([8.5, 7.0, 1.0, 0.3])
    spec = O.make_spec(1, (0, 34, 0, 34, 0, 34), (20, 20, 20))
    want = O.ghost_reference(spec, ids, xyz, r, q, qr, max_nbr=4096)
    assert want["nfaces"].max() > 520          # over 1,024 vertices
    con = voro_amd.container_poly(0, 34, 0, 34, 0, 34, 20, 20, 20, False, False, False, 8)
    con.put_array(ids, xyz, r)
    got = con.compute_ghost_cells(q, qr, max_nbr=4096)
    assert np.array_equal(got["exists"], want["exists"])
    assert got["exists"][2] == 0 and got["vol"][2] == 0 and np.all(got["cent"][2] == 0)
    e = want["exists"] == 1
    for k in ("vol", "cent", "nfaces"):
        assert np.array_equal(got[k][e], want[k][e]), k
    for i in np.nonzero(e)[0]:
        nf = want["nfaces"][i]
        assert np.array_equal(got["nbr"][i, :nf], want["nbr"][i, :nf]), i


@pytest.mark.gpu
def test_gpu_find_on_empty_container_and_non_finite_queries(oracle_mod):
    """No particle: every query is not found (src/container.cc:260-261), at
    once, on a 100^3-block grid; NaN and infinite query points are not found
    either."""
    O = oracle_mod
    _ref_or_skip(O)
    con = voro_amd.container(0, 10, 0, 10, 0, 10, 100, 100, 100, False, False, False, 8)
    f, pid, pos = con.find_voronoi_cells(np.array([[1.0, 2.0, 3.0], [np.nan, 1, 1], [5, 5, 5]]))
    assert not f.any()
    xyz = I.uniform_box(132, 5000, (0, 0, 0), (10, 10, 10))
    ids = np.arange(5000, dtype=np.int32)
    con2 = voro_amd.container(0, 10, 0, 10, 0, 10, 10, 10, 10, False, True, False, 8)
    con2.put_array(ids, xyz)
    q = np.array([[np.nan, 1, 1], [1, np.inf, 1], [2, 2, -np.inf], [4.5, 4.5, 4.5], [1, 1e300, 1]])
    f2, pid2, pos2 = con2.find_voronoi_cells(q)
    assert f2.tolist() == [False, False, False, True, False]
    want = O.find_reference(O.make_spec(0, (0, 10, 0, 10, 0, 10), (10, 10, 10), periodic=(0, 1, 0)), ids, xyz, None,
                            q[3:4])
    assert want[0][0] and pid2[3] == want[1][0] and np.array_equal(pos2[3], want[2][0])


# ---- print_custom and import ---------------------------------------------

FORMAT = "%i %q %v %s %n %c %C 100%% %z"


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [0, 1])
def test_gpu_print_custom_matches_reference_text(oracle_mod, tmp_path, kind):
    """print_custom's text, byte for byte, with the particles read through
    import_() (which must give the reference's particles exactly)."""
    O = oracle_mod
    _ref_or_skip(O)
    n = 2000
    xyz = I.uniform_box(51 + kind, n, (0, 0, 0), (10, 10, 10))
    r = I.radii_lognormal(51, n) * 0.5 if kind else None
    ids = np.arange(n, dtype=np.int32) + 11
    src = tmp_path / "pack.txt"
    with open(src, "w") as f:
        for t in range(n):
            f.write("%d %.17g %.17g %.17g%s\n" % (ids[t], *xyz[t], (" %.17g" % r[t]) if kind else ""))
    want = tmp_path / "ref.txt"
    fmt = FORMAT + " %r"
    assert O.print_custom_reference(O.make_spec(kind, (0, 10, 0, 10, 0, 10), (5, 5, 5)), ids, xyz, r, fmt, str(want))
    cls = voro_amd.container_poly if kind else voro_amd.container
    con = cls(0, 10, 0, 10, 0, 10, 5, 5, 5, False, False, False, 8)
    con.import_(str(src))
    assert con.print_custom(fmt) == want.read_text()


@pytest.mark.gpu
def test_gpu_import_stops_at_a_bad_record(tmp_path):
    """A record that does not parse is an import error; the records before
    it were put (the reference exits at that point)."""
    src = tmp_path / "bad.txt"
    src.write_text("1 0.5 0.5 0.5\n2 1.5 1.5 1.5\n3 x 0 0\n")
    con = voro_amd.container(0, 4, 0, 4, 0, 4, 2, 2, 2, False, False, False, 8)
    with pytest.raises(voro_amd.VoroGPUError) as e:
        con.import_(str(src))
    assert e.value.status == voro_amd.FILE_ERROR
    con.compute_all_cells()
    assert sorted(con.results()["id"].tolist()) == [1, 2]


# ---- particle_order / c_loop_order, pre_container --------------------------

def test_pre_container_guess_optimal():
    pc = voro_amd.pre_container(0, 10, 0, 20, 0, 5, False, False, True)
    for t, p in enumerate(I.uniform_box(61, 1000, (0, 0, -3), (10, 20, 8))):
        pc.put(t, *p)
    # z periodic: every particle kept; the grid of src/pre_container.cc:57-63
    assert pc.total_particles() == 1000
    s = (1000 / (5.6 * 10 * 20 * 5)) ** (1 / 3.0)
    assert pc.guess_optimal() == (int(10 * s + 1), int(20 * s + 1), int(5 * s + 1))
    pc2 = voro_amd.pre_container(0, 1, 0, 1, 0, 1, False, False, False)
    pc2.put(0, 0.5, 0.5, 1.5)
    assert pc2.total_particles() == 0


@pytest.mark.gpu
def test_gpu_loop_order_and_pre_container():
    xyz = I.uniform_box(62, 4000, (-1, 0, 0), (11, 10, 10))     # some outside in x: dropped
    ids = np.arange(4000, dtype=np.int32) + 100
    con = voro_amd.container(0, 10, 0, 10, 0, 10, 5, 5, 5, False, False, False, 8)
    po = voro_amd.particle_order()
    con.put_array(ids[:1000], xyz[:1000])
    con.put_array(ids[1000:], xyz[1000:], order=po)
    con.compute_all_cells()
    vl = voro_amd.c_loop_order(con, po)
    seen = []
    if vl.start():
        while True:
            seen.append(vl.pid())
            if not vl.inc():
                break
    inside = (xyz[:, 0] >= 0) & (xyz[:, 0] < 10)
    assert seen == [int(i) for i, k in zip(ids[1000:], inside[1000:]) if k]
    # pre_container -> container: the same cells as putting directly
    pc = voro_amd.pre_container(0, 10, 0, 10, 0, 10, False, False, False)
    for i, p in zip(ids, xyz):
        pc.put(int(i), *p)
    nx, ny, nz = pc.guess_optimal()
    a = voro_amd.container(0, 10, 0, 10, 0, 10, nx, ny, nz, False, False, False, 8)
    pc.setup(a)
    a.compute_all_cells()
    b = voro_amd.container(0, 10, 0, 10, 0, 10, nx, ny, nz, False, False, False, 8)
    b.put_array(ids[inside], xyz[inside])
    b.compute_all_cells()
    ra, rb = a.results(), b.results()
    for k in ("id", "vol", "nbr"):
        assert np.array_equal(ra[k], rb[k]), k
