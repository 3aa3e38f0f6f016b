"""Seeded synthetic particle sets for the benchmark workloads and the tests.

Every draw is a counter-based splitmix64 hash of (seed, stream, index), so a
given (seed, stream) gives bit-identical FP64 inputs in any process, on any
host; ``double = (u >> 11) * 2**-53``.

Workloads (box, grid, particle count):
  * ``unit_cube``      -- [0,1]^3, 6x6x6 blocks, 1,000 uniform particles.
  * ``uniform``        -- [0,100]^3, 58^3 blocks (~5 particles per block),
                          1,000,000 uniform particles.
  * ``clusters``       -- same box and grid; 64 isotropic Gaussian clusters
                          (sigma ~ U[2,8], Dirichlet(1) weights) plus a 30%
                          uniform background; points outside the box redrawn.
  * ``periodic_poly``  -- triclinic periodic box bx=by=bz=60, bxy=5, bxz=3,
                          byz=4, 74^3 blocks, 2,000,000 particles, radii
                          clip(lognormal(ln 0.5, 0.25), 0.2, 1.2).
  * ``fcc_sphere``     -- one perturbed FCC lattice cut by a sphere wall.
"""
import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_SEEDMUL = np.uint64(0xD1B54A32D192ED03)


def _mix(z):
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def u64(seed, stream, n, start=0):
    """n 64-bit words for (seed, stream), counter start..start+n-1."""
    with np.errstate(over="ignore"):
        base = _mix(np.uint64(seed) * _SEEDMUL + np.uint64(stream) + _GOLDEN)
        ctr = np.arange(start, start + n, dtype=np.uint64)
        return _mix(base + ctr * _GOLDEN)


def uniform(seed, stream, n, start=0):
    """n doubles in [0,1): (u >> 11) * 2^-53."""
    return (u64(seed, stream, n, start) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed, stream, n):
    """Standard normals by the Box-Muller transform of two uniform streams."""
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 1, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def uniform_box(seed, n, lo, hi):
    """n uniform points in the axis-aligned box [lo,hi)^3, shape (n,3)."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    xyz = np.empty((n, 3), dtype=np.float64)
    for a in range(3):
        xyz[:, a] = lo[a] + (hi[a] - lo[a]) * uniform(seed, 10 + a, n)
    return xyz


def clusters(seed, n, box=100.0, n_clusters=64, background=0.3):
    """Gaussian-cluster mixture plus uniform background, all inside [0,box)^3."""
    centres = uniform_box(seed * 7919 + 1, n_clusters, (0, 0, 0), (box, box, box))
    sig = 2.0 + 6.0 * uniform(seed, 40, n_clusters)
    w = -np.log1p(-uniform(seed, 41, n_clusters))        # Dirichlet(1) = normalised Exp(1)
    w = w / w.sum()
    cdf = np.cumsum(w)
    out = np.empty((n, 3), dtype=np.float64)
    filled = 0
    rnd = 0
    while filled < n:
        m = n - filled
        s = 1000 + 16 * rnd
        is_bg = uniform(seed, s, m) < background
        which = np.minimum(np.searchsorted(cdf, uniform(seed, s + 1, m), side="right"), n_clusters - 1)
        p = np.empty((m, 3))
        ub = uniform_box(seed * 131 + rnd + 3, m, (0, 0, 0), (box, box, box))
        for a in range(3):
            g = normal(seed, s + 2 + 2 * a, m)
            p[:, a] = np.where(is_bg, ub[:, a], centres[which, a] + sig[which] * g)
        ok = np.all((p >= 0.0) & (p < box), axis=1)
        take = p[ok]
        out[filled:filled + len(take)] = take
        filled += len(take)
        rnd += 1
    return out


def radii_lognormal(seed, n, mu=np.log(0.5), sigma=0.25, lo=0.2, hi=1.2):
    return np.clip(np.exp(mu + sigma * normal(seed, 60, n)), lo, hi)


def fcc_sphere(seed, n_target, spacing=1.0, jitter=0.05):
    """Perturbed FCC lattice with about n_target points inside a sphere of
    radius 0.45 L centred in the cube [0,L)^3.  Returns (xyz, L)."""
    # An FCC cell holds 4 points; the sphere keeps ~0.38 of the cube.
    m = max(2, int(round((n_target / (4 * 0.3817)) ** (1.0 / 3.0))))
    L = m * spacing
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) + 0.25
    ii, jj, kk = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
    cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float64)
    pts = (cells[:, None, :] + base[None, :, :]).reshape(-1, 3) * spacing
    k = len(pts)
    for a in range(3):
        pts[:, a] += jitter * spacing * normal(seed, 70 + 2 * a, k)
    c = 0.5 * L
    r = np.sqrt(((pts - c) ** 2).sum(axis=1))
    keep = (r < 0.45 * L) & np.all((pts >= 0) & (pts < L), axis=1)
    return pts[keep], L


def fcc_batch(seed, count, n_lo=100, n_hi=20000, spacing=1.0):
    """The batched workload: `count` independent FCC-jitter containers.
    Container c targets N_c = round(exp(U[ln n_lo, ln n_hi])) particles: an
    FCC lattice of spacing a with m = round(cbrt(N_c / (4 * 0.3817))) cells a
    side (L = m a; the wall sphere keeps ~0.3817 of the cube), Gaussian
    jitter of sigma ~ U[0.02, 0.2] a, kept inside the wall sphere of radius
    0.45 L at the box centre; grid round(cbrt(N/5))^3 for the N kept.
    Returns a list of (L, grid, xyz)."""
    lnN = np.log(n_lo) + (np.log(n_hi) - np.log(n_lo)) * uniform(seed, 200, count)
    sig = 0.02 + 0.18 * uniform(seed, 201, count)
    out = []
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]]) + 0.25
    for c in range(count):
        nt = int(round(np.exp(lnN[c])))
        m = max(2, int(round((nt / (4 * 0.3817)) ** (1.0 / 3.0))))
        L = m * spacing
        ii, jj, kk = np.meshgrid(np.arange(m), np.arange(m), np.arange(m), indexing="ij")
        cells = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1).astype(np.float64)
        pts = (cells[:, None, :] + base[None, :, :]).reshape(-1, 3) * spacing
        k = len(pts)
        for a in range(3):
            pts[:, a] += sig[c] * spacing * normal(seed * 100003 + c, 300 + 2 * a, k)
        cen = 0.5 * L
        rr = np.sqrt(((pts - cen) ** 2).sum(axis=1))
        keep = (rr < 0.45 * L) & np.all((pts >= 0) & (pts < L), axis=1)
        pts = pts[keep]
        g = max(1, int(round((len(pts) / 5.0) ** (1.0 / 3.0))))
        out.append((L, g, pts))
    return out


WORKLOADS = {
    "unit_cube": dict(kind=0, box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), grid=(6, 6, 6), n=1000),
    "uniform": dict(kind=0, box=(0.0, 100.0, 0.0, 100.0, 0.0, 100.0), grid=(58, 58, 58), n=1_000_000),
    "clusters": dict(kind=0, box=(0.0, 100.0, 0.0, 100.0, 0.0, 100.0), grid=(58, 58, 58), n=1_000_000),
    "periodic_poly": dict(kind=3, box=(60.0, 5.0, 60.0, 3.0, 4.0, 60.0), grid=(74, 74, 74), n=2_000_000),
}


def make(workload, seed, n=None):
    """Returns (ids, xyz, r) for a named workload; r is None for plain ones."""
    w = WORKLOADS[workload]
    n = w["n"] if n is None else n
    ids = np.arange(n, dtype=np.int32)
    b = w["box"]
    if workload == "clusters":
        return ids, clusters(seed, n, box=b[1]), None
    if workload == "periodic_poly":
        xyz = uniform_box(seed, n, (0, 0, 0), (b[0], b[2], b[5]))
        return ids, xyz, radii_lognormal(seed, n)
    xyz = uniform_box(seed, n, (b[0], b[2], b[4]), (b[1], b[3], b[5]))
    return ids, xyz, None
