"""Runs the HIP engine through its C ABI and returns oracle-shaped results."""
import numpy as np

import voro_amd


def engine_run(kind, geom, grid, ids, xyz, r=None, periodic=(0, 0, 0), spheres=None, flags=None, stats=False,
               first_tier=0):
    if kind >= 2:
        cls = voro_amd.container_periodic_poly if kind == 3 else voro_amd.container_periodic
        con = cls(*geom, *grid, 8)
    else:
        cls = voro_amd.container_poly if kind == 1 else voro_amd.container
        con = cls(*geom, *grid, *[bool(p) for p in periodic], 8)
    for s in spheres or []:
        con.add_wall(voro_amd.wall_sphere(*s))
    con.put_array(ids, xyz, r)
    if first_tier:
        con.set_first_tier(first_tier)
    f = voro_amd.F_NEIGHBORS | voro_amd.F_CSR if flags is None else flags
    if stats:
        f |= voro_amd.F_STATS
    con.compute_all_cells(f)
    out = dict(con.results())
    if stats:
        out.update({"st_" + k: v for k, v in con.stats().items()})
    out["_con"] = con
    return out
