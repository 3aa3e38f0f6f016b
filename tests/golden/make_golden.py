"""Writes the golden fixtures of tests/golden/ from the reference-compiled
oracle (oracle/_ref/libvororef.so, built from the reference sources by
oracle/Makefile).  Run from the repository root:

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  unit_cube_seed{1,2,3}.npz -- 1,000 uniform particles in [0,1]^3, 6^3 blocks:
      inputs (ids, xyz) and the reference outputs in c_loop_all order
      (id, ijk, q, exists, vol, cent, nfaces, nbr_off, nbr).
  poly_small.npz           -- 2,000 particles, container_poly [0,10]^3, 8^3 blocks.
  wall_small.npz           -- perturbed FCC in a sphere wall, ~2,000 particles.
  prd_poly_small.npz       -- container_periodic_poly(60,5,60,3,4,60), 8^3 blocks,
      3,000 particles drawn in [0,60)^3 (remapped by put), lognormal radii.
  prd_tri_small.npz        -- container_periodic(10,1.5,10,0.7,-1.2,10), 5^3 blocks,
      1,000 particles drawn in [-2,12)^3.
  worklist_sha256.txt      -- SHA-256 of the reference worklist table values.
  packs/                   -- the particle packings the reference's examples read
      (examples/basic/pack_ten_cube, examples/custom/pack_six_cube and
      pack_six_cube_poly, examples/walls/pack_cylinder and pack_torus),
      copied byte for byte as input data ("python make_golden.py packs").
"""
import hashlib
import os
import re
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402
from voro_amd import inputs as I  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def save(name, spec_args, ids, xyz, r=None):
    sp = O.make_spec(*spec_args[:3], **spec_args[3])
    out = O.run("reference", sp, ids, xyz, r)
    d = dict(ids=ids, xyz=xyz, **{k: v for k, v in out.items() if k != "pos"})
    if r is not None:
        d["r"] = r
    d["spec_kind"] = np.array(sp.kind)
    d["spec_geom"] = np.array(list(sp.geom))
    d["spec_grid"] = np.array(list(sp.grid))
    np.savez_compressed(os.path.join(HERE, name), **d)
    print("wrote", name, len(out["id"]), "cells")


def main():
    O.build()
    for seed in (1, 2, 3):
        ids, xyz, _ = I.make("unit_cube", seed)
        save("unit_cube_seed%d.npz" % seed, (0, (0, 1, 0, 1, 0, 1), (6, 6, 6), {}), ids, xyz)
    n = 2000
    ids = np.arange(n, dtype=np.int32)
    xyz = I.uniform_box(11, n, (0, 0, 0), (10, 10, 10))
    r = I.radii_lognormal(11, n)
    save("poly_small.npz", (1, (0, 10, 0, 10, 0, 10), (8, 8, 8), {}), ids, xyz, r)
    pts, L = I.fcc_sphere(5, 2000, 1.0, 0.05)
    g = int(round((len(pts) / 5) ** (1 / 3)))
    save("wall_small.npz", (0, (0, L, 0, L, 0, L), (g, g, g),
                            {"spheres": [[L / 2, L / 2, L / 2, 0.45 * L, -99]]}),
         np.arange(len(pts), dtype=np.int32), pts)
    n = 3000
    ids = np.arange(n, dtype=np.int32)
    save("prd_poly_small.npz", (3, (60, 5, 60, 3, 4, 60), (8, 8, 8), {}), ids,
         I.uniform_box(21, n, (0, 0, 0), (60, 60, 60)), I.radii_lognormal(21, n))
    n = 1000
    ids = np.arange(n, dtype=np.int32)
    save("prd_tri_small.npz", (2, (10, 1.5, 10, 0.7, -1.2, 10), (5, 5, 5), {}), ids,
         I.uniform_box(22, n, (-2, -2, -2), (12, 12, 12)))
    src = open(os.path.join(O.REF_ROOT, "src", "v_base_wl.cc")).read()
    body = src[src.index("{") + 1:src.rindex("}")]
    vals = [int(x, 0) for x in re.findall(r"0x[0-9a-f]+|\b\d+\b", body)]
    h = hashlib.sha256(np.array(vals, dtype=np.uint32).tobytes()).hexdigest()
    with open(os.path.join(HERE, "worklist_sha256.txt"), "w") as f:
        f.write(h + "\n")
    print("worklist sha256", h)


PACKS = ["basic/pack_ten_cube", "custom/pack_six_cube", "custom/pack_six_cube_poly", "walls/pack_cylinder",
         "walls/pack_torus"]


def copy_packs():
    os.makedirs(os.path.join(HERE, "packs"), exist_ok=True)
    for rel in PACKS:
        data = open(os.path.join(O.REF_ROOT, "examples", rel), "rb").read()
        with open(os.path.join(HERE, "packs", os.path.basename(rel)), "wb") as f:
            f.write(data)
        print("copied", rel, len(data), "bytes")


if __name__ == "__main__":
    if sys.argv[1:] == ["packs"]:
        copy_packs()
    else:
        main()
