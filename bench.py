"""Benchmark: cells/s of the per-particle cell computation on one MI355X.

Workload: 1,000,000 uniform FP64 particles (seed 42) in the non-periodic cube
[0,100]^3 with 58^3 blocks (~5 per block), voronoicell_neighbor outputs
(volume, centroid, face count, neighbour ids in loop-order CSR).  A step is
one pass of the hot path over the particles already resident in HBM:
stable block binning + every cell + output compaction.

Two paths are timed on the same particles: the plain one (expanding
worklist, as the reference) and the learned one (candidate lists cut at the
depth the predictor gives, cells whose list is too short recomputed by the
plain path; results are checked identical).  The predictor is first trained
on the GPU on the plain path's statistics for other draws of the workload.
`value` is the learned path's throughput, the plain path's is under "plain".

Prints one JSON line (see the keys below).  The CPU baseline is the
reference's own container::compute_all_cells() (built from the reference
sources into oracle/_ref), timed on this host, one thread, on the same
particles.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Per-cell algorithmic work of the cell kernel ("Measurement"):
FLOP_PER_CELL = 7.6e3          # FP64 operations per cell (uniform workload)
BYTES_PER_CELL_FIXED = 36 + 44  # own particle record in + fixed outputs out
MI355X_HBM_GBS = 8000.0
MI355X_FP64_TFLOPS = 78.6


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample", type=int, default=1_000_000,
                    help="particles of the workload timed on the CPU")
    ap.add_argument("--occ", type=int, default=0, help="fast-tier waves per SIMD (2,3,4; 0 = default)")
    ap.add_argument("--waves", type=int, default=0, help="fast-tier persistent wavefronts (0 = default)")
    ap.add_argument("--train-seeds", type=int, nargs="*", default=[1, 2],
                    help="draws of the workload whose plain-path statistics train the predictor")
    ap.add_argument("--epochs", type=int, default=400, help="predictor training steps (full-batch Adam)")
    ap.add_argument("--fallback", type=float, default=0.02,
                    help="fraction of training blocks the depth margin allows to be under-predicted")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (scale for the FETCH_SIZE and "
                         "WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (--gpus 1); multi-GPU runs are not part of it",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        a = torch.ones(n, dtype=torch.uint8, device="cuda:0")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n}))
        return

    import voro_amd
    from voro_amd import inputs as I

    ids, xyz, _ = I.make("uniform", 42, n=args.n)
    dev = torch.device("cuda:0")
    d_ids = torch.from_numpy(ids).to(dev)
    d_xyz = torch.from_numpy(xyz).to(dev)
    torch.cuda.synchronize()
    G = max(1, int(round(58 * (args.n / 1e6) ** (1 / 3))))

    def make_container(di, dx):
        c = voro_amd.container(0, 100, 0, 100, 0, 100, G, G, G, False, False, False, 8)
        c.put_device(len(di), di.data_ptr(), dx.data_ptr())
        if args.occ or args.waves:
            c.set_launch(args.waves, args.occ)
        return c

    flags = voro_amd.F_NEIGHBORS | voro_amd.F_CSR

    # The predictor is trained here, on the GPU, on the statistics the plain
    # path records for two other draws of the same workload (not timed).
    t_train = time.perf_counter()
    pred = voro_amd.predictor()
    for seed in args.train_seeds:
        tids, txyz, _ = I.make("uniform", seed, n=args.n)
        tc = make_container(torch.from_numpy(tids).to(dev), torch.from_numpy(txyz).to(dev))
        tc.compute_all_cells(flags | voro_amd.F_STATS)
        pred.add_samples(tc)
        del tc
    train_loss = pred.train(epochs=args.epochs, lr=3e-3, seed=1, fallback=args.fallback)
    t_train = time.perf_counter() - t_train

    con = make_container(d_ids, d_xyz)

    def timed(fl):
        for _ in range(args.warmup):
            con.compute_all_cells(fl)
        torch.cuda.synchronize()
        tier_ms, gather_ms = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            con.compute_all_cells(fl)
            tier_ms.append(con.timings()["fast_ms"])
            if fl & voro_amd.F_LEARNED:
                gather_ms.append(con.learned_stats()["gather_ms"])
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        return dt, float(np.mean(tier_ms)), float(np.mean(gather_ms)) if gather_ms else 0.0

    dt_plain, plain_tier_ms, _ = timed(flags)
    r = con.results()
    ncells = len(r["id"])
    nnbr = len(r["nbr"])
    plain_cells_s = ncells * args.steps / dt_plain
    plain_ref = (r["vol"].copy(), r["nbr"].copy())
    con.set_predictor(pred)
    dt, tier_ms, gather_ms = timed(flags | voro_amd.F_LEARNED)
    ls = con.learned_stats()
    r = con.results()
    identical = bool(np.array_equal(r["vol"], plain_ref[0]) and np.array_equal(r["nbr"], plain_ref[1]))
    cells_s = ncells * args.steps / dt
    # roofline of the list-driven cell kernel alone (its HIP events, gathering excluded)
    kern_s = (tier_ms - gather_ms) / 1e3
    bytes_launch = ncells * BYTES_PER_CELL_FIXED + 4.0 * nnbr
    hbm_gbs = bytes_launch / kern_s / 1e9
    fp64_tf = ncells * FLOP_PER_CELL / kern_s / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_k_cells_uniform1m.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        if O.reference() is not None:
            m = min(args.cpu_sample, args.n)
            sp = O.make_spec(0, (0, 100, 0, 100, 0, 100), (G, G, G))
            if m < args.n:
                # the first m particles of the same draw are a uniform subset;
                # scale the grid to keep ~5 particles per block
                g = max(1, int(round(G * (m / args.n) ** (1 / 3))))
                sp = O.make_spec(0, (0, 100, 0, 100, 0, 100), (g, g, g))
            secs = O.time_reference(sp, ids[:m], xyz[:m], loop="compute_all_cells")
            cpu = {"value": m / secs, "unit": "cells/s", "cores": 1, "kind": "reference",
                   "sample": "container::compute_all_cells() on %d of the uniform particles, %.1f s" % (m, secs)}

    line = {
        "metric": "cells/s at 1M uniform pts, learned vs plain GPU path (bit-exact nbrs, 1e-12 vol)",
        "value": cells_s,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "uniform 1M particles in [0,100]^3, 58^3 blocks, seed 42, neighbour outputs",
                   "n_particles": int(args.n), "grid": [G, G, G], "path": "learned",
                   "fast_occupancy": args.occ, "fast_waves": args.waves},
        "plain": {"value": plain_cells_s, "ms_per_step": dt_plain / args.steps * 1e3, "tier_ms": plain_tier_ms},
        "learned": {"speedup_vs_plain": cells_s / plain_cells_s, "fallbacks": ls["fallbacks"],
                    "fallback_rate": ls["fallbacks"] / max(ncells, 1), "list_entries": ls["list_entries"],
                    "gather_ms": gather_ms, "tier_ms": tier_ms, "identical_to_plain": identical,
                    "train_seconds": t_train, "train_loss": train_loss, "train_seeds": list(args.train_seeds)},
        "roofline": {"bound": "hbm", "achieved": hbm_gbs, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                     "frac": hbm_gbs / MI355X_HBM_GBS, "traffic": traffic,
                     "kernel": "k_cells_list<mono>", "kernel_ms": kern_s * 1e3},
        "roofline_fp64": {"achieved": fp64_tf, "peak": MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
                          "frac": fp64_tf / MI355X_FP64_TFLOPS},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
