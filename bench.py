"""Benchmark: cells/s of the per-particle cell computation on one MI355X.

Workload: 1,000,000 uniform FP64 particles (seed 42) in the non-periodic cube
[0,100]^3 with 58^3 blocks (~5 per block), voronoicell_neighbor outputs
(volume, centroid, face count, neighbour ids in loop-order CSR).  A step is
one pass of the hot path over the particles already resident in HBM:
stable block binning + every cell + output compaction.

Two paths are timed on the same particles: the plain one (expanding
worklist, as the reference; first tier one cell per wavefront) and the
learned one (candidate lists cut at the depth the predictor gives, cells
whose list is too short recomputed by the plain path; results are checked
identical).  The predictor is first trained on the GPU on the plain path's
statistics for other draws of the workload.  `value` is the learned path's
throughput, the plain path's is under "plain", with the plain path's
one-cell-per-lane first tier timed beside it.

Prints one JSON line (see the keys below).  The CPU baseline is the
reference's own container::compute_all_cells() (built from the reference
sources into oracle/_ref), timed on this host on the same particles, on all
the cores this process may use (z-slabs of blocks through
c_loop_subset::setup_intbox, one container copy per thread; median of three
runs), with the one-thread time and the voronoicell_neighbor output loop
beside it.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Per-cell algorithmic work of the cell kernel on the uniform workload:
# 7.6 kFLOP FP64 per cell (operation counts of the reference algorithm: plane
# tests, new vertices, radius refreshes, region tests, output walks), and
# compulsory HBM bytes: the own particle record and id in (36 B), the fixed
# outputs out (volume 8, centroid 24, face count 4, status 4, neighbour
# start 4 = 44 B), plus 4 B per neighbour id.
FLOP_PER_CELL = 7.6e3
BYTES_PER_CELL_FIXED = 36 + 44
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
    ap.add_argument("--workload", choices=["uniform", "periodic_poly", "batch", "clusters", "find"], default="uniform",
                    help="uniform: the headline measurement (default); periodic_poly / batch: only that "
                         "secondary workload, for profiling")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the secondary workloads (periodic radical 2M, batch of 4096 containers)")
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

    if args.workload != "uniform":
        args.no_cpu = True
        print(json.dumps(secondary_workloads(args, torch, only=args.workload)))
        return

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

    con.set_first_tier(1)
    dt_lane, lane_tier_ms, _ = timed(flags)
    con.set_first_tier(0)
    # the project's autotuner picks the plain path's launch configuration on
    # this container (untimed setup), then the learned path's depth margin
    wl = "uniform %d particles, %d^3 blocks" % (args.n, G)
    tune_plain = con.autotune(flags, reps=2, config=wl)
    dt_plain, plain_tier_ms, _ = timed(flags)
    plain_occ = tune_plain["best"]["occupancy"]
    plain_kernel = ("k_wcells<1, mono, %d>" % plain_occ if tune_plain["best"]["first_tier"] == "wavefront"
                    else "k_cells<Tier3, mono, %d>" % plain_occ)
    r = con.results()
    ncells = len(r["id"])
    nnbr = len(r["nbr"])
    plain_cells_s = ncells * args.steps / dt_plain
    plain_ref = (r["vol"].copy(), r["nbr"].copy())
    con.set_predictor(pred)
    tune_learned = con.autotune(flags | voro_amd.F_LEARNED, reps=2, config=wl + ", learned")
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
    # and of the plain path's wavefront-cell kernel
    wave_s = plain_tier_ms / 1e3
    wave_fp64 = ncells * FLOP_PER_CELL / wave_s / 1e12
    # the calibrated FETCH_SIZE + WRITE_SIZE of each kernel variant, measured
    # with rocprofv3 on this workload (profiles/); the variant the autotuner
    # chose is looked up
    def traffic_of(name, kernel):
        tpath = os.path.join(ROOT, "profiles", name)
        if not os.path.exists(tpath):
            return None
        return json.load(open(tpath)).get("kernels", {}).get(kernel, {}).get("bytes_per_launch")
    lt = tune_learned["best"]
    list_occ = 4 if lt["occupancy"] <= 4 else min(6, lt["occupancy"])
    list_kernel = "k_wcells_list<mono, %d>" % list_occ
    traffic = traffic_of("traffic_k_wcells_list_uniform1m.json", list_kernel)
    wave_traffic = traffic_of("traffic_k_wcells_uniform1m.json", plain_kernel)

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
            cpu = cpu_reference(O, sp, ids[:m], xyz[:m], None, m, "uniform particles")

    others = None if args.headline_only else secondary_workloads(args, torch)

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
        "plain": {"value": plain_cells_s, "ms_per_step": dt_plain / args.steps * 1e3, "tier_ms": plain_tier_ms,
                  "first_tier": plain_kernel, "autotune": tune_plain,
                  "per_lane_first_tier": {"value": ncells * args.steps / dt_lane,
                                          "ms_per_step": dt_lane / args.steps * 1e3, "tier_ms": lane_tier_ms},
                  "roofline_fp64": {"kernel": plain_kernel, "kernel_ms": plain_tier_ms,
                                    "achieved": wave_fp64, "peak": MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
                                    "frac": wave_fp64 / MI355X_FP64_TFLOPS},
                  "roofline_hbm": {"achieved": bytes_launch / wave_s / 1e9, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                                   "frac": bytes_launch / wave_s / 1e9 / MI355X_HBM_GBS, "traffic": wave_traffic}},
        "learned": {"speedup_vs_plain": cells_s / plain_cells_s, "autotune": tune_learned, "fallbacks": ls["fallbacks"],
                    "fallback_rate": ls["fallbacks"] / max(ncells, 1), "list_entries": ls["list_entries"],
                    "gather_ms": gather_ms, "tier_ms": tier_ms, "identical_to_plain": identical,
                    "train_seconds": t_train, "train_loss": train_loss, "train_seeds": list(args.train_seeds)},
        "roofline": {"bound": "mfma", "achieved": fp64_tf, "peak": MI355X_FP64_TFLOPS, "unit": "TFLOP/s",
                     "frac": fp64_tf / MI355X_FP64_TFLOPS, "traffic": traffic,
                     "kernel": list_kernel, "kernel_ms": kern_s * 1e3,
                     "compute": "FP64 vector arithmetic, no matrix instructions; the peak is MI355X's FP64 rate",
                     "limiter": "scalar-instruction issue (see the instruction-mix profile)",
                     "flop_per_cell": FLOP_PER_CELL},
        "roofline_hbm": {"achieved": hbm_gbs, "peak": MI355X_HBM_GBS, "unit": "GB/s",
                         "frac": hbm_gbs / MI355X_HBM_GBS, "traffic": traffic, "bytes_per_launch": bytes_launch},
        "cpu_baseline": cpu,
        "other_workloads": others,
    }
    print(json.dumps(line))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    """The host cores this process may use (its affinity set, at most 16)."""
    from oracle import oracle as O
    return O.default_threads()


def cpu_reference(O, sp, ids, xyz, r, m, what, runs=3):
    """The reference on this host: container::compute_all_cells() on all the
    process's cores (one container copy per thread, each computing a z-slab
    of blocks via c_loop_subset::setup_intbox; median of `runs`), the
    voronoicell_neighbor output loop the same way, and compute_all_cells()
    on one thread once."""
    T = host_threads()
    mt = sorted(O.time_reference_mt(sp, ids, xyz, r, threads=T, loop="compute_all_cells") for _ in range(runs))
    nb = sorted(O.time_reference_mt(sp, ids, xyz, r, threads=T, loop="neighbor") for _ in range(runs))
    one = O.time_reference(sp, ids, xyz, r, loop="compute_all_cells")
    med, nmed = mt[len(mt) // 2], nb[len(nb) // 2]
    return {"value": m / med, "unit": "cells/s", "cores": T, "kind": "reference", "cpu": cpu_model(),
            "sample": "container::compute_all_cells() over %d %s on %d threads (z-slabs of blocks, one container "
                      "copy each), median of %d runs: %.3f s" % (m, what, T, runs, med),
            "neighbor_loop": {"value": m / nmed, "seconds_median": nmed,
                              "what": "voronoicell_neighbor loop with volume, centroid, face count, neighbours"},
            "one_thread": {"value": m / one, "seconds": one}}


def secondary_workloads(args, torch, only=None):
    """The radical periodic workload (container_periodic_poly(60,5,60,3,4,60),
    2,000,000 particles, 74^3 blocks), the batched one (4,096 FCC-jitter
    containers with sphere walls, ~15.5M particles, one graph launch after
    autotuning), the clustered one with the predictor and 1M find_voronoi_cell
    queries: GPU rates (inputs resident on the device), and the reference
    timed on bounded samples on this host."""
    import voro_amd
    out = {}
    flags = voro_amd.F_NEIGHBORS | voro_amd.F_CSR
    specs = None
    if only is None or only == "periodic_poly":
        periodic_poly(out, flags, torch, args.steps if only else 1)
    if only is None or only == "batch":
        specs = batch(out, flags, args.steps if only else 1)
    if only is None or only == "clusters":
        clusters_learned(out, flags, torch)
    if only is None or only == "find":
        find_cells(out, args.steps if only else 3, not args.no_cpu)
    if only is None and not args.no_cpu and specs is not None:
        cpu_samples(out, specs)
    return out


def find_cells(out, steps, cpu):
    """find_voronoi_cell: 1,000,000 uniform query points in the 1M uniform
    container ([0,100]^3, 58^3 blocks), queries and answers resident on the
    device; beside it the reference library's container::find_voronoi_cell
    on this host, one thread, over the same queries."""
    import torch
    import voro_amd
    from voro_amd import inputs as I
    ids, xyz, _ = I.make("uniform", 42)
    q = I.uniform_box(7, 1000000, (0, 0, 0), (100, 100, 100))
    con = voro_amd.container(0, 100, 0, 100, 0, 100, 58, 58, 58, False, False, False, 8)
    con.put_array(ids, xyz)
    dq = torch.from_numpy(q).to("cuda:0")
    found = torch.zeros(len(q), dtype=torch.int32, device="cuda:0")
    pid = torch.zeros(len(q), dtype=torch.int32, device="cuda:0")
    pos = torch.zeros((len(q), 3), dtype=torch.float64, device="cuda:0")
    L = voro_amd.lib()

    def run():
        voro_amd._check(L.vgpu_find_voronoi_cells(con._h, len(q), C.cast(dq.data_ptr(), C.POINTER(C.c_double)),
                                                  C.cast(found.data_ptr(), C.POINTER(C.c_int)),
                                                  C.cast(pid.data_ptr(), C.POINTER(C.c_int)),
                                                  C.cast(pos.data_ptr(), C.POINTER(C.c_double)), 1),
                        "vgpu_find_voronoi_cells")
    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    out["find_voronoi_cell_1m"] = {"queries_per_s": len(q) / dt, "ms": dt * 1e3, "queries": len(q),
                                   "found": int(found.sum().item())}
    if cpu:
        from oracle import oracle as O
        if O.reference() is not None:
            m = len(q)
            sp = O.make_spec(0, (0, 100, 0, 100, 0, 100), (58, 58, 58))
            f, p, _, secs = O.find_reference(sp, ids, xyz, None, q[:m], timed=True)
            same = bool(np.array_equal(p[f], pid[:m].cpu().numpy()[f]))
            out["find_voronoi_cell_1m"]["cpu_baseline"] = {
                "value": m / secs, "unit": "queries/s", "cores": 1, "kind": "reference",
                "sample": "container::find_voronoi_cell on the %d queries, %.2f s" % (m, secs),
                "same_particles": same}


def periodic_poly(out, flags, torch, steps):
    """The radical periodic workload, plain and with the radius-input
    predictor (trained on the plain path's statistics for seeds 1 and 2,
    evaluated on seed 42; results checked identical to the plain path's)."""
    import voro_amd
    from voro_amd import inputs as I

    def make(seed):
        ids, xyz, r = I.make("periodic_poly", seed)
        c = voro_amd.container_periodic_poly(60, 5, 60, 3, 4, 60, 74, 74, 74)
        d = [torch.from_numpy(a).to("cuda:0") for a in (ids, xyz, r)]
        torch.cuda.synchronize()
        c.put_device(len(ids), d[0].data_ptr(), d[1].data_ptr(), d[2].data_ptr())
        c._keep = d
        return c

    def timed(c, fl):
        c.compute_all_cells(fl)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            c.compute_all_cells(fl)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    con = make(42)
    tune_plain = con.autotune(flags, reps=1, config="periodic radical 2M")
    dt = timed(con, flags)
    res = con.results()
    n = len(res["id"])
    plain_ms = {k: float(v) for k, v in con.timings().items()}
    ref = (res["vol"].copy(), res["nbr"].copy())
    t0 = time.perf_counter()
    pred = voro_amd.predictor(with_radius=True)
    for seed in (1, 2):
        c = make(seed)
        c.compute_all_cells(flags | voro_amd.F_STATS)
        pred.add_samples(c)
        del c
    loss = pred.train(epochs=300, lr=3e-3, seed=1, fallback=0.02)
    t_train = time.perf_counter() - t0
    con.set_predictor(pred)
    tune_learned = con.autotune(flags | voro_amd.F_LEARNED, reps=1, config="periodic radical 2M, learned")
    dtl = timed(con, flags | voro_amd.F_LEARNED)
    r = con.results()
    ls = con.learned_stats()
    out["periodic_poly_2m"] = {"cells_per_s": n / dt, "ms": dt * 1e3,
                               "volume_sum_rel_err": abs(float(res["vol"].sum()) - 216000.0) / 216000.0,
                               "timings_ms": plain_ms, "autotune_best": tune_plain["best"],
                               "learned": {
                                   "autotune_best": tune_learned["best"],
                                   "cells_per_s": n / dtl, "ms": dtl * 1e3, "speedup_vs_plain": dt / dtl,
                                   "fallbacks": ls["fallbacks"], "fallback_rate": ls["fallbacks"] / n,
                                   "given_no_list": ls["routed"], "list_entries": ls["list_entries"],
                                   "train_seeds": [1, 2], "eval_seed": 42, "train_seconds": t_train,
                                   "train_loss": loss, "timings_ms": {k: float(v) for k, v in con.timings().items()},
                                   "identical_to_plain": bool(np.array_equal(r["vol"], ref[0])
                                                              and np.array_equal(r["nbr"], ref[1]))}}


def batch(out, flags, steps):
    import voro_amd
    from voro_amd import inputs as I
    t0 = time.perf_counter()
    specs = I.fcc_batch(11, 4096)
    cons = []
    for L, g, pts in specs:
        c = voro_amd.container(0, L, 0, L, 0, L, g, g, g, False, False, False, 8)
        c.add_wall(voro_amd.wall_sphere(L / 2, L / 2, L / 2, 0.45 * L, -99))
        c.put_array(np.arange(len(pts), dtype=np.int32), pts)
        cons.append(c)
    b = voro_amd.container_batch(cons)
    setup = time.perf_counter() - t0
    tune = voro_amd.container_batch(cons[:256]).autotune(flags)
    b.set_config(0, tune["occupancy"], tune["chunk"])

    def timed(fl):
        b.compute(fl, graph=True)
        t0 = time.perf_counter()
        for _ in range(steps):
            b.compute(fl, graph=True)
        return (time.perf_counter() - t0) / steps

    dt = timed(flags)
    n = sum(len(s[2]) for s in specs)
    sample = list(range(0, len(specs), 97))
    ref = [b.results(i)["vol"].copy() for i in sample]
    ref_nb = [b.results(i)["nbr"].copy() for i in sample]
    # the learned path: a block predictor trained on another draw of the
    # workload (512 containers, seed 12), the batch's statistics as samples
    t0 = time.perf_counter()
    tcons = []
    for L, g, pts in I.fcc_batch(12, 512):
        c = voro_amd.container(0, L, 0, L, 0, L, g, g, g, False, False, False, 8)
        c.add_wall(voro_amd.wall_sphere(L / 2, L / 2, L / 2, 0.45 * L, -99))
        c.put_array(np.arange(len(pts), dtype=np.int32), pts)
        tcons.append(c)
    tb = voro_amd.container_batch(tcons)
    tb.compute(flags | voro_amd.F_STATS)
    pred = voro_amd.predictor()
    pred.add_samples(tb)
    del tb, tcons
    loss = pred.train(epochs=300, lr=3e-3, seed=1, fallback=0.02)
    t_train = time.perf_counter() - t0
    b.set_predictor(pred)
    dtl = timed(flags | voro_amd.F_LEARNED)
    ls = b.learned_stats()
    same = all(np.array_equal(b.results(i)["vol"], r) and np.array_equal(b.results(i)["nbr"], q)
               for i, r, q in zip(sample, ref, ref_nb))
    out["batch_4096"] = {"cells_per_s": n / dt, "ms": dt * 1e3, "containers": len(specs), "particles": n,
                         "graph": True, "autotuned": tune, "setup_s": setup,
                         "learned": {"cells_per_s": n / dtl, "ms": dtl * 1e3, "speedup_vs_plain": dt / dtl,
                                     "graph": True, "fallbacks": ls["fallbacks"], "fallback_rate": ls["fallbacks"] / n,
                                     "given_no_list": ls["routed"], "list_records": ls["list_entries"],
                                     "train": "512 containers of seed 12", "train_seconds": t_train,
                                     "train_loss": loss,
                                     "identical_to_plain_on_sampled_containers": bool(same)}}
    return specs


def clusters_learned(out, flags, torch):
    """The clustered workload (1M particles, 64 Gaussian clusters + 30%
    uniform background, 58^3 blocks): the predictor trained on the plain
    path's statistics for seeds 1..10, evaluated on seed 101 against the
    plain path (results checked identical)."""
    import voro_amd
    from voro_amd import inputs as I
    dev = "cuda:0"

    def make(seed):
        ids, xyz, _ = I.make("clusters", seed)
        d_ids = torch.from_numpy(ids).to(dev)
        d_xyz = torch.from_numpy(xyz).to(dev)
        c = voro_amd.container(0, 100, 0, 100, 0, 100, 58, 58, 58, False, False, False, 8)
        c.put_device(len(ids), d_ids.data_ptr(), d_xyz.data_ptr())
        c._keep = (d_ids, d_xyz)
        return c

    t0 = time.perf_counter()
    pred = voro_amd.predictor()
    for seed in range(1, 11):
        c = make(seed)
        c.compute_all_cells(flags | voro_amd.F_STATS)
        pred.add_samples(c)
        del c
    loss = pred.train(epochs=400, lr=3e-3, seed=1, fallback=0.02)
    t_train = time.perf_counter() - t0
    con = make(101)
    reps = 3

    def timed(fl):
        con.compute_all_cells(fl)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            con.compute_all_cells(fl)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / reps

    tune_plain = con.autotune(flags, reps=1, config="clustered 1M")
    dt_plain = timed(flags)
    plain_ms = {k: float(v) for k, v in con.timings().items()}
    ref = con.results()
    ref = (ref["vol"].copy(), ref["nbr"].copy())
    con.set_predictor(pred)
    tune_learned = con.autotune(flags | voro_amd.F_LEARNED, reps=1, config="clustered 1M, learned")
    dt = timed(flags | voro_amd.F_LEARNED)
    r = con.results()
    n = len(r["id"])
    ls = con.learned_stats()
    out["clusters_learned"] = {
        "plain_cells_per_s": n / dt_plain, "learned_cells_per_s": n / dt, "speedup_vs_plain": dt_plain / dt,
        "fallbacks": ls["fallbacks"], "fallback_rate": ls["fallbacks"] / n, "train_seeds": list(range(1, 11)),
        "eval_seed": 101, "train_seconds": t_train, "train_loss": loss, "plain_timings_ms": plain_ms,
        "learned_timings_ms": {k: float(v) for k, v in con.timings().items()},
        "autotune_best": {"plain": tune_plain["best"], "learned": tune_learned["best"]},
        "identical_to_plain": bool(np.array_equal(r["vol"], ref[0]) and np.array_equal(r["nbr"], ref[1]))}


def cpu_samples(out, specs):
    """The reference (oracle/_ref) on bounded samples of the secondary
    workloads, on all of this process's cores."""
    import threading
    from concurrent.futures import ThreadPoolExecutor
    from oracle import oracle as O
    from voro_amd import inputs as I
    if O.reference() is None:
        return
    T = host_threads()
    # periodic: the same lattice shape at 1/8 the volume and particle count
    # (same density), 250,000 particles, 37^3 blocks; each thread its own
    # container, computing the cells of its z-slab of primary blocks
    ids2, xyz2, r2 = I.make("periodic_poly", 42, 250000)
    xyz2 = xyz2 * 0.5
    sp = O.make_spec(3, (30, 2.5, 30, 1.5, 2.0, 30), (37, 37, 37))
    mt = sorted(O.time_reference_mt(sp, ids2, xyz2, r2, threads=T) for _ in range(3))[1]
    one = O.time_reference(sp, ids2, xyz2, r2, loop="compute_all_cells")
    out["periodic_poly_2m"]["cpu_baseline"] = {
        "value": 250000 / mt, "unit": "cells/s", "cores": T, "kind": "reference", "cpu": cpu_model(),
        "sample": "container_periodic_poly(30,2.5,30,1.5,2,30) with 250,000 particles at the same density, "
                  "compute_all_cells() split over %d threads (z-slabs, one container copy each), median of 3: "
                  "%.3f s" % (T, mt),
        "one_thread": {"value": 250000 / one, "seconds": one}}
    # batch: containers shared out over the threads, each thread computing
    # whole containers (compute_all_cells() each); the rate is the cells over
    # the busiest thread's compute time
    sample = specs[:512]
    busy = {}
    lock = threading.Lock()

    def one_container(spec):
        L, g, pts = spec
        sp = O.make_spec(0, (0, L, 0, L, 0, L), (g, g, g), spheres=[[L / 2, L / 2, L / 2, 0.45 * L, -99]])
        t = O.time_reference(sp, np.arange(len(pts), dtype=np.int32), pts, loop="compute_all_cells")
        with lock:
            busy[threading.get_ident()] = busy.get(threading.get_ident(), 0.0) + t
    with ThreadPoolExecutor(max_workers=T) as ex:
        list(ex.map(one_container, sample))
    m = sum(len(s[2]) for s in sample)
    wall = max(busy.values())
    out["batch_4096"]["cpu_baseline"] = {
        "value": m / wall, "unit": "cells/s", "cores": T, "kind": "reference", "cpu": cpu_model(),
        "sample": "the first %d containers (%d particles), compute_all_cells() each, shared out over %d threads; "
                  "busiest thread %.2f s of compute, all threads %.2f s" % (len(sample), m, T, wall,
                                                                            sum(busy.values()))}
    # clustered: the evaluation draw itself, all cores
    if "clusters_learned" in out:
        ids3, xyz3, _ = I.make("clusters", 101)
        sp = O.make_spec(0, (0, 100, 0, 100, 0, 100), (58, 58, 58))
        mt = sorted(O.time_reference_mt(sp, ids3, xyz3, None, threads=T) for _ in range(3))[1]
        out["clusters_learned"]["cpu_baseline"] = {
            "value": len(ids3) / mt, "unit": "cells/s", "cores": T, "kind": "reference", "cpu": cpu_model(),
            "sample": "container::compute_all_cells() over the 1,000,000 clustered particles of seed 101 on %d "
                      "threads (z-slabs), median of 3: %.3f s" % (T, mt)}


if __name__ == "__main__":
    main()
