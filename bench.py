"""Benchmark of the particle hot path on one MI355X.

Workload (default): the periodic Gaussian-mixture cell list -- 3D,
N=4,000,000 positions from 64 isotropic Gaussians (sigma 0.02, wrapped) in
the periodic unit cube, radius 0.01, n_particles_in_leaf 3.95 (100^3 cells).
One step = CellListOrdered build + radius search over all N particles
(neighbour counts, row offsets and the neighbour index lists).  The metric is
neighbour pairs per second.  Inputs are generated on the host and uploaded
before timing; the timed region starts with the positions resident in HBM.

Also reported (outside the timed steps): the Wendland C2 sparse operator on 8
right-hand sides in FP64 and FP32-compensated, and a CPU baseline -- the
oracle (C++/OpenMP restatement of the reference's CPU search) timed on the
host on a bounded sample of centres.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_VALU_PEAK_TFLOPS = 78.6


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=4_000_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-matvec", action="store_true")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only; run with --gpus 1 (multi-GPU sharding of the "
              "cell list is described in DESIGN.md but not built)", file=sys.stderr)
        sys.exit(2)
    import torch
    import aboria_amd as ab
    from aboria_amd import workloads as wl

    cfg = wl.MIXTURE
    n = args.n
    n_leaf = cfg["n_leaf"] * n / 4_000_000
    r = cfg["radius"]
    log("generating %d mixture positions" % n)
    pos_np = wl.mixture_positions(n, cfg["n_centres"], cfg["sigma"], seed=cfg["seed"])
    dev = torch.device("cuda", 0)
    pos = torch.tensor(pos_np, device=dev)
    ctx = ab.Context(0)
    stream = torch.cuda.current_stream(dev)
    cells = ab.CellListOrdered(ctx, cfg["low"], cfg["high"], cfg["periodic"], n_leaf)
    cells.update_positions(pos)
    sp = cells.positions()
    search = ab.Search(cells, sp, r)
    total = search.total
    col = torch.empty(int(total * 1.02) + 1024, dtype=torch.int32, device=dev)
    log("grid %s, pairs %d (%.1f per particle)" % (cells.grid, total, total / n))

    fill_ms = []
    count_ms = []
    build_ms = []

    def step(record):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e2 = torch.cuda.Event(enable_timing=True)
        e3 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        cells.update_positions(pos)
        e1.record(stream)
        search.run(cells, cells.positions(), r)
        e2.record(stream)
        search.fill(col)
        e3.record(stream)
        if record:
            torch.cuda.synchronize()
            build_ms.append(e0.elapsed_time(e1))
            count_ms.append(e1.elapsed_time(e2))
            fill_ms.append(e2.elapsed_time(e3))

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    # a second, per-phase timed pass (events) for the breakdown
    for _ in range(max(1, args.steps)):
        step(True)
    total = search.total
    value = total * args.steps / elapsed
    log("step %.3f ms, build %.3f ms, count %.3f ms, fill %.3f ms" % (
        ms_per_step, np.median(build_ms), np.median(count_ms), np.median(fill_ms)))

    nb = cells.n_buckets
    fill_med = float(np.median(fill_ms))
    fill_bytes = 24 * n + 8 * nb + 8 * n + 4 * total
    roofline = {
        "kernel": "k_search<3,fill>",
        "bound": "hbm",
        "achieved": fill_bytes / (fill_med * 1e-3) / 1e9,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "traffic": None,
    }
    roofline["frac"] = roofline["achieved"] / roofline["peak"]

    extra = {"build_ms": float(np.median(build_ms)), "count_ms": float(np.median(count_ms)),
             "fill_ms": fill_med, "pairs": int(total), "grid": list(cells.grid)}

    if not args.no_matvec:
        K = ab.SparseOperator(ctx, cells, sp, r, ab.Kernel.wendland_c2(r))
        x = torch.randn(8, n, dtype=torch.float64, device=dev)
        y = torch.zeros_like(x)
        for prec, name in ((ab.FP64, "fp64"), (ab.FP32_COMPENSATED, "fp32c")):
            K.apply_colmajor(x, y, prec)
            torch.cuda.synchronize()
            ts = []
            for _ in range(3):
                a = torch.cuda.Event(enable_timing=True)
                b = torch.cuda.Event(enable_timing=True)
                a.record(stream)
                K.apply_colmajor(x, y, prec)
                b.record(stream)
                torch.cuda.synchronize()
                ts.append(a.elapsed_time(b))
            extra["matvec8_%s_ms" % name] = float(np.median(ts))
        log("sparse matvec 8 rhs: fp64 %.3f ms, fp32c %.3f ms" % (
            extra["matvec8_fp64_ms"], extra["matvec8_fp32c_ms"]))

    cpu = None
    if not args.no_cpu:
        import oracle
        threads = oracle.nthreads_default()
        log("cpu baseline: oracle build + search sample on %d threads" % threads)
        oc = oracle.Cells(cfg["low"], cfg["high"], cfg["periodic"], n_leaf)
        tb = time.perf_counter()
        oc.update(pos_np.copy())
        build_s = time.perf_counter() - tb
        rng = np.random.default_rng(99)
        sample = 20000
        idx = rng.choice(n, sample, replace=False)
        tq = time.perf_counter()
        rp, _ = oc.search(oc.positions[idx], r)
        q_s = time.perf_counter() - tq
        # scale the sample so the timed CPU work is a few seconds
        if q_s < 2.0:
            sample = min(n, int(sample * 2.0 / max(q_s, 1e-3)))
            idx = rng.choice(n, sample, replace=False)
            tq = time.perf_counter()
            rp, _ = oc.search(oc.positions[idx], r)
            q_s = time.perf_counter() - tq
        pairs_sample = int(rp[-1])
        cpu = {"value": pairs_sample / q_s, "unit": "pairs/s", "cores": threads, "kind": "port",
               "sample": "oracle search of %d random centres of the N=%d mixture cell list "
                         "(%d pairs, %.2f s); serial oracle build of all N: %.2f s"
                         % (sample, n, pairs_sample, q_s, build_s)}
        log("cpu: %.3e pairs/s" % cpu["value"])

    out = {
        "metric": "pairs/s (cell-list query)",
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "cell-list build + radius search, 3D periodic Gaussian mixture",
                   "n": n, "radius": r, "n_particles_in_leaf": n_leaf, "grid": list(cells.grid),
                   "n_centres": cfg["n_centres"], "sigma": cfg["sigma"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
