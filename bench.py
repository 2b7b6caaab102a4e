"""Benchmark of the particle hot path on one MI355X.

Headline (the JSON line's metric/value): the periodic Gaussian-mixture cell
list -- 3D, N=4,000,000 positions from 64 isotropic Gaussians (sigma 0.02,
wrapped) in the periodic unit cube, radius 0.01, n_particles_in_leaf 3.95
(100^3 cells).  One step = CellListOrdered build + radius search over all N
particles (neighbour counts, row offsets and the neighbour index lists).  The
metric is neighbour pairs per second.  Inputs are generated on the host and
uploaded before timing; the timed region starts with the positions resident
in HBM.

Also reported under "extra" (each outside the timed steps, each its own
workload):
  sparse      the Wendland C2 sparse operator of the same cell list on 8
              right-hand sides, FP64 and FP32-compensated
  c1          the reference's CPU-runnable case: 100,000 uniform points,
              radius 0.04, build + search and a Gaussian sparse matvec,
              beside the CPU oracle
  fmm         black-box Chebyshev FMM, N=2,000,000 uniform, 1/(r+1e-3),
              leaves <= 64, orders 6-8 uncompressed and with the M2L
              compressed at several tolerances, 16 RHS: matvec/s in FP64 and
              mixed precision (the M2L autotuner's choice reported),
              relative error against an FP64 direct sum on 10,000 seeded
              targets, and the fastest form at <= 1e-6
  rbf_batch   512 uneven 2D multiquadric (c=0.05) RBF systems, batched
              GMRES(50) on a bounded iteration budget, FP64 and FP32
              compensated operator (FP64 residual / orthogonalisation);
              Schwarz-preconditioned FP64 GMRES over the whole batch
              (Arnoldi steps replayed from HIP graphs), the same with
              iterative refinement (double-double residual), and the time to
              1e-8 (set-up + solve) of the problems with N < 2,000 as a
              batch of their own, with graphs and with a host check per step
  schwarz     2D jittered grid N=400,056, Gaussian sigma=0.01 (+1e-5 I) and a
              narrow variant sigma=0.002, GMRES to 1e-8 on a 300-iteration
              budget: plain Schwarz (with and without the coarse grid of
              set_coarse_grid_n(32)) vs the learned patch solves (the network
              trained on the patch residual with the fused training step,
              graph-replayed; the Schwarz fallback), the exact patch solves,
              the exact least-squares approximate inverse and none; setup /
              training / solve times and iterations
and a CPU baseline of the headline -- the oracle (C++/OpenMP restatement of
the reference's CPU search) timed on the host on a bounded sample of centres.
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
SECTIONS = ("search", "sparse", "c1", "fmm", "rbf", "schwarz", "copy")
DEFAULT_SECTIONS = ("search", "sparse", "c1", "fmm", "rbf", "schwarz")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--n", type=int, default=4_000_000)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--sections", default=",".join(DEFAULT_SECTIONS),
                   help="comma-separated subset of %s ('copy' is the calibration copy "
                        "of the HBM traffic counters, not run by default)" % (SECTIONS,))
    p.add_argument("--fused", action="store_true",
                   help="time the single-traversal search (abh_search_run_fused) as the step "
                        "instead of count pass + fill pass")
    p.add_argument("--fmm-orders", default="6,7,8",
                   help="Chebyshev orders of the fmm section (the first is reported at the top)")
    p.add_argument("--fmm-tols", default="none,1e-6,2e-7,1e-7",
                   help="M2L compression tolerances of the fmm section ('none': uncompressed)")
    p.add_argument("--no-rbf-schwarz", action="store_true",
                   help="skip the batched Schwarz part of the rbf section")
    p.add_argument("--rbf-iters", type=int, default=100,
                   help="GMRES iteration budget of the batched RBF systems")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Timer:
    """HIP-event timing on the context stream (torch's current stream)."""

    def __init__(self, torch, stream):
        self.torch, self.stream = torch, stream

    def __call__(self, fn, reps=1):
        ev = [self.torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ev[0].record(self.stream)
        for _ in range(reps):
            fn()
        ev[1].record(self.stream)
        self.torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]) / reps


def bench_search(args, torch, ab, wl, ctx, stream, extra):
    cfg = wl.MIXTURE
    n = args.n
    n_leaf = cfg["n_leaf"] * n / 4_000_000
    r = cfg["radius"]
    log("generating %d mixture positions" % n)
    pos_np = wl.mixture_positions(n, cfg["n_centres"], cfg["sigma"], seed=cfg["seed"])
    dev = ctx.device
    pos = torch.tensor(pos_np, device=dev)
    cells = ab.CellListOrdered(ctx, cfg["low"], cfg["high"], cfg["periodic"], n_leaf)
    cells.update_positions(pos)
    sp = cells.positions()
    search = ab.Search(cells, sp, r)
    total = search.total
    col = torch.empty(int(total * 1.02) + 1024, dtype=torch.int32, device=dev)
    log("grid %s, pairs %d (%.1f per particle)" % (cells.grid, total, total / n))
    build_ms, count_ms, fill_ms = [], [], []

    def step(record):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        ev[0].record(stream)
        cells.update_positions(pos)
        ev[1].record(stream)
        if not args.fused:
            search.run(cells, cells.positions(), r)
            ev[2].record(stream)
            search.fill(col)
        else:
            ev[2].record(stream)
            search.run_fused(cells, cells.positions(), r, col)
        ev[3].record(stream)
        if record:
            torch.cuda.synchronize()
            build_ms.append(ev[0].elapsed_time(ev[1]))
            count_ms.append(ev[1].elapsed_time(ev[2]))
            fill_ms.append(ev[2].elapsed_time(ev[3]))

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step(False)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = elapsed * 1e3 / args.steps
    for _ in range(max(1, args.steps)):  # per-phase breakdown (events)
        step(True)
    total = search.total
    value = total * args.steps / elapsed
    log("step %.3f ms, build %.3f ms, count %.3f ms, %s %.3f ms" % (
        ms_per_step, np.median(build_ms), np.median(count_ms),
        "fused search" if args.fused else "fill", np.median(fill_ms)))
    nb = cells.n_buckets
    fill_med = float(np.median(fill_ms))
    # algorithmic bytes of the fill launch: positions (24 B) + bucket
    # begin/end (8 B per bucket) + row offsets (8 B) read, 4 B per pair written
    fill_bytes = 24 * n + 8 * nb + 8 * n + 4 * total
    # measured HBM bytes per fill launch (rocprofv3 FETCH_SIZE / WRITE_SIZE,
    # scaled by a plain copy of known size), recorded beside the profiles
    traffic = None
    valu_insts = None
    tpath = os.path.join(ROOT, "profiles", "k_search_fill_traffic.json")
    if os.path.exists(tpath) and n == 4_000_000:
        with open(tpath) as fh:
            tj = json.load(fh)
        traffic = tj.get("hbm_bytes_per_launch")
        valu_insts = tj.get("valu_insts_per_launch")
    roofline = {"kernel": "k_search<3,fill>", "bound": "hbm",
                "achieved": fill_bytes / (fill_med * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "traffic": traffic,
                "bytes_per_launch": int(fill_bytes), "launch_ms": fill_med}
    roofline["frac"] = roofline["achieved"] / roofline["peak"]
    if valu_insts:
        # the fill's radius tests are FP64 vector work: its SQ_INSTS_VALU per
        # launch (rocprofv3) at 4 cycles per wave64 instruction on the 1,024
        # SIMDs at 2.4 GHz, over this run's launch time
        roofline["valu_issue_frac"] = valu_insts * 4 / (1024 * 2.4e9 * fill_med * 1e-3)
    extra["search"] = {"build_ms": float(np.median(build_ms)),
                       "count_ms": float(np.median(count_ms)), "fill_ms": fill_med,
                       "pairs": int(total), "grid": list(cells.grid),
                       "mode": "fused" if args.fused else "count + fill"}
    if not args.fused:
        # the single-traversal form of the same search, for comparison
        search.run_fused(cells, cells.positions(), r, col)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        ts = []
        for _ in range(3):
            ev[0].record(stream)
            search.run_fused(cells, cells.positions(), r, col)
            ev[1].record(stream)
            torch.cuda.synchronize()
            ts.append(ev[0].elapsed_time(ev[1]))
        extra["search"]["fused_search_ms"] = float(np.median(ts))
        log("single-traversal search %.3f ms (count + fill %.3f ms)" % (
            np.median(ts), np.median(count_ms) + fill_med))
    config = {"workload": "cell-list build + radius search, 3D periodic Gaussian mixture",
              "n": n, "radius": r, "n_particles_in_leaf": n_leaf, "grid": list(cells.grid),
              "n_centres": cfg["n_centres"], "sigma": cfg["sigma"]}
    state = {"cells": cells, "sp": sp, "pos_np": pos_np, "n_leaf": n_leaf, "r": r, "n": n}
    return value, ms_per_step, roofline, config, state


def bench_sparse(torch, ab, ctx, timer, st, extra):
    n, r = st["n"], st["r"]
    K = ab.SparseOperator(ctx, st["cells"], st["sp"], r, ab.Kernel.wendland_c2(r))
    x = torch.randn(8, n, dtype=torch.float64, device=ctx.device)
    y = torch.zeros_like(x)
    out = {}
    for prec, name in ((ab.FP64, "fp64"), (ab.FP32_COMPENSATED, "fp32c")):
        K.apply_colmajor(x, y, prec)
        out["matvec8_%s_ms" % name] = timer(lambda: K.apply_colmajor(x, y, prec), reps=3)
    log("sparse matvec 8 rhs: fp64 %.3f ms, fp32c %.3f ms" % (out["matvec8_fp64_ms"],
                                                               out["matvec8_fp32c_ms"]))
    # cached CSR form: assembled once, then streamed by every apply
    out["nnz"] = K.nnz
    for prec, name in ((ab.FP64, "fp64"), (ab.FP32_COMPENSATED, "fp32c")):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        K.cache(prec)
        torch.cuda.synchronize()
        out["cache_%s_setup_ms" % name] = (time.perf_counter() - t0) * 1e3
        K.apply_colmajor(x, y, prec)
        out["cached_matvec8_%s_ms" % name] = timer(lambda: K.apply_colmajor(x, y, prec), reps=3)
        K.cache(prec, enable=False)
    log("sparse matvec 8 rhs, cached CSR (%d pairs): fp64 %.3f ms (setup %.1f ms), fp32c %.3f ms "
        "(setup %.1f ms)" % (out["nnz"], out["cached_matvec8_fp64_ms"], out["cache_fp64_setup_ms"],
                             out["cached_matvec8_fp32c_ms"], out["cache_fp32c_setup_ms"]))
    K.close()
    extra["sparse"] = out


def bench_c1(args, torch, ab, wl, ctx, stream, extra):
    """The reference's CPU-runnable case: 100,000 uniform points in the unit
    cube (non-periodic), cell side = radius = 0.04 (25^3 buckets), a radius
    search for every particle and one Gaussian sparse matvec
    exp(-|dx|^2 / 0.02^2) on b ~ U[0,1); GPU step = build + count + fill,
    and the operator apply, beside the CPU oracle's build + search + matvec
    (all host threads)."""
    c = wl.UNIFORM
    pos_np = wl.uniform_positions(c["n"], seed=c["seed"])
    b_np = wl.uniform_rhs(c["n"], seed=c["seed"] + 1)[:, 0].copy()
    pos = torch.tensor(pos_np, device=ctx.device)
    cells = ab.CellListOrdered(ctx, c["low"], c["high"], c["periodic"], c["n_leaf"])
    cells.update_positions(pos)
    search = ab.Search(cells, cells.positions(), c["radius"])
    col = torch.empty(search.total + 1024, dtype=torch.int32, device=ctx.device)
    kern = ab.Kernel.gaussian(c["gauss_scale"])
    op = ab.SparseOperator(ctx, cells, cells.positions(), c["radius"], kern)
    b = torch.tensor(b_np, device=ctx.device)

    def step():
        cells.update_positions(pos)
        search.run(cells, cells.positions(), c["radius"])
        search.fill(col)

    for _ in range(3):
        step()
    torch.cuda.synchronize()
    reps = 20
    t0 = time.perf_counter()
    for _ in range(reps):
        step()
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) * 1e3 / reps
    y = torch.zeros_like(b)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    op.apply(b, y)
    ev[0].record(stream)
    for _ in range(reps):
        op.apply(b, y)
    ev[1].record(stream)
    torch.cuda.synchronize()
    mv_ms = ev[0].elapsed_time(ev[1]) / reps
    out = {"n": c["n"], "pairs": int(search.total), "grid": list(cells.grid),
           "build_search_ms": step_ms, "pairs_per_s": search.total / (step_ms * 1e-3),
           "matvec_ms": mv_ms}
    log("c1: build + search %.3f ms (%d pairs, %.3g pairs/s), Gaussian matvec %.3f ms" % (
        step_ms, search.total, out["pairs_per_s"], mv_ms))
    if not args.no_cpu:
        import oracle
        t0 = time.perf_counter()
        oc = oracle.Cells(c["low"], c["high"], c["periodic"], c["n_leaf"])
        oc.update(pos_np.copy())
        rp, _ = oc.search(oc.positions, c["radius"])
        t1 = time.perf_counter()
        oc.sparse_matvec(oc.positions, c["radius"], (1, c["gauss_scale"]), b_np[oc.order])
        t2 = time.perf_counter()
        out["cpu"] = {"build_search_ms": (t1 - t0) * 1e3, "matvec_ms": (t2 - t1) * 1e3,
                      "cores": oracle.nthreads_default(), "kind": "port",
                      "sample": "oracle build (serial) + search + Gaussian matvec of all "
                                "100,000 particles"}
        log("c1 cpu: build + search %.1f ms, matvec %.1f ms" % ((t1 - t0) * 1e3,
                                                               (t2 - t1) * 1e3))
    op.close()
    search.close()
    cells.close()
    extra["c1"] = out


def bench_fmm(args, torch, ab, wl, ctx, timer, extra):
    """Chebyshev FMM on N=2,000,000 uniform points, 16 right-hand sides, at the
    configured order 6 and at orders 7 and 8, FP64 and MIXED, each also with
    its M2L operators compressed (abh_fmm_compress, tolerances 1e-6, 2e-7
    and 1e-7): error against the FP64 direct sum on 10,000 seeded targets and
    matvec/s (one matvec = one right-hand side).  'at_1e-6' names the fastest
    form whose error is at most 1e-6."""
    c = wl.FMM
    pos = torch.tensor(wl.fmm_positions(c["n"], c["seed"]), device=ctx.device)
    kern = ab.Kernel.smooth_inv_r(c["eps"])
    x = torch.tensor(wl.fmm_rhs(c["n"], c["nrhs"], c["seed"] + 1), device=ctx.device)
    rng = np.random.default_rng(c["seed"] + 2)
    idx = torch.tensor(np.sort(rng.choice(c["n"], c["n_check"], replace=False)),
                       device=ctx.device)
    ref = ab.dense_apply(ctx, pos[idx].contiguous(), pos, kern, x)  # FP64 direct sum
    out = {"n_check": c["n_check"]}
    # the device HyperOctree alone: build (tags, sort, levels) and the dual
    # traversal's interaction lists on the same points
    tree = ab.HyperOctree(ctx, [0] * 3, [1] * 3, None, c["n_leaf"])
    tree.update_positions(pos)
    build_ms = timer(lambda: tree.update_positions(pos), reps=5)
    lists_ms = timer(lambda: tree.interaction_lists(), reps=3)
    ti = tree.info
    out["octree"] = {"build_ms": build_ms, "lists_ms": lists_ms, "levels": ti.levels,
                     "n_nodes": ti.n_nodes, "n_leaves": ti.n_leaves}
    log("octree: build %.2f ms, interaction lists %.2f ms (%d levels, %d nodes, %d leaves)" % (
        build_ms, lists_ms, ti.levels, ti.n_nodes, ti.n_leaves))
    tree.close()
    best = None
    orders = [int(v) for v in args.fmm_orders.split(",") if v]
    tols = [None if v == "none" else float(v) for v in args.fmm_tols.split(",") if v]
    for order in orders:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = ab.FmmOperator(ctx, pos, [0] * 3, [1] * 3, c["n_leaf"], order, kern)
        torch.cuda.synchronize()
        o = {"setup_s": time.perf_counter() - t0, "info": f.info}
        for tol in tols:
            tag = ""
            if tol is not None:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                kp, ranks = f.compress(tol)
                torch.cuda.synchronize()
                tag = "_compressed%.0e" % tol
                o["compressed%.0e" % tol] = {"setup_s": time.perf_counter() - t0,
                                             "padded_rank": kp, "ranks": ranks}
            for prec, name in ((ab.FP64, "fp64"), (ab.MIXED, "mixed")):
                if tol is None and order > 6 and prec == ab.MIXED:
                    continue  # the same FP64 M2L at these orders
                y = torch.zeros_like(x)
                f.apply_colmajor(x, y, prec)
                err = float(torch.linalg.norm(y[:, idx] - ref) / torch.linalg.norm(ref))
                del y
                ms = timer(lambda: f.apply_colmajor(x, torch.zeros_like(x), prec),
                           reps=3 if (order <= 6 or tol is not None) else 1)
                mps = c["nrhs"] / (ms * 1e-3)
                o[name + tag] = {"apply_ms": ms, "matvec_per_s": mps, "rel_err_vs_direct": err}
                if prec == ab.MIXED and tol is not None:
                    tt, tms = f.m2l_tuning()
                    o[name + tag]["m2l_tiles_per_block"] = tt
                    o[name + tag]["m2l_tuning_ms"] = tms
                log("fmm order %d %s%s: %d rhs in %.1f ms (%.1f matvec/s), rel err %.2e%s" % (
                    order, name, tag, c["nrhs"], ms, mps, err,
                    " (M2L tuner: %d tile(s)/block, %s ms)" % (tt, tms)
                    if prec == ab.MIXED and tol is not None else ""))
                if err <= 1e-6 and (best is None or mps > best["matvec_per_s"]):
                    best = {"order": order, "precision": name + tag, "matvec_per_s": mps,
                            "rel_err_vs_direct": err}
        out["order%d" % order] = o
        f.close()
        torch.cuda.empty_cache()
    # the first order's figures at the top level as well
    first = out["order%d" % orders[0]]
    out.update({k: first[k] for k in ("setup_s", "info", "fp64", "mixed") if k in first})
    out["at_1e-6"] = best
    log("fmm at 1e-6: %s" % best)
    extra["fmm"] = out


def bench_rbf(args, torch, ab, wl, ctx, timer, extra):
    c = wl.RBF_BATCH
    off, pos, f = wl.rbf_batch_problems(c["nprob"], c["n_min"], c["n_max"], c["seed"])
    P = torch.tensor(pos, device=ctx.device)
    b = torch.tensor(f, device=ctx.device)
    op = ab.DenseBatchOperator(ctx, off, P, ab.Kernel.multiquadric(c["c"]))
    sizes = np.diff(off)
    out = {"nprob": int(c["nprob"]), "total_points": int(off[-1]),
           "pairs_per_matvec": int((sizes.astype(np.float64) ** 2).sum()),
           "iteration_budget": args.rbf_iters}
    for prec, name in ((ab.FP64, "fp64"), (ab.FP32_COMPENSATED, "fp32c")):
        y = torch.zeros_like(b)
        op.apply(b, y, prec)
        mv = timer(lambda: op.apply(b, y, prec))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, stats = op.gmres(b, restart=c["restart"], max_iters=args.rbf_iters, tol=c["tol"],
                            precision=prec)
        torch.cuda.synchronize()
        solve_s = time.perf_counter() - t0
        res = np.array([s["residual"] for s in stats])
        conv = int(sum(s["converged"] for s in stats))
        out[name] = {"matvec_ms": mv, "pairs_per_s": out["pairs_per_matvec"] / (mv * 1e-3),
                     "solve_s": solve_s, "converged": conv,
                     "residual_median": float(np.median(res)),
                     "residual_max": float(res.max())}
        log("rbf %s: matvec %.1f ms, %d iters in %.1f s, converged %d/%d, residual median "
            "%.2e max %.2e" % (name, mv, args.rbf_iters, solve_s, conv, c["nprob"],
                               np.median(res), res.max()))
        out[name]["converged_by_size"] = by_size(sizes, stats)
    if args.no_rbf_schwarz:
        op.close()
        extra["rbf_batch"] = out
        return
    # restricted additive Schwarz per problem (signed local factorisations)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sw = ab.SchwartzPreconditioner.for_dense_batch(op, [0.0, 0.0], [1.0, 1.0],
                                                   c["schwarz_n_leaf"])
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    info = sw.info
    xs = torch.empty_like(b)
    ap = timer(lambda: sw.apply(b, xs))
    out["schwarz"] = {"n_leaf": c["schwarz_n_leaf"], "setup_s": setup_s, "apply_ms": ap,
                      "n_domains": info["n_domains"], "max_size": info["max_size"],
                      "factor_bytes": info["factor_bytes"],
                      "failed_domains": info["failed_domains"]}
    log("rbf schwarz: %d domains (max %d), %.1f GB inverse rows, setup %.1f s, apply %.2f ms"
        % (info["n_domains"], info["max_size"], info["factor_bytes"] / 1e9, setup_s, ap))
    # preconditioned solve until every problem has converged or stagnated
    # (two restarts in a row without a 10% residual cut: the FP64 floor)
    budget = max(args.rbf_iters, 150)
    out["stall_restarts"] = 2
    out["schwarz_iteration_budget"] = budget
    for prec, name in ((ab.FP64, "fp64_schwarz"),):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x, stats = op.gmres(b, restart=c["restart"], max_iters=budget, tol=c["tol"],
                            precision=prec, preconditioner=sw, stall_restarts=2, graph=True)
        torch.cuda.synchronize()
        solve_s = time.perf_counter() - t0
        res = np.array([s["residual"] for s in stats])
        conv = int(sum(s["converged"] for s in stats))
        its = np.array([s["iterations"] for s in stats])
        out[name] = {"solve_s": solve_s, "converged": conv, "graph": True,
                     "residual_median": float(np.median(res)), "residual_max": float(res.max()),
                     "iterations_converged_max": int(its[[s["converged"] for s in stats]].max())
                     if conv else None,
                     "converged_by_size": by_size(sizes, stats)}
        out[name]["iterations_max"] = int(its.max())
        log("rbf %s: finished in %.1f s (max %d iters), converged %d/%d, residual median %.2e "
            "max %.2e, %s" % (name, solve_s, int(its.max()), conv, c["nprob"], np.median(res),
                              res.max(), out[name]["converged_by_size"]))
    # iterative refinement: the residual in double-double, Schwarz-GMRES(50)
    # corrections to 1e-3 (one cycle each), four residual evaluations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x, stats = op.refine(b, tol=c["tol"], max_outer=4, inner_tol=1e-3, inner_iters=50,
                         preconditioner=sw, graph=True)
    torch.cuda.synchronize()
    solve_s = time.perf_counter() - t0
    res = np.array([s["residual"] for s in stats])
    conv = int(sum(s["converged"] for s in stats))
    its = np.array([s["iterations"] for s in stats])
    out["fp64_schwarz_refine"] = {
        "solve_s": solve_s, "converged": conv, "max_outer": 4, "inner_tol": 1e-3,
        "inner_iters": 50, "residual_median": float(np.median(res)),
        "iterations_max": int(its.max()), "converged_by_size": by_size(sizes, stats)}
    log("rbf refine (double-double residual, Schwarz-GMRES corrections): %.1f s, converged "
        "%d/%d, %s" % (solve_s, conv, c["nprob"], out["fp64_schwarz_refine"]["converged_by_size"]))
    sw.close()
    op.close()
    # time to 1e-8 on the problems FP64 can solve to 1e-8 (N < 2,000: the
    # rounding of the residual exceeds 1e-8 ||b|| above a few thousand), as a
    # batch of their own: Schwarz set-up + preconditioned GMRES(50)
    sel = [i for i, nn in enumerate(sizes) if nn < 2000]
    sub_off = np.concatenate([[0], np.cumsum(sizes[sel])])
    sub_pos = np.concatenate([pos[off[i]:off[i + 1]] for i in sel])
    sub_b = np.concatenate([f[off[i]:off[i + 1]] for i in sel])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sop = ab.DenseBatchOperator(ctx, sub_off, torch.tensor(sub_pos, device=ctx.device),
                                ab.Kernel.multiquadric(c["c"]))
    ssw = ab.SchwartzPreconditioner.for_dense_batch(sop, [0.0, 0.0], [1.0, 1.0],
                                                    c["schwarz_n_leaf"])
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    sbt = torch.tensor(sub_b, device=ctx.device)
    _, sstats = sop.gmres(sbt, restart=c["restart"], max_iters=budget, tol=c["tol"],
                          precision=ab.FP64, preconditioner=ssw, stall_restarts=2, graph=True)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    # the same solve with a host check after every Arnoldi step
    _, estats = sop.gmres(sbt, restart=c["restart"], max_iters=budget, tol=c["tol"],
                          precision=ab.FP64, preconditioner=ssw, stall_restarts=2)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    sconv = int(sum(st["converged"] for st in sstats))
    out["time_to_1e-8_small"] = {
        "problems": len(sel), "sizes": "1000-1999", "points": int(sub_off[-1]),
        "converged": sconv, "setup_s": t1 - t0, "solve_s": t2 - t1, "total_s": t2 - t0,
        "graph": True, "solve_s_eager_steps": t3 - t2,
        "eager_converged": int(sum(st["converged"] for st in estats)),
        "iterations_max": int(max(st["iterations"] for st in sstats))}
    log("rbf time to 1e-8, the %d problems with N < 2000: %d converged, setup %.2f s + solve "
        "%.3f s (graph-captured cycles; %.3f s with a host check per step)"
        % (len(sel), sconv, t1 - t0, t2 - t1, t3 - t2))
    ssw.close()
    sop.close()
    extra["rbf_batch"] = out


def by_size(sizes, stats):
    """Converged / total problems per size octave [1000, 2000), [2000, 4000) ..."""
    edges = [1000, 2000, 4000, 8000, 16000, 32000, 64001]
    out = {}
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = [i for i, n in enumerate(sizes) if lo <= n < hi]
        out["%d-%d" % (lo, hi - 1 if hi == 64001 else hi)] = "%d/%d" % (
            sum(stats[i]["converged"] for i in sel), len(sel))
    return out


def bench_schwarz(torch, ab, wl, ctx, extra):
    """Schwarz-preconditioned GMRES on the jittered grid (N=400,056): the
    configured kernel width sigma=0.01 and a narrow variant sigma=0.002 (same
    points), each with the plain Schwarz preconditioner, the same with the
    coarse-grid domain (set_coarse_grid_n(32)), the learned patch
    solves (the network trained on the patch residual; Schwarz fallback), the
    exact patch solves, the exact least-squares approximate inverse and no
    preconditioner, on bounded iteration budgets."""
    from aboria_amd import learned
    c = wl.SCHWARZ
    shift, k = c["shift"], c["k"]
    pos = wl.jittered_grid(c["nx"], c["ny"], c["jitter"], c["seed"])
    train_pos = wl.jittered_grid(c["nx"], c["ny"], c["jitter"], c["seed"] + 100)
    sync = torch.cuda.synchronize

    def timed(fn):
        sync()
        t0 = time.perf_counter()
        r = fn()
        sync()
        return r, time.perf_counter() - t0

    out = {}
    for name, sigma, leaf in (("sigma_0.01", c["sigma"], c["n_leaf"]), ("sigma_0.002", 0.002, 64.0)):
        radius = c["cutoff"] * sigma
        kern = ab.Kernel.gaussian(sigma)
        cl = ab.CellListOrdered(ctx, [0, 0], [1, 1], [False, False], leaf)
        cl.update_positions(torch.tensor(pos, device=ctx.device))
        op = ab.SparseOperator(ctx, cl, cl.positions(), radius, kern)
        b = torch.tensor(np.random.default_rng(c["seed"] + 1).normal(size=cl.n_alive),
                         device=ctx.device)
        r = {"n": int(cl.n_alive), "sigma": sigma, "n_leaf": leaf, "grid": list(cl.grid)}
        sw, t_sw = timed(lambda: ab.SchwartzPreconditioner(ctx, cl, kern, radius, shift))
        (_, st), t = timed(lambda: op.gmres(b, diag_shift=shift, preconditioner=sw, restart=50,
                                            max_iters=300))
        r["schwarz"] = {"setup_s": t_sw, "solve_s": t, "info": sw.info, **st}
        # with the coarse-grid domain of set_coarse_grid_n(32) (1,089
        # lattice points; seeded sample)
        swc, t_swc = timed(lambda: ab.SchwartzPreconditioner(ctx, cl, kern, radius, shift,
                                                             coarse_grid_n=32, seed=c["seed"]))
        (_, st), t = timed(lambda: op.gmres(b, diag_shift=shift, preconditioner=swc, restart=50,
                                            max_iters=300))
        r["schwarz_coarse32"] = {"setup_s": t_swc, "solve_s": t, "info": swc.info, **st}
        swc.close()
        # the network standing in for the patch local solves, trained on the
        # patch residual over a second seeded grid
        (X, M, _, oscale), t_data = timed(lambda: learned.training_set_patch(
            ctx, [train_pos], kern, radius, shift, 16.0, k, feature_scale=sigma,
            max_rows=400_000, seed=c["seed"]))
        (net, losses), t_train = timed(lambda: learned.train_residual(
            ctx, X, M, epochs=60, batch=1024, graph=True, fused=True))
        t_train_eager = None
        if name == "sigma_0.002":  # the same training, one kernel per layer operation
            (net_e, _), t_train_eager = timed(lambda: learned.train_residual(
                ctx, X, M, epochs=60, batch=1024, graph=True))
            net_e.close()
        del X, M
        ai, t_ai = timed(lambda: ab.ApproxInverse.learned(ctx, cl, net, k, sigma, oscale,
                                                          bf16=True))
        (_, rep), t = timed(lambda: ab.gmres_with_fallback(op, b, shift, ai, sw, restart=50,
                                                           max_iters=400, learned_iters=300))
        r["learned"] = {"training_rows": 400_000, "training_set_s": t_data,
                        "train_s": t_train, "train_graph": True, "train_fused": True,
                        "train_s_unfused": t_train_eager,
                        "loss_first": float(losses[0]),
                        "loss_last": float(losses[-1]), "setup_s": t_ai, "solve_s": t,
                        "learned": rep["learned"], "fallback": rep["fallback"]}
        pa, t_pa = timed(lambda: ab.ApproxInverse.patch(ctx, cl, kern, radius, shift, k))
        (_, st), t = timed(lambda: op.gmres(b, diag_shift=shift, preconditioner=pa, restart=50,
                                            max_iters=300))
        r["patch"] = {"setup_s": t_pa, "solve_s": t, **st}
        ls, t_ls = timed(lambda: ab.ApproxInverse.lsq(ctx, cl, kern, radius, shift, k))
        (_, st), t = timed(lambda: op.gmres(b, diag_shift=shift, preconditioner=ls, restart=50,
                                            max_iters=300))
        r["lsq"] = {"setup_s": t_ls, "solve_s": t, **st}
        (_, st), t = timed(lambda: op.gmres(b, diag_shift=shift, restart=50, max_iters=300))
        r["none"] = {"solve_s": t, **st}
        log("schwarz %s: plain %d its (%s) setup %.2f s solve %.2f s | coarse32 %d its (%s) "
            "res %.2e | learned %s fallback %s "
            "(train %.1f s fused%s, setup %.3f s, solve %.2f s) | patch %d its (%s) | lsq %d its "
            "(%s) | none %d its (%s)"
            % (name, r["schwarz"]["iterations"], r["schwarz"]["converged"], t_sw,
               r["schwarz"]["solve_s"], r["schwarz_coarse32"]["iterations"],
               r["schwarz_coarse32"]["converged"], r["schwarz_coarse32"]["residual"],
               rep["learned"], rep["fallback"], t_train,
               "" if t_train_eager is None else " / %.1f s unfused" % t_train_eager, t_ai,
               r["learned"]["solve_s"], r["patch"]["iterations"], r["patch"]["converged"],
               r["lsq"]["iterations"], r["lsq"]["converged"], r["none"]["iterations"],
               r["none"]["converged"]))
        out[name] = r
        for h in (ai, pa, ls, sw, net, op, cl):
            h.close()
        torch.cuda.empty_cache()
    extra["schwarz"] = out


def bench_copy(torch, ctx, timer, extra):
    """A plain device copy of known size (2 GiB read + 2 GiB written), the
    calibration kernel for the FETCH_SIZE / WRITE_SIZE counters."""
    n = (2 << 30) // 8
    x = torch.ones(n, dtype=torch.float64, device=ctx.device)
    y = torch.empty_like(x)
    y.copy_(x)
    ms = timer(lambda: y.copy_(x), reps=3)
    extra["copy"] = {"bytes_read": n * 8, "bytes_written": n * 8, "ms": ms,
                     "gbs": 2 * n * 8 / (ms * 1e-3) / 1e9}
    log("copy 2 GiB: %.3f ms, %.0f GB/s (read + write)" % (ms, extra["copy"]["gbs"]))


def cpu_baseline(st, value_unit="pairs/s"):
    import oracle
    from aboria_amd import workloads as wl
    cfg = wl.MIXTURE
    n, r, pos_np = st["n"], st["r"], st["pos_np"]
    threads = oracle.nthreads_default()
    log("cpu baseline: oracle build + search sample on %d threads" % threads)
    oc = oracle.Cells(cfg["low"], cfg["high"], cfg["periodic"], st["n_leaf"])
    tb = time.perf_counter()
    oc.update(pos_np.copy())
    build_s = time.perf_counter() - tb
    rng = np.random.default_rng(99)
    sample = 20000
    idx = rng.choice(n, sample, replace=False)
    tq = time.perf_counter()
    rp, _ = oc.search(oc.positions[idx], r)
    q_s = time.perf_counter() - tq
    if q_s < 10.0:  # scale the sample to about 10 s of CPU work
        sample = min(n, int(sample * 10.0 / max(q_s, 1e-3)))
        idx = rng.choice(n, sample, replace=False)
        tq = time.perf_counter()
        rp, _ = oc.search(oc.positions[idx], r)
        q_s = time.perf_counter() - tq
    pairs = int(rp[-1])
    log("cpu: %.3e pairs/s" % (pairs / q_s))
    st["oc"] = oc
    return {"value": pairs / q_s, "unit": value_unit, "cores": threads, "kind": "port",
            "sample": "oracle search of %d random centres of the N=%d mixture cell list "
                      "(%d pairs, %.2f s); serial oracle build of all N: %.2f s"
                      % (sample, n, pairs, q_s, build_s)}


def cpu_sparse(st, extra):
    """C2 operator on the CPU oracle: one right-hand side over a row sample,
    scaled to all N rows (the reference applies one RHS per call)."""
    import oracle
    oc, n, r = st["oc"], st["n"], st["r"]
    rng = np.random.default_rng(98)
    sample = 20000
    x = rng.standard_normal(n)
    for _ in range(2):
        idx = rng.choice(n, sample, replace=False)
        t0 = time.perf_counter()
        oc.sparse_matvec(oc.positions[idx], r, (2, r), x)
        t = time.perf_counter() - t0
        if t >= 5.0:
            break
        sample = min(n, int(sample * 8.0 / max(t, 1e-3)))
    ms = t * 1e3 * n / sample
    extra.setdefault("sparse", {})["cpu"] = {
        "ms_per_rhs_all_rows": ms, "cores": oracle.nthreads_default(), "kind": "port",
        "sample": "oracle Wendland C2 matvec, 1 RHS, %d of %d rows (%.2f s), scaled to all rows"
                  % (sample, n, t)}
    log("cpu sparse: %.0f ms per RHS over all rows" % ms)


def cpu_fmm(wl, extra):
    """C3 FMM on the CPU oracle at N = 20,000 (the first points of the C3
    set; same order, leaf size and kernel), one right-hand side."""
    import oracle
    c = wl.FMM
    n = 20000
    pos = wl.fmm_positions(c["n"], c["seed"])[:n]
    q = wl.fmm_rhs(c["n"], 1, c["seed"] + 1)[0][:n]
    ot = oracle.Octree([0, 0, 0], [1, 1, 1], pos, c["n_leaf"])
    t0 = time.perf_counter()
    ot.fmm_matvec(c["order"], (4, c["eps"]), q)
    t = time.perf_counter() - t0
    extra.setdefault("fmm", {})["cpu"] = {
        "n": n, "matvec_per_s": 1.0 / t, "cores": oracle.nthreads_default(), "kind": "port",
        "sample": "oracle FMM (M2L kernels recomputed per pair, as the reference) at N=%d, "
                  "1 RHS: %.2f s" % (n, t)}
    log("cpu fmm N=%d: %.2f s per matvec" % (n, t))


def cpu_rbf(wl, extra):
    """C4 batched matvec on the CPU oracle: the problems in batch order until
    about 1.5e10 pairs, pairs/s scaled to the batch's sum of N^2."""
    import oracle
    c = wl.RBF_BATCH
    off, pos, f = wl.rbf_batch_problems(c["nprob"], c["n_min"], c["n_max"], c["seed"])
    sizes = np.diff(off)
    pairs = t = 0.0
    used = 0
    for p in range(len(sizes)):
        a, b = off[p], off[p + 1]
        t0 = time.perf_counter()
        oracle.dense_matvec(pos[a:b], pos[a:b], (3, c["c"]), f[a:b])
        t += time.perf_counter() - t0
        pairs += float(sizes[p]) ** 2
        used += 1
        if pairs >= 1.5e10:
            break
    rate = pairs / t
    total = float((sizes.astype(np.float64) ** 2).sum())
    extra.setdefault("rbf_batch", {})["cpu"] = {
        "pairs_per_s": rate, "batched_matvec_s": total / rate,
        "cores": oracle.nthreads_default(), "kind": "port",
        "sample": "oracle dense multiquadric matvec of %d problems (%.3g pairs, %.2f s), "
                  "scaled to the batch's %.3g pairs" % (used, pairs, t, total)}
    log("cpu rbf: %.3e pairs/s, %.1f s per batched matvec" % (rate, total / rate))


def cpu_schwarz(wl, extra):
    """C5 plain Schwarz on the CPU oracle, sigma = 0.002 variant (the one the
    method solves): factorisation and one application over all N."""
    import oracle
    s = wl.SCHWARZ
    pos = wl.jittered_grid(s["nx"], s["ny"], s["jitter"], s["seed"])
    sigma = 0.002
    oc = oracle.Cells([0.0, 0.0], [1.0, 1.0], [0, 0], 64.0)
    oc.update(pos.copy())
    t0 = time.perf_counter()
    sw = oc.schwarz((1, sigma), s["cutoff"] * sigma, s["shift"])
    t1 = time.perf_counter()
    sw.apply(np.random.default_rng(1).normal(size=oc.n_alive))
    t2 = time.perf_counter()
    extra.setdefault("schwarz", {})["cpu"] = {
        "setup_s": t1 - t0, "apply_ms": (t2 - t1) * 1e3, "n_domains": sw.n_domains,
        "cores": oracle.nthreads_default(), "kind": "port",
        "sample": "oracle SchwartzPreconditioner<LLT> on the N=%d jittered grid, sigma=%g, "
                  "64 particles per bucket" % (oc.n_alive, sigma)}
    log("cpu schwarz sigma=%g: setup %.1f s, apply %.1f ms" % (sigma, t1 - t0, (t2 - t1) * 1e3))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only (the box has one); multi-GPU sharding of the "
              "cell list, FMM and solvers is a design note, not built", file=sys.stderr)
        sys.exit(2)
    sections = [s for s in args.sections.split(",") if s]
    for s in sections:
        if s not in SECTIONS:
            print("unknown section %s" % s, file=sys.stderr)
            sys.exit(2)
    import torch
    import aboria_amd as ab
    from aboria_amd import workloads as wl

    ctx = ab.Context(0)
    stream = torch.cuda.current_stream(ctx.device)
    timer = Timer(torch, stream)
    extra = {}
    value = ms_per_step = roofline = config = cpu = None
    st = None
    if "search" in sections or "sparse" in sections:
        value, ms_per_step, roofline, config, st = bench_search(args, torch, ab, wl, ctx, stream,
                                                                extra)
    if "sparse" in sections:
        bench_sparse(torch, ab, ctx, timer, st, extra)
    if st is not None and not args.no_cpu:
        cpu = cpu_baseline(st)
        if "sparse" in sections:
            cpu_sparse(st, extra)
    if st is not None:
        st["cells"].close()
        st.clear()
    torch.cuda.empty_cache()
    if "c1" in sections:
        bench_c1(args, torch, ab, wl, ctx, stream, extra)
    if "fmm" in sections:
        bench_fmm(args, torch, ab, wl, ctx, timer, extra)
        torch.cuda.empty_cache()
    if "rbf" in sections:
        bench_rbf(args, torch, ab, wl, ctx, timer, extra)
        torch.cuda.empty_cache()
    if "schwarz" in sections:
        bench_schwarz(torch, ab, wl, ctx, extra)
    if "copy" in sections:
        bench_copy(torch, ctx, timer, extra)
    if not args.no_cpu:
        # the CPU oracle beside the GPU numbers (bounded samples, host cores)
        if "fmm" in sections:
            cpu_fmm(wl, extra)
        if "rbf" in sections:
            cpu_rbf(wl, extra)
        if "schwarz" in sections:
            cpu_schwarz(wl, extra)

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
        "config": config,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
