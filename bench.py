"""Benchmark: converged Green-function columns per second at relative residual 1e-9.

Workload (one MI355X): the batched energy-contour solve.  The 32x32x32-block cell (128^3 grid, h = 0.25 Bohr)
with 64 seeded Gaussians and 128 atoms at numax 5; 128 (E, k, radius) instances = 32 contour energies from
E = -1 Ha to E_F = 0 Ha (Im E 0.005 .. 0.3 Ha) x 4 seeded k-points, truncation radius drawn from
{6, 8, 10, 12} Bohr per instance; one source block (64 columns) per instance; 8th-order kinetic stencil.
The cell is wrapped by every radius (2 r <= 32 Bohr), so the k-point leaves the operator unchanged, as in the
reference.  A step solves the 32 instances of one k-point as one batched plan (column groups = instances)
with FP32 COCG (the operator is complex symmetric: k is inert under Wrap) inside FP64 refinement; steps cycle
over the 4 k-points, so every 4 steps cover the 128 instances.  Every column is verified with the FP64 action
and only columns whose true residual meets 1e-9 are counted.

Each step's plan is built, uploaded and given the tile shapes the warm-up tuned before its timer starts; its
solver state is released after the step, outside the timer.  The value is the converged columns of the timed
steps over the sum of their solve times.

Beside it: the FP64 solution of every batch (FP64 COCG, batched) and the mixed solution's error against it; on a
sample of 32 instances spread over the 4 batches (every energy once) the plain FP64 path (one plan and FP64 tfQMR
solve per instance, as the reference's contour loop calls its solver per energy); on batch 0 the FP64 tfQMR and
mixed tfQMR batched solves; the local kernel's roofline from HIP events on its launch stream; and the oracle's
CPU TFQMR (single thread) on a bounded sample.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BLOCK_BYTES = {64: 2 * 64 * 64 * 8, 32: 2 * 64 * 64 * 4}
# FP32 local-kernel tile shapes (kind, columns, lines) timed by --mode action
FP32_SHAPES = [(0, 16, 1), (0, 16, 2), (0, 16, 4), (0, 32, 1), (0, 32, 2), (0, 32, 4), (1, 32, 0), (1, 64, 0),
               (2, 16, 3), (2, 32, 3), (2, 64, 3), (2, 16, 2), (2, 32, 2), (2, 64, 2)]
FP64_SHAPES = [(0, 16, 2), (1, 32, 0), (2, 16, 3), (2, 32, 3), (2, 16, 2), (2, 32, 2)]
# HBM bytes per launch of the FP32 local kernel per tile shape on this workload (rocprofv3 FETCH_SIZE and
# WRITE_SIZE of --mode action, scaled by a plain copy's counts); the command that made them is in the file
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "local_fp32_contour32_traffic.json")
# MFMA utilisation of the matrix-core forms of the SHO projection and add-back on this workload (rocprofv3
# SQ_VALU_MFMA_BUSY_CYCLES, SQ_INSTS_VALU_MFMA_MOPS_* of --mode dyadic); the commands are in the file
MFMA_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "dyadic_mfma_contour32_counters.json")
KERNEL_OF_KIND = {0: "k_local", 1: "k_plane", 2: "k_lines"}
SOURCE = (16, 16, 16)
NBATCH = 4


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4, help="batches of one k-point's contour instances (cycling k)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=32, help="instances per batch (<= 32)")
    ap.add_argument("--maxit", type=int, default=8000)
    ap.add_argument("--fp32-maxit", type=int, default=0,
                    help="operator applications allowed to one FP32 solve (0: 1200 for COCG, whose column groups then "
                         "go on in FP64, 3000 for tfQMR)")
    ap.add_argument("--fp32-check", type=int, default=-1,
                    help="COCG: applications after which column groups whose residual is still above 1e-3 leave the "
                         "FP32 solve for FP64 (-1: 2/5 of the FP32 budget; 0: no check)")
    ap.add_argument("--method", choices=["cocg", "tfqmr"], default="cocg",
                    help="inner Krylov method of the mixed solve")
    ap.add_argument("--inner-tol", type=float, default=0.0,
                    help="residual reduction asked of one FP32 solve (0: COCG with residual replacement goes to the "
                         "final tolerance in one solve; tfQMR uses 1e-4)")
    ap.add_argument("--action-reps", type=int, default=10)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solves")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-halfsteps", type=int, default=10, help="half-steps of the timed CPU TFQMR sample")
    ap.add_argument("--cpu-cores", type=int, default=16,
                    help="host threads of the all-cores CPU sample (at most the logical CPUs; 1 skips it)")
    ap.add_argument("--budget-s", type=float, default=420.0,
                    help="wall seconds after which the optional batch-0 comparisons are skipped")
    ap.add_argument("--mode", choices=["bench", "action", "dyadic", "contour", "batch128", "batch"], default="bench",
                    help="action: only apply the action a few times (for profiler runs); dyadic: the action with "
                         "each engine of the dyadic term (for counter runs); batch128: the 128 instances in one plan; "
                         "batch: one mixed solve of the batch --batch (for profiler runs)")
    ap.add_argument("--batch", type=int, default=0, help="k-point batch of --mode batch")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def batch(instances, k, n):
    b = k % NBATCH
    return instances[32 * b: 32 * b + n]


def batch_plan(G, sysm, inst):
    p = G.GreenPlan(sysm, r_trunc=max(d["radius"] for d in inst), fd_range=(4, 4, 4), sources=[SOURCE] * len(inst),
                    source_radii=[d["radius"] for d in inst])
    import numpy as np
    p.set_inputs(np.array([d["E"] for d in inst]), inst[0]["k"])
    return p


def time_action(plan, fp, part, reps, stream):
    """Average duration of `reps` applications (part 3: whole action, 2: local kernel) from HIP events recorded
    on the launch stream."""
    import torch
    dt = torch.float64 if fp == 64 else torch.float32
    x = torch.randn((plan.nnzb, 2, 64, 64), dtype=dt, device="cuda")
    y = torch.empty_like(x)
    plan.multiply_part(y, x, 3, fp, stream)   # coefficients for part 2, warm caches of the code objects
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    for _ in range(reps):
        plan.multiply_part(y, x, part, fp, stream)
    ev[1].record(stream)
    torch.cuda.synchronize()
    del x, y
    return ev[0].elapsed_time(ev[1]) / 1e3 / reps


def rel_error(A, B):
    """max |A - B| / max |B| over the whole batch, and per column group (64 columns of one instance)."""
    import torch
    d = (A - B).abs().amax(dim=(1, 2, 3))
    return float((A - B).abs().max() / B.abs().max()), d


def fp32_policy(args):
    """(FP32 budget, progress check) of the mixed solve: COCG 1200 applications and a check at 2/5 of the budget,
    tfQMR 3000 and no check, unless given on the command line."""
    budget = args.fp32_maxit or (1200 if args.method == "cocg" else 3000)
    if args.method != "cocg":
        return budget, 0
    return budget, (args.fp32_check if args.fp32_check >= 0 else 2 * budget // 5)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the Green-function column groups would shard across GPUs with no exchange); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from angstromcube_amd import gfa as G
    from angstromcube_amd import synthetic as S

    t_begin = time.time()
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    t0 = time.time()
    sysm = S.contour_32()
    instances = S.contour_instances()
    n = min(32, args.instances)
    log(f"# system built in {time.time() - t0:.1f} s")

    if args.mode == "action":
        action_mode(G, sysm, instances, n, args, stream, dev)
        return
    if args.mode == "dyadic":
        p = batch_plan(G, sysm, batch(instances, 0, n))
        dyadic_engines(p, args, stream)
        return
    if args.mode == "contour":
        contour_mode(G, sysm, args)
        return

    inner = args.inner_tol if args.inner_tol > 0 else (1e-30 if args.method == "cocg" else 1e-4)
    fp32_budget, fp32_check = fp32_policy(args)
    opts = dict(tolerance=1e-9, max_iterations=args.maxit, max_refinements=6, inner_tolerance=inner,
                fp32_max_iterations=fp32_budget, fp32_progress_check=fp32_check, method=args.method)

    if args.mode == "batch128":
        batch128_mode(G, sysm, instances, opts)
        return
    if args.mode == "batch":
        batch_mode(G, sysm, instances, n, args, opts)
        return

    # ---- warmup (the first one also tunes the tile shapes and dyadic engines of both precisions) ----
    shapes, engines = None, None

    def pin(p, fps=(32, 64)):   # the tuned local-kernel shapes and dyadic engines
        if shapes:
            for fp in fps:
                p.set_tile_shape(fp, *shapes[fp])
                p.set_dyadic_engine(fp, engines[fp])

    for w in range(args.warmup):
        p = batch_plan(G, sysm, batch(instances, w, n))
        pin(p)
        _, st = p.solve(fp_bits=32, autotune=True, **opts)
        if shapes is None:
            shapes = {fp: p.tile_shape(fp)[:3] for fp in (32, 64)}
            engines = {fp: p.dyadic_engine(fp) for fp in (32, 64)}
        log(f"# warmup {w}: {st['converged_columns']}/{st['total_columns']} columns, {st['seconds']:.2f} s, "
            f"tile shapes {shapes}, dyadic engines {engines}")
        p.close()

    # ---- the measured path: batched mixed-precision solves, one plan per step ----
    t_mixed = 0.0
    cols_mixed = 0
    per_batch = {}
    errors = {}
    fp64_cocg = {}
    per_inst = []
    per_batch_hs = {}
    for k in range(args.steps):
        b = k % NBATCH
        inst = batch(instances, k, n)
        p = batch_plan(G, sysm, inst)   # untimed: host plan, device upload, workspaces, pinned tile shapes
        pin(p)
        p.prepare(32)
        p.prepare(64)
        keep = (b not in errors) and not args.no_fp64
        X32 = torch.empty((p.nnzb, 2, 64, 64), dtype=torch.float64, device=dev) if keep else None
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _, st = p.solve(fp_bits=32, autotune=True, X=X32, **opts)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t1
        t_mixed += dt
        cols_mixed += st["converged_columns"]
        per_batch.setdefault(b, []).append(st["converged_columns"] / dt)
        log(f"# step {k} (k-point {b}, {p.nnzb} blocks): {st['converged_columns']}/{st['total_columns']} columns "
            f"in {dt:.2f} s, {st['iterations']} applications, max residual {st['max_residual']:.2e}, "
            f"solver {p.sizes.solver_bytes / 1e9:.1f} GB")
        hs_inst = [int(h) for h in st["column_halfsteps"].reshape(-1, 64).max(axis=1)]
        if k < NBATCH:
            per_batch_hs[b] = hs_inst
            log(f"#   operator applications per instance: {hs_inst}; host syncs {st['host_syncs']}, "
                f"graph captures {st['graph_captures']}")
        if k == 0:
            per_inst = hs_inst
        if keep:
            # the FP64 solution of the same batch (FP64 COCG, batched), untimed for the mixed rate
            p.release_solver()
            X64 = torch.empty_like(X32)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            _, s64 = p.solve(fp_bits=64, tolerance=1e-9, max_iterations=args.maxit, max_refinements=2,
                             method="cocg" if args.method == "cocg" else "tfqmr", X=X64)
            torch.cuda.synchronize()
            t64 = time.perf_counter() - t2
            fp64_cocg[b] = s64["converged_columns"] / t64
            e, per_group = rel_error(X32, X64)
            errors[b] = e
            log(f"#   fp64 {'cocg' if args.method == 'cocg' else 'tfqmr'} batched: {s64['converged_columns']} "
                f"columns in {t64:.2f} s; |X_mixed - X_fp64| / |X_fp64| = {e:.2e}")
            del X32, X64, per_group
        p.close()
    value = cols_mixed / t_mixed

    # ---- FP64 comparisons ----
    fp64 = {}
    first = batch(instances, 0, n)
    if not args.no_fp64:
        fp64["batched_columns_per_s_by_batch"] = {str(b): round(v, 3) for b, v in sorted(fp64_cocg.items())}
        # the plain FP64 path: one plan and FP64 tfQMR solve per instance, on 32 instances spread over the batches
        sample = [instances[32 * b + i] for b in range(NBATCH) for i in range(b, 32, NBATCH)]
        cols, t_seq, hs_seq, nb_seq = 0, 0.0, [], []
        for d in sample:
            p1 = batch_plan(G, sysm, [d])
            pin(p1, (64,))
            p1.prepare(64)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            _, s1 = p1.solve(fp_bits=64, tolerance=1e-9, max_iterations=args.maxit, max_refinements=2)
            torch.cuda.synchronize()
            t_seq += time.perf_counter() - t1
            cols += s1["converged_columns"]
            hs_seq.append(int(s1["column_halfsteps"].max()))
            nb_seq.append(int(p1.nnzb))
            p1.close()
        log(f"# fp64 one instance at a time ({len(sample)} instances, every energy once): {cols} columns in "
            f"{t_seq:.2f} s")
        fp64.update(columns_per_s=round(cols / t_seq, 3), seconds=round(t_seq, 3), converged_columns=cols,
                    sample="32 instances: from batch b the energies b, b+4, ..., b+28")
        # batch 0 with the other methods, while the wall-clock budget allows
        if time.time() - t_begin < args.budget_s:
            p = batch_plan(G, sysm, first)
            pin(p)
            p.prepare(32)
            p.prepare(64)
            for name, kw in (("tfqmr_batched_columns_per_s",
                              dict(fp_bits=64, tolerance=1e-9, max_iterations=args.maxit, max_refinements=2)),
                             ("mixed_tfqmr_batched_columns_per_s",
                              dict(fp_bits=32, tolerance=1e-9, max_iterations=args.maxit, max_refinements=6,
                                   inner_tolerance=1e-4, fp32_max_iterations=1000, method="tfqmr"))):
                if time.time() - t_begin > args.budget_s:
                    break
                p.release_solver()
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                _, sx = p.solve(**kw)
                torch.cuda.synchronize()
                tx = time.perf_counter() - t1
                fp64[name + "_batch0"] = round(sx["converged_columns"] / tx, 3)
                log(f"# batch 0 {name}: {sx['converged_columns']} columns in {tx:.2f} s")
            p.close()

    # ---- roofline of the local kernel (potential + kinetic + SHO add-back), HIP events on its stream ----
    plan = batch_plan(G, sysm, first)
    pin(plan)
    nnzb = plan.nnzb
    t_loc64 = time_action(plan, 64, 2, args.action_reps, stream)
    t_loc32 = time_action(plan, 32, 2, args.action_reps, stream)
    shape32 = plan.tile_shape(32)[:3]   # the FP32 local kernel the mixed solve ran
    t_act64 = time_action(plan, 64, 3, args.action_reps, stream)
    t_act32 = time_action(plan, 32, 3, args.action_reps, stream)
    nrows = plan.nRows
    alg64 = nnzb * 2 * BLOCK_BYTES[64] + nrows * 64 * 8
    alg32 = nnzb * 2 * BLOCK_BYTES[32] + nrows * 64 * 8
    # the whole action with the dyadic term on each engine: vector ALUs (sum-factorised) and matrix cores (dense)
    engine_ms = {}
    for fp in (32, 64):
        used = plan.dyadic_engine(fp)
        for eng in (0, 1):
            plan.set_dyadic_engine(fp, eng)
            engine_ms[f"{('vector', 'matrix')[eng]}_fp{fp}"] = round(
                time_action(plan, fp, 3, args.action_reps, stream) * 1e3, 3)
        plan.set_dyadic_engine(fp, used)
    plan.close()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(G, sysm, sample if not args.no_fp64 else first,
                           hs_seq if not args.no_fp64 else per_inst, args)

    out = {
        "metric": "Gf columns solved/s at resid 1e-9 (mixed+batched vs FP64); action % HBM peak",
        "value": round(value, 3),
        "unit": "Gf columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mixed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "none (one GPU)",
        "vs_baseline": None,
        "dtype": f"c32 {args.method} + f64 refinement (groups over the FP32 budget of {fp32_budget}"
                 + (f", or above residual 1e-3 after {fp32_check} applications," if fp32_check else "")
                 + " go on in FP64)",
        "data": "synthetic (seeded Gaussian potential, random atom matrices, Gauss-Legendre contour)",
        "config": {"workload": "batched contour: 32^3-block cell, 128 atoms numax 5, 32 energies -1..0 Ha "
                               "(Im 0.005..0.3) x 4 k-points, radii {6,8,10,12} Bohr, one source block each; "
                               "a step = the 32 instances of one k-point in one plan",
                   "instances_per_step": n, "columns_per_step": 64 * n},
        "columns_per_s_by_batch": {str(b): round(float(np.mean(v)), 3) for b, v in sorted(per_batch.items())},
        "mixed_vs_fp64_max_rel_error": max(errors.values()) if errors else None,
        "mixed_vs_fp64_rel_error_by_batch": {str(b): v for b, v in sorted(errors.items())} or None,
        "mixed_half_steps_per_instance_batch0": per_inst,
        "mixed_applications_per_instance_by_batch": {str(b): v for b, v in sorted(per_batch_hs.items())},
        "fp64": fp64 or None,
        "mixed_over_fp64": round(value / fp64["columns_per_s"], 3) if fp64 else None,
        "action": {"fp64_ms": round(t_act64 * 1e3, 3), "fp32_ms": round(t_act32 * 1e3, 3),
                   "local_fp64_ms": round(t_loc64 * 1e3, 3), "local_fp32_ms": round(t_loc32 * 1e3, 3),
                   "fp64_GBs": round(alg64 / t_act64 / 1e9, 1), "fp32_GBs": round(alg32 / t_act32 / 1e9, 1),
                   "nnzb_batch0": int(nnzb),
                   "local_fp64_frac": round(alg64 / t_loc64 / 1e9 / HBM_PEAK_GBS, 4)},
        "dyadic_engines": {"used": {f"fp{fp}": ("vector", "matrix")[engines[fp]] for fp in (32, 64)} if engines
                           else None,
                           "action_ms": engine_ms,
                           "mfma": mfma_of(nnzb)},
        "roofline": {"bound": "hbm", "achieved": round(alg32 / t_loc32 / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg32 / t_loc32 / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic_of(shape32, nnzb),
                     "kernel": f"FP32 local kernel, tile shape {shape32} ({KERNEL_OF_KIND.get(shape32[0])}: "
                               "potential + 8th-order kinetic + SHO add-back), the one the mixed solve ran, on the "
                               "batch-0 plan; algorithmic bytes per launch = nnzb * 2 * 32 KiB + nRows * 512 B"},
        "cpu_baseline": cpu,
        "wall_s": round(time.time() - t_begin, 1),
    }
    print(json.dumps(out))


def action_mode(G, sysm, instances, n, args, stream, dev):
    import torch
    p = batch_plan(G, sysm, batch(instances, 0, n))
    for part in (3, 2):
        t = time_action(p, 64, part, args.action_reps, stream)
        log(f"# action fp64 part {part}: {t * 1e3:.3f} ms for {p.nnzb} blocks")
    t = time_action(p, 32, 3, args.action_reps, stream)
    log(f"# action fp32 part 3: {t * 1e3:.3f} ms for {p.nnzb} blocks")
    for shape in FP32_SHAPES:   # each FP32 local-kernel shape, in this order (per-dispatch counters follow it)
        p.set_tile_shape(32, *shape)
        t = time_action(p, 32, 2, args.action_reps, stream)
        log(f"# local fp32 shape {shape}: {t * 1e3:.3f} ms for {p.nnzb} blocks")
    for shape in FP64_SHAPES:
        p.set_tile_shape(64, *shape)
        t = time_action(p, 64, 2, args.action_reps, stream)
        log(f"# local fp64 shape {shape}: {t * 1e3:.3f} ms for {p.nnzb} blocks")
    p.set_tile_shape(32, 1, 32, 0)   # the default shapes again for the engine comparison
    p.set_tile_shape(64, 2, 16, 2)
    dyadic_engines(p, args, stream)
    # a plain device copy of one FP64 vector (nnzb * 64 KiB each way): the calibration of HBM counters
    x = torch.randn((p.nnzb, 2, 64, 64), dtype=torch.float64, device=dev)
    y = torch.empty_like(x)
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"# copy calibration: {x.numel() * 8} bytes read and written per copy")


def dyadic_engines(p, args, stream):
    """The dyadic term's engines on plan p: part 1 (projection and atom matrices), part 2 (local kernel with the
    add-back inside it, or followed by the matrix-core add-back) and the whole action, in FP32 and FP64."""
    for fp in (32, 64):
        for eng in (0, 1):
            p.set_dyadic_engine(fp, eng)
            t1, t2, t3 = (time_action(p, fp, part, args.action_reps, stream) for part in (1, 2, 3))
            log(f"# dyadic engine {('vector', 'matrix')[eng]} fp{fp}: projection {t1 * 1e3:.3f} ms, local + add-back "
                f"{t2 * 1e3:.3f} ms, action {t3 * 1e3:.3f} ms")
        p.set_dyadic_engine(fp, 0)


def batch_mode(G, sysm, instances, n, args, opts):
    """One mixed solve of k-point batch --batch with the default tile shapes, and its per-instance statistics."""
    import torch
    p = batch_plan(G, sysm, batch(instances, args.batch, n))
    p.prepare(32)
    p.prepare(64)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, st = p.solve(fp_bits=32, **opts)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    hs = st["column_halfsteps"].reshape(-1, 64).max(axis=1)
    inst = batch(instances, args.batch, n)
    out = {"mode": "batch", "batch": args.batch, "nnzb": int(p.nnzb), "seconds": round(dt, 3),
           "columns_per_s": round(st["converged_columns"] / dt, 3), "applications": int(st["iterations"]),
           "host_syncs": int(st["host_syncs"]), "graph_captures": int(st["graph_captures"]),
           "instances": [{"E": [round(d["E"].real, 4), round(d["E"].imag, 4)], "radius": d["radius"],
                          "applications": int(h)} for d, h in zip(inst, hs)]}
    p.close()
    print(json.dumps(out))


def batch128_mode(G, sysm, instances, opts):
    """The 128 instances (4 k-points x 32 energies) in one plan and one solve (the k-point is inert under Wrap)."""
    import torch
    p = batch_plan(G, sysm, instances)
    sb, pb = p.solve_memory(32, opts["method"])
    log(f"# 128-instance plan: {p.nnzb} blocks, solver up to {sb / 1e9:.1f} GB, plan {pb / 1e9:.1f} GB")
    p.prepare(32)
    p.prepare(64)
    torch.cuda.synchronize()
    t = time.perf_counter()
    _, st = p.solve(fp_bits=32, autotune=True, **opts)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t
    out = {"mode": "batch128", "instances": len(instances), "nnzb": int(p.nnzb),
           "converged_columns": int(st["converged_columns"]), "total_columns": int(st["total_columns"]),
           "seconds": round(dt, 3), "columns_per_s": round(st["converged_columns"] / dt, 3),
           "max_residual": st["max_residual"], "solver_bytes": int(p.sizes.solver_bytes),
           "plan_bytes": int(p.sizes.device_bytes), "applications": int(st["iterations"])}
    p.close()
    print(json.dumps(out))


def contour_mode(G, sysm, args):
    """A whole contour integration (gfa_contour_density): the reference's default 32-point energy mesh, Gamma, 8
    source blocks (2x2x2 at the cell centre), radius 8 Bohr.  Batched mixed precision (8 energies x 8 sources
    per plan) against the reference's structure (one FP64 tfQMR solve of the 8 sources per energy point); prints
    one JSON line with both rates and the densities' agreement."""
    import numpy as np
    import torch
    from angstromcube_amd import contour as K
    E, wE = K.energy_mesh()
    kw = dict(nsources=(2, 2, 2), kpoints=[(0.0, 0.0, 0.0, 1.0)], E_Fermi=0.0, tolerance=1e-9)
    runs = {}
    for name, opts in (("batched_mixed_cocg", dict(max_groups=64, fp_bits=32, method="cocg", max_refinements=6,
                                                    inner_tolerance=1e-30, fp32_max_iterations=1200,
                                                    max_iterations=8000, autotune=True)),
                       ("per_energy_fp64_tfqmr", dict(max_groups=8, fp_bits=64, method="tfqmr", max_refinements=2,
                                                      max_iterations=8000))):
        torch.cuda.synchronize()
        t = time.perf_counter()
        rho, res, st = K.contour_density(sysm, 8.0, E, wE, **kw, **opts)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t
        runs[name] = (rho, st, dt)
        log(f"# contour {name}: {st['converged_columns']}/{st['total_columns']} columns in {dt:.2f} s, "
            f"{st['solves']} solves, max residual {st['max_residual']:.2e}")
    (r1, s1, t1), (r2, s2, t2) = runs["batched_mixed_cocg"], runs["per_energy_fp64_tfqmr"]
    out = {"mode": "contour", "energy_points": int(len(E)), "kpoints": 1, "sources": int(s1["sources"]),
           "radius_bohr": 8.0, "columns": int(s1["total_columns"]),
           "batched_mixed_columns_per_s": round(s1["converged_columns"] / t1, 2), "batched_mixed_seconds": round(t1, 3),
           "per_energy_fp64_columns_per_s": round(s2["converged_columns"] / t2, 2),
           "per_energy_fp64_seconds": round(t2, 3), "speedup": round(t2 / t1, 3),
           "density_max_rel_diff": float(np.abs(r1 - r2).max() / np.abs(r2).max()),
           "electrons_in_source_blocks": float(r2.sum() * sysm.dvol)}
    print(json.dumps(out))


def traffic_of(shape, nnzb):
    """HBM bytes (read + written) per launch of the FP32 local kernel of this tile shape, from the committed
    rocprofv3 counter summary of the same workload; None if it covers another workload or shape."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    e = t.get("shapes", {}).get(",".join(str(v) for v in shape))
    if t.get("nnzb") != int(nnzb) or e is None:
        return None
    return int(e["fetch_bytes"] + e["write_bytes"])


def mfma_of(nnzb):
    """MFMA utilisation and rate of the matrix-core SHO kernels from the committed rocprofv3 counter summary of the
    same workload; None if it covers another workload."""
    try:
        with open(MFMA_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("nnzb") != int(nnzb):
        return None
    return {k: {"mfma_util": e["mfma_util"], "TFLOPs": e["tflops"], "peak_TFLOPs": e["peak_tflops"]}
            for k, e in t.get("kernels", {}).items()}


def cpu_model():
    """The host CPU's model name (Linux), else the machine type."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def cpu_baseline(G, sysm, sample, half_steps, args):
    """The oracle's CPU TFQMR (a C restatement of the reference's action with the same TFQMR as the device: actions,
    vector updates and inner products), single thread, timed for --cpu-halfsteps half-steps on the radius-6
    instance of the sample; the time per block and half-step is scaled to the sample's instances by their block
    counts and the half-steps their plain FP64 GPU solves took (the same algorithm), giving columns per second."""
    from oracle import oracle as O
    small = min(sample, key=lambda d: d["radius"])
    p = G.GreenPlan(sysm, r_trunc=small["radius"], fd_range=(4, 4, 4), sources=[SOURCE])
    p.set_inputs(small["E"], small["k"])
    o = O.Oracle(p, small["E"], small["k"])
    nb = p.nnzb
    nhs = max(2, args.cpu_halfsteps)
    t = time.perf_counter()
    _, _, m = o.tfqmr(tol=1e-30, maxit=nhs)
    t_run = time.perf_counter() - t
    # the run applies the operator once per half-step, once to start and once for the final residual
    per_block_hs = t_run / ((m + 2) * nb)
    # all cores: the same run on every host core at once, one per thread (the oracle's C solve releases the GIL
    # and only reads the shared oracle state); throughput per block and half-step over all of them
    ncores = max(1, min(args.cpu_cores, os.cpu_count() or 1))
    t_all = None
    if ncores > 1:
        import threading
        gate = threading.Barrier(ncores + 1)
        def run():
            gate.wait()
            o.tfqmr(tol=1e-30, maxit=nhs)
        ths = [threading.Thread(target=run) for _ in range(ncores)]
        for th in ths:
            th.start()
        gate.wait()
        t = time.perf_counter()
        for th in ths:
            th.join()
        t_all = time.perf_counter() - t
    p.close()
    nblocks = {}
    for d in sample:
        r = d["radius"]
        if r not in nblocks:
            q = G.GreenPlan(sysm, r_trunc=r, fd_range=(4, 4, 4), sources=[SOURCE])
            nblocks[r] = q.nnzb
            q.close()
    work = sum(nblocks[d["radius"]] * (h + 2) for d, h in zip(sample, half_steps))   # blocks x half-steps
    out = {"value": round(64.0 * len(sample) / (per_block_hs * work), 6), "unit": "Gf columns/s", "cores": 1,
           "kind": "port",
           "sample": f"oracle CPU TFQMR, {m} half-steps on the radius-{small['radius']:g} instance ({nb} blocks) in "
                     f"{t_run:.1f} s (actions, vector updates and inner products), scaled by block count and by "
                     f"the FP64 GPU solves' half-steps to the {len(sample)} instances of the plain FP64 sample; "
                     f"host {cpu_model()}, {os.cpu_count()} logical CPUs"}
    if t_all is not None:
        per_all = t_all / ((m + 2) * nb * ncores)
        out["all_cores"] = {"value": round(64.0 * len(sample) / (per_all * work), 6), "unit": "Gf columns/s",
                            "cores": ncores,
                            "sample": f"the same run on {ncores} threads at once in {t_all:.1f} s (one per core, "
                                      "as independent instances), scaled the same way"}
    return out


if __name__ == "__main__":
    main()
