"""Benchmark: converged Green-function columns per second at relative residual 1e-9.

Workload (one MI355X): the batched energy-contour solve.  The 32x32x32-block cell (128^3 grid, h = 0.25 Bohr)
with 64 seeded Gaussians and 128 atoms at numax 5; 128 (E, k, radius) instances = 32 contour energies from
E = -1 Ha to E_F = 0 Ha (Im E 0.005 .. 0.3 Ha) x 4 seeded k-points, truncation radius drawn from
{6, 8, 10, 12} Bohr per instance; one source block (64 columns) per instance; 8th-order kinetic stencil.
The cell is wrapped by every radius (2 r <= 32 Bohr), so the k-point leaves the operator unchanged, as in the
reference.  A step solves the 32 instances of one k-point as one batched plan (column groups = instances)
with FP32 COCG (the operator is complex symmetric: k is inert under Wrap) inside FP64 refinement; every column
is verified with the FP64 action and only columns whose true residual meets 1e-9 are counted.

Beside it: the plain FP64 path on the same 32 instances (one plan and tfQMR solve per instance, as the reference's
contour loop calls its solver per energy), the FP64 batched tfQMR and COCG solves, the local kernel's roofline from HIP events
on its launch stream, and the reference's own CPU action (oracle/_ref, single thread) on a bounded sample.
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
FP32_SHAPES = [(0, 16, 1), (0, 16, 2), (0, 16, 4), (0, 32, 1), (0, 32, 2), (0, 32, 4), (1, 32, 0), (1, 64, 0)]
# HBM bytes per launch of the FP32 local kernel per tile shape on this workload (rocprofv3 FETCH_SIZE and
# WRITE_SIZE of --mode action, scaled by a plain copy's counts); the command that made them is in the file
TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "local_fp32_contour32_traffic.json")
SOURCE = (16, 16, 16)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="batches of one k-point's contour instances")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--instances", type=int, default=32, help="instances per batch (<= 32)")
    ap.add_argument("--maxit", type=int, default=8000)
    ap.add_argument("--fp32-maxit", type=int, default=3000, help="operator applications allowed to one FP32 solve")
    ap.add_argument("--method", choices=["cocg", "tfqmr"], default="cocg",
                    help="inner Krylov method of the mixed solve (the FP64 comparison always runs tfQMR)")
    ap.add_argument("--inner-tol", type=float, default=0.0,
                    help="residual reduction asked of one FP32 solve (0: COCG with residual replacement goes to the "
                         "final tolerance in one solve; tfQMR uses 1e-4)")
    ap.add_argument("--action-reps", type=int, default=10)
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 comparison solves")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-actions", type=int, default=4)
    ap.add_argument("--mode", choices=["bench", "action", "contour"], default="bench",
                    help="action: only apply the action a few times (for profiler runs)")
    return ap.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def batch(instances, k, n):
    return instances[32 * (k % 4): 32 * (k % 4) + n]


def batch_plan(G, sysm, inst):
    import numpy as np
    p = G.GreenPlan(sysm, r_trunc=max(d["radius"] for d in inst), fd_range=(4, 4, 4), sources=[SOURCE] * len(inst),
                    source_radii=[d["radius"] for d in inst])
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

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    t0 = time.time()
    sysm = S.contour_32()
    instances = S.contour_instances()
    n = min(32, args.instances)
    log(f"# system built in {time.time() - t0:.1f} s")

    if args.mode == "action":
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
        # a plain device copy of one FP64 vector (nnzb * 64 KiB each way): the calibration of HBM counters
        x = torch.randn((p.nnzb, 2, 64, 64), dtype=torch.float64, device=dev)
        y = torch.empty_like(x)
        for _ in range(2):
            y.copy_(x)
        torch.cuda.synchronize()
        log(f"# copy calibration: {x.numel() * 8} bytes read and written per copy")
        return

    if args.mode == "contour":
        contour_mode(G, sysm, args)
        return

    inner = args.inner_tol if args.inner_tol > 0 else (1e-30 if args.method == "cocg" else 1e-4)
    opts = dict(tolerance=1e-9, max_iterations=args.maxit, max_refinements=6, inner_tolerance=inner,
                fp32_max_iterations=args.fp32_maxit, method=args.method)
    # ---- warmup (also tunes the tile shapes of both precisions) ----
    for w in range(args.warmup):
        p = batch_plan(G, sysm, batch(instances, w, n))
        _, st = p.solve(fp_bits=32, autotune=True, **opts)
        log(f"# warmup {w}: {st['converged_columns']}/{st['total_columns']} columns, {st['seconds']:.2f} s")
        p.close()

    # ---- the measured path: batched mixed-precision solves ----
    plans = [batch_plan(G, sysm, batch(instances, k, n)) for k in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    stats = []
    for k in range(args.steps):
        _, st = plans[k].solve(fp_bits=32, autotune=True, **opts)
        stats.append(st)
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t_start
    cols_mixed = sum(s["converged_columns"] for s in stats)
    per_inst = []
    for k, st in enumerate(stats):
        hs = st["column_halfsteps"].reshape(-1, 64).max(axis=1)
        log(f"# step {k}: {st['converged_columns']}/{st['total_columns']} columns in {st['seconds']:.2f} s, "
            f"{st['iterations']} half-steps, max residual {st['max_residual']:.2e}")
        if k == 0:
            per_inst = [int(h) for h in hs]
    nnzb = plans[0].nnzb
    for p in plans[1:]:
        p.close()
    plan = plans[0]
    first = batch(instances, 0, n)

    # ---- FP64: the same instances batched, and one at a time (the plain drop-in path) ----
    fp64 = {}
    if not args.no_fp64:
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        _, s64 = plan.solve(fp_bits=64, autotune=True, tolerance=1e-9, max_iterations=args.maxit, max_refinements=2)
        t_b64 = time.perf_counter() - t1
        log(f"# fp64 batched: {s64['converged_columns']} columns in {t_b64:.2f} s")
        fp64["batched_columns_per_s"] = round(s64["converged_columns"] / t_b64, 3)
        fp64["batched_half_steps"] = [int(h) for h in s64["column_halfsteps"].reshape(-1, 64).max(axis=1)]
        if args.method == "cocg":   # the same batch in FP64 with the mixed path's inner method
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            _, sc = plan.solve(fp_bits=64, autotune=True, tolerance=1e-9, max_iterations=args.maxit, max_refinements=2,
                               method="cocg")
            t_c = time.perf_counter() - t1
            log(f"# fp64 cocg batched: {sc['converged_columns']} columns in {t_c:.2f} s")
            fp64["cocg_batched_columns_per_s"] = round(sc["converged_columns"] / t_c, 3)
            # the mixed path with the reference's method (FP32 tfQMR inside FP64 refinement), same batch
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            _, sm = plan.solve(fp_bits=32, autotune=True, tolerance=1e-9, max_iterations=args.maxit, max_refinements=6,
                               inner_tolerance=1e-4, fp32_max_iterations=1000, method="tfqmr")
            t_m = time.perf_counter() - t1
            log(f"# mixed tfqmr batched: {sm['converged_columns']} columns in {t_m:.2f} s")
            fp64["mixed_tfqmr_batched_columns_per_s"] = round(sm["converged_columns"] / t_m, 3)
        cols, t_seq = 0, 0.0
        for d in first:   # plan built untimed (as the batched plans), device upload and solve timed
            p1 = batch_plan(G, sysm, [d])
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            _, s1 = p1.solve(fp_bits=64, tolerance=1e-9, max_iterations=args.maxit, max_refinements=2)
            t_seq += time.perf_counter() - t1
            cols += s1["converged_columns"]
            p1.close()
        log(f"# fp64 one instance at a time: {cols} columns in {t_seq:.2f} s")
        fp64.update(columns_per_s=round(cols / t_seq, 3), seconds=round(t_seq, 3), converged_columns=cols,
                    actions=int(s64["actions"]))

    # ---- roofline of the local kernel (potential + kinetic + SHO add-back), HIP events on its stream ----
    t_loc64 = time_action(plan, 64, 2, args.action_reps, stream)
    t_loc32 = time_action(plan, 32, 2, args.action_reps, stream)
    shape32 = plan.tile_shape(32)[:3]   # the FP32 local kernel the mixed solve ran (autotuned)
    t_act64 = time_action(plan, 64, 3, args.action_reps, stream)
    t_act32 = time_action(plan, 32, 3, args.action_reps, stream)
    nrows = plan.nRows
    alg64 = nnzb * 2 * BLOCK_BYTES[64] + nrows * 64 * 8
    alg32 = nnzb * 2 * BLOCK_BYTES[32] + nrows * 64 * 8
    plan.close()

    value = cols_mixed / t_mixed
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(sysm, first, per_inst, fp64.get("batched_half_steps"), args)

    out = {
        "metric": "Gf columns solved/s at resid 1e-9 (mixed+batched vs FP64); action % HBM peak",
        "value": round(value, 3),
        "unit": "Gf columns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mixed / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"c32 {args.method} + f64 refinement (per-group FP64 fallback)",
        "data": "synthetic (seeded Gaussian potential, random atom matrices, Gauss-Legendre contour)",
        "config": {"workload": "batched contour: 32^3-block cell, 128 atoms numax 5, 32 energies -1..0 Ha "
                               "(Im 0.005..0.3) x 4 k-points, radii {6,8,10,12} Bohr, one source block each",
                   "instances_per_step": n, "columns_per_step": 64 * n, "nnzb_per_step": int(nnzb)},
        "mixed_half_steps_per_instance": per_inst,
        "fp64": fp64 or None,
        "mixed_over_fp64": round(value / fp64["columns_per_s"], 3) if fp64 else None,
        "mixed_over_fp64_batched": round(value / fp64["batched_columns_per_s"], 3) if fp64 else None,
        "action": {"fp64_ms": round(t_act64 * 1e3, 3), "fp32_ms": round(t_act32 * 1e3, 3),
                   "local_fp64_ms": round(t_loc64 * 1e3, 3), "local_fp32_ms": round(t_loc32 * 1e3, 3),
                   "fp64_GBs": round(alg64 / t_act64 / 1e9, 1), "fp32_GBs": round(alg32 / t_act32 / 1e9, 1)},
        "roofline": {"bound": "hbm", "achieved": round(alg32 / t_loc32 / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg32 / t_loc32 / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": traffic_of(shape32, nnzb),
                     "kernel": f"FP32 local kernel, tile shape {shape32} ({'k_local' if shape32[0] == 0 else 'k_plane'}: "
                               "potential + 8th-order kinetic + SHO add-back), the one the mixed solve ran; "
                               "algorithmic bytes per launch = nnzb * 2 * 32 KiB + nRows * 512 B"},
        "cpu_baseline": cpu,
    }
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
                                                    inner_tolerance=1e-30, fp32_max_iterations=3000,
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


def cpu_baseline(sysm, first, half_steps_mixed, half_steps_fp64, args):
    """The reference's own CPU action (built from its sources into oracle/_ref; the oracle's C restatement if
    that build is absent), single thread, on one radius-8 instance; scaled by block count to every instance
    and by the FP64 solve's actions per instance to converged columns per second (vector updates of the CPU
    solve not counted, so this is an upper bound of the CPU rate)."""
    import numpy as np
    from angstromcube_amd import gfa as G
    from oracle import refgreen as R
    nblocks = {}
    for d in first:
        r = d["radius"]
        if r not in nblocks:
            p = G.GreenPlan(sysm, r_trunc=r, fd_range=(4, 4, 4), sources=[SOURCE])
            nblocks[r] = p.nnzb
            p.close()
    sample = next(d for d in first if d["radius"] == 8.0) if any(d["radius"] == 8.0 for d in first) else first[0]
    kind = "reference"
    x = np.random.default_rng(3).standard_normal
    try:
        if not R.available():
            raise RuntimeError("no reference build")
        rp = R.RefPlan(sysm, r_trunc=sample["radius"], fd_range=(4, 4, 4), nsrc=(1, 1, 1))
        rp.set_inputs(sample["E"])
        nb = rp.nnzb
        xv = x((nb, 2, 64, 64))
        t = time.perf_counter()
        for _ in range(args.cpu_actions):
            rp.action(xv)
        t_action = (time.perf_counter() - t) / args.cpu_actions
    except Exception as e:
        log(f"# cpu baseline falls back to the oracle: {e}")
        kind = "port"
        from oracle import oracle as O
        p = G.GreenPlan(sysm, r_trunc=sample["radius"], fd_range=(4, 4, 4), sources=[SOURCE])
        o = O.Oracle(p, sample["E"])
        nb = p.nnzb
        xv = x((nb, 2, 64, 64))
        t = time.perf_counter()
        for _ in range(args.cpu_actions):
            o.action(xv)
        t_action = (time.perf_counter() - t) / args.cpu_actions
        p.close()
    per_block = t_action / nb
    hs = half_steps_fp64 or half_steps_mixed
    if not hs:
        return {"value": None, "unit": "Gf columns/s", "cores": 1, "kind": kind,
                "sample": f"{args.cpu_actions} CPU actions of {t_action:.2f} s; no iteration counts"}
    t_cpu = sum(per_block * nblocks[d["radius"]] * (h + 2) for d, h in zip(first, hs))
    return {"value": round(64.0 * len(first) / t_cpu, 6), "unit": "Gf columns/s", "cores": 1, "kind": kind,
            "sample": f"{args.cpu_actions} single-thread CPU actions on the radius-{sample['radius']:g} instance "
                      f"({nb} blocks, {t_action:.2f} s each), scaled by block count to the {len(first)} instances "
                      f"and by the FP64 GPU solve's actions per instance; vector updates not counted; "
                      f"host {platform.processor() or platform.machine()}"}


if __name__ == "__main__":
    main()
