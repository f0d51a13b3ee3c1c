"""Benchmark of the LandSlide-HySEA solver step on one MI355X.

Headline workload: the large single-GPU case, 4096x4096 (plane shelf + 8
seeded Gaussian bumps, 3 seeded slides of 20-80 m, water at rest, walls, CFL
0.5, r 0.5, Coulomb 10 deg), forward steps of the fused step.  A "step" is
one solver time step over the whole grid: the four edge classes, the cell
update and the on-device dt reduction; in FP32C its launches are the
sediment-free tiles' kernel and the general tiles' kernel, which also closes
the step (tile_classes reports how many class tiles took the general one).

Prints one JSON line.  value = Mcell-updates/s of the headline path (FP32
with compensated state when available, else FP64) with the state resident in
HBM; the FP64 path is timed beside it.  roofline.achieved uses algorithmic
bytes per cell-update (104 B: read 6 state fields + depth, write 6 fields)
over the step kernels' average duration measured with HIP events on the
library's stream; roofline.traffic is the HBM bytes per step that rocprofv3's
FETCH_SIZE / WRITE_SIZE counters measured for the same workload.  Extra keys:
the 2048x2048 precision-comparison workload's forward steps (c2_2048), the
adjoint workloads (1024x1024 and 4096x4096), the 64-scenario ensemble (a
short timed run and the workload's full 5000 steps), the FP32C-vs-FP64
precision report at 2048x2048, and cpu_baseline (the FP64 oracle on the host
cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from dataclasses import replace

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPU baseline's OpenMP threads stay on their cores (read once, when the
# OpenMP runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

from landslidehysea_amd import scenario as S          # noqa: E402
from landslidehysea_amd import solver as L            # noqa: E402

# algorithmic bytes per cell-update: FP64 reads 6x8 B state + 8 B depth and writes
# 6x8 B; FP32c reads 6 values + 6 compensations (12x4 B) + 8 B depth and writes 12x4 B
BYTES_PER_CELL = 104
HBM_PEAK_GBS = 8000.0
# HBM bytes per FP32C step at 2048x2048 and 4096x4096 (the sediment-free and
# the general tiles' launches together), measured with rocprofv3 FETCH_SIZE
# (x2.0) and WRITE_SIZE (x1.0), the scale factors of a plain 256 MiB copy
# (profiles/fused_step_fp32c_2048x2048_counters.txt,
# profiles/fused_step_fp32c_4096x4096_counters.txt,
# profiles/copy_256MiB_counter_calibration.txt); below the algorithmic bytes
# because sediment-free tiles neither read their zero layer-2 planes nor write
# them again where the output buffer already holds them
TRAFFIC_BYTES_FP32C = {2048: 255.9e6, 4096: 976.3e6}
# workloads: grid side -> (seed, slides)
WORKLOADS = {2048: (2048, 4), 4096: (4096, 3)}
NX = NY = 2048      # the precision-comparison workload
SEED = 2048
HEAD_N = 4096       # the headline workload


def make_case(nx=NX, ny=NY, precision=L.FP64):
    seed, slides = WORKLOADS.get(nx, (SEED, 4))
    sc = S.make_scenario(nx, ny, seed=seed, n_slides=slides, thickness=(20.0, 80.0))
    H, _ = S.hmin_shift(sc.H)
    prm = L.Params(nx=nx, ny=ny, dx=sc.dx, dy=sc.dy, angle_deg=(10.0,), precision=precision)
    return sc, np.ascontiguousarray(sc.state), np.ascontiguousarray(H), prm


def time_gpu(precision, steps, warmup, n=HEAD_N):
    sc, W, H, prm = make_case(n, n, precision=precision)
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.initial_dt()
        if warmup:
            sim.step(warmup)
        sim.sync()
        t0 = time.perf_counter()
        total_ms, _ = sim.time_steps(steps, per_kernel=False)
        sim.sync()
        wall = time.perf_counter() - t0
        _, kernel_ms = sim.time_steps(max(steps // 2, 5), per_kernel=True)
        classes = sim.tile_classes()
        t, dtn, done = sim.time()
    N = n * n
    return {"n": n, "total_ms": total_ms, "wall_s": wall, "kernel_ms": kernel_ms, "classes": classes,
            "mcells": N * steps / (total_ms * 1e-3) / 1e6, "t": t, "steps_done": done,
            "achieved_gbs": BYTES_PER_CELL * N / (kernel_ms * 1e-3) / 1e9}


def time_adjoint(precision, steps=4000, snaps=64, n=1024, ngauges=16):
    """Adjoint workload: 1024x1024, one slide, gauge misfit over `steps` steps,
    binomial checkpointing with `snaps` HBM snapshots."""
    sc = S.make_scenario(n, n, seed=1024, n_slides=1)
    H, _ = S.hmin_shift(sc.H)
    prm = L.Params(nx=n, ny=n, dx=sc.dx, dy=sc.dy, angle_deg=(10.0,), precision=precision)
    gauges = S.gauge_cells(sc, ngauges, seed=16)
    obs = np.zeros((steps, len(gauges)))
    with L.Simulation(prm) as sim:
        sim.upload(np.ascontiguousarray(sc.state), np.ascontiguousarray(H))
        t0 = time.perf_counter()
        res = sim.adjoint(steps, snaps, gauges, obs)
        wall = time.perf_counter() - t0
    fwd, rev = res["forward_ms"], res["reverse_ms"]
    return {"grid": f"{n}x{n}", "steps": steps, "snapshots": snaps, "forward_ms": round(fwd, 2),
            "reverse_ms": round(rev, 2), "cost_vs_forward": round((fwd + rev) / fwd, 3),
            "mcells_fwd_plus_adjoint": round(n * n * steps / ((fwd + rev) * 1e-3) / 1e6, 2),
            "re_advances": res["advances"], "wall_s": round(wall, 2)}


def precision_report(steps=1000, angle=2.0, ngauges=16):
    """FP32C against FP64 on the 2048x2048 construction: relative L2 of the
    gauge signal (eta - sea level at rest; gauges among the wet cells within
    40 cells of a slide), of the eta field and of h_s, and the maximum runup;
    beside it the spread that rounding the initial h_s to FP32 size
    (x (1 + 1e-7 N(0,1))) alone causes in the FP64 solver."""
    from scipy import ndimage
    sc, W, H, p64 = make_case()
    p64 = replace(p64, angle_deg=(angle,))
    p32 = replace(p64, precision=L.FP32C)
    near = ndimage.distance_transform_edt(W[3] <= 0.0) <= 40.0
    cand = np.flatnonzero((near & (W[3] <= 0.0) & (W[0] > 1.0)).ravel())
    gauges = np.sort(np.random.default_rng(16).choice(cand, ngauges, replace=False)).astype(np.int32)

    def run(W0, prm):
        with L.Simulation(prm) as sim:
            sim.upload(W0, H)
            sim.track_eta_max(True)
            sim.set_gauges(gauges, steps)
            sim.step(steps)
            return sim.download(), sim.gauges(), sim.eta_max()[0]

    def rel(a, b):
        d = np.linalg.norm(b)
        return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))

    W64, g64, e64 = run(W, p64)
    W32, g32, e32 = run(W, p32)
    Wp = W.copy()
    Wp[3] = Wp[3] * (1.0 + 1e-7 * np.random.default_rng(7).standard_normal(Wp[3].shape))
    Wq, gq, _ = run(Wp, p64)
    rest = float((W[0] + W[3] - H)[W[0] > 1.0].mean())
    wet = W64[0] + W64[3] > 0.0

    def errs(Wx, gx):
        return {"gauges": rel(gx - rest, g64 - rest),
                "eta": rel(Wx[0] + Wx[3] - H - rest, W64[0] + W64[3] - H - rest),
                "h_s": rel(Wx[3], W64[3])}

    return {"steps": steps, "friction_deg": angle, "gauges": ngauges,
            "gauge_signal_max_m": float(np.abs(g64 - rest).max()),
            "fp32c_vs_fp64": {k: float(f"{v:.3e}") for k, v in errs(W32, g32).items()},
            "fp64_input_rounding_spread": {k: float(f"{v:.3e}") for k, v in errs(Wq, gq).items()},
            "max_runup_m": {"fp64": float(e64[wet].max()) - rest, "fp32c": float(e32[wet].max()) - rest}}


def time_large_adjoint(adj_steps=200, snaps=100, n=4096):
    """4096x4096 workload: the discrete adjoint (FP32C) of a bounded prefix of
    its 20000 steps with `snaps` HBM snapshots (0.8 GB each)."""
    sc, W, H, prm = make_case(n, n, precision=L.FP32C)
    gauges = S.gauge_cells(sc, 16, seed=16)
    obs = np.zeros((adj_steps, len(gauges)))
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        res = sim.adjoint(adj_steps, snaps, gauges, obs)
    fwd, rev = res["forward_ms"], res["reverse_ms"]
    return {"grid": f"{n}x{n}", "steps": adj_steps, "snapshots": snaps, "forward_ms": round(fwd, 2),
            "reverse_ms": round(rev, 2), "cost_vs_forward": round((fwd + rev) / fwd, 3),
            "re_advances": res["advances"]}


def time_ensemble(steps=200, warmup=16, n=64, tune_steps=16):
    """Ensemble workload: 64 seeded scenarios (sides U{256..1536}, one slide of
    volume logU[1e6, 5e7] m^3, friction U[6, 16] deg, r U[0.45, 0.6]) stepped
    together in FP32C by the batched kernel, the step loop replayed from a
    captured graph; the tile shape is autotuned first (the autotuner restores
    the state), then `warmup` untimed and `steps` timed steps (warmup 0 and
    5000 steps: every scenario runs the workload's 5000 steps)."""
    specs, build = S.make_ensemble(n)
    params = []
    for spec in specs:
        params.append(L.Params(nx=spec["nx"], ny=spec["ny"], dx=S.DEFAULT_SPACING_DEG * S.DEG_TO_M,
                               dy=S.DEFAULT_SPACING_DEG * S.DEG_TO_M, angle_deg=(spec["angle_deg"],),
                               r=spec["r"], precision=L.FP32C))
    cells = sum(p.nx * p.ny for p in params)
    with L.Batch(params) as b:
        for i, spec in enumerate(specs):
            sc, _ = build(spec)
            H, _ = S.hmin_shift(sc.H)
            b.upload(i, np.ascontiguousarray(sc.state), np.ascontiguousarray(H))
            del sc, H
        b.initial_dt()
        best, ms = b.autotune(tune_steps)
        if warmup:
            b.step(warmup)
        b.sync()
        total_ms = b.time_steps(steps)
        t0 = [b.time(i) for i in (0, n - 1)]
    shapes = L.batch_tile_shapes()
    return {"scenarios": n, "cells": cells, "steps": steps, "total_ms": round(total_ms, 3),
            "ms_per_step": round(total_ms / steps, 4), "mcells": round(cells * steps / (total_ms * 1e-3) / 1e6, 2),
            "tile": list(shapes[best]), "autotune_ms": {f"{s[0]}x{s[1]}/{s[2]}": round(float(m), 3)
                                                        for s, m in zip(shapes, ms)},
            "achieved_gbs": round(BYTES_PER_CELL * cells * steps / (total_ms * 1e-3) / 1e9, 1),
            "steps_done_first_last": [t0[0][2], t0[1][2]]}


def host_cpu():
    """CPU model and physical core count of the host (/proc/cpuinfo)."""
    model, cores = "unknown", set()
    try:
        phys = core = None
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "model name":
                    model = v
                elif k == "physical id":
                    phys = v
                elif k == "core id":
                    core = v
                elif not k and phys is not None:
                    cores.add((phys, core))
                    phys = core = None
    except OSError:
        pass
    return model, len(cores)


def cpu_baseline(c2_steps=20, reps=3):
    """FP64 oracle (OpenMP) on the host, as the CPU baseline is planned:
    threads bound close to their cores (OMP_PROC_BIND=close,
    OMP_PLACES=cores, set when this process starts, before any OpenMP runtime
    reads them), one thread per core of this process's CPU share
    (OMP_NUM_THREADS where the host sets it, else the cores this process may
    run on); the step loop only, timed by the oracle's omp_get_wtime.
    value = median of `reps` timings of a c2_steps-step prefix of the
    2048x2048 precision-comparison workload; c1 = the 400x300 plain-FP64
    workload in full (2000 steps, one timing)."""
    from oracle import binding as ob
    ob.build()
    try:
        share = len(os.sched_getaffinity(0))
    except Exception:
        share = os.cpu_count() or 1
    cores = int(os.environ.get("OMP_NUM_THREADS", share) or share)
    cores = max(1, min(cores, share))
    ob.set_threads(cores)
    model, phys = host_cpu()

    def rate(sc, steps, nrep):
        H, _ = S.hmin_shift(sc.H)
        W0 = np.ascontiguousarray(sc.state)
        p = ob.make_params(sc.nx, sc.ny, sc.dx, sc.dy, angle_deg=10.0)
        dt = ob.initial_dt(p, W0, H)
        out = []
        for _ in range(nrep):
            W = W0.copy()
            t0 = ob.wtime()
            ob.run(p, W, H, steps, dt=dt, record=False)
            out.append(sc.nx * sc.ny * steps / (ob.wtime() - t0) / 1e6)
        return out

    seed, slides = WORKLOADS[2048]
    c2 = rate(S.make_scenario(2048, 2048, seed=seed, n_slides=slides, thickness=(20.0, 80.0)), c2_steps, reps)
    c1 = rate(S.make_scenario(400, 300, seed=1), 2000, 1)
    return {"value": float(np.median(c2)), "unit": "Mcell-updates/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle, 2048x2048 precision-comparison workload, {c2_steps}-step prefix, "
                      f"median of {reps}",
            "runs": [round(x, 3) for x in c2], "c1_400x300_2000_steps": round(c1[0], 3),
            "cpu_model": model, "host_physical_cores": phys,
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}


def copy_probe(mib=256):
    """Plain device-to-device copy of a known size: the calibration kernel for the
    FETCH_SIZE / WRITE_SIZE counters (torch's copy launches one copy kernel)."""
    import torch
    n = mib * 1024 * 1024 // 4
    x = torch.ones(n, device="cuda", dtype=torch.float32)
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4}), flush=True)


def summarize(dbs):
    """Per-kernel summary of rocprofv3 result databases (kernel trace and counters)."""
    import sqlite3
    for db in dbs:
        c = sqlite3.connect(db)
        print(f"# {db}")
        try:
            rows = c.execute("select name, total_calls, total_duration, average, percentage from top_kernels").fetchall()
            if rows:
                print("kernel | calls | total_us | avg_us | pct")
                for name, n, tot, avg, pct in rows:
                    print(f"{name[:110]} | {n} | {tot:.1f} | {avg:.3f} | {pct:.2f}")
        except sqlite3.Error:
            pass
        try:
            rows = c.execute("select kernel_name, counter_name, count(*), avg(value) from counters_collection "
                             "group by kernel_name, counter_name").fetchall()
            if rows:
                print("kernel | counter | dispatches | avg_value (KiB for FETCH_SIZE/WRITE_SIZE, else counts)")
                for name, cn, n, v in rows:
                    print(f"{name[:110]} | {cn} | {n} | {v:.1f}")
        except sqlite3.Error:
            pass


def gaps(argv):
    """Idle time between consecutive kernels in a rocprofv3 kernel trace
    (--kernel-trace --output-format csv): `--gaps PATTERN LAST csv...` takes the
    last LAST dispatches whose name contains PATTERN (the steady state of a
    timed loop), in start order, and reports per kernel its count and mean
    duration, and per transition (kernel -> next kernel) the count, mean,
    median and maximum gap from one kernel's end to the next one's start."""
    import csv
    import statistics
    pat, last, files = argv[0], int(argv[1]), argv[2:]
    rows = []
    for fn in files:
        with open(fn, newline="") as f:
            for r in csv.DictReader(f):
                name = r["Kernel_Name"]
                if pat in name:
                    short = name.split("(")[0].split("<")[0].replace("void ", "").split("::")[-1]
                    rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short))
    rows.sort()
    rows = rows[-last:]
    if len(rows) < 2:
        print("fewer than two matching dispatches")
        return
    span_us = (rows[-1][1] - rows[0][0]) / 1e3
    busy_us = sum(e - s for s, e, _ in rows) / 1e3
    print(f"dispatches: {len(rows)}; first start to last end {span_us:.1f} us; kernels busy {busy_us:.1f} us "
          f"({100 * busy_us / span_us:.2f}%); idle between kernels {span_us - busy_us:.1f} us")
    per = {}
    for s, e, n in rows:
        per.setdefault(n, []).append((e - s) / 1e3)
    print("kernel | dispatches | mean_us")
    for n, d in per.items():
        print(f"{n} | {len(d)} | {statistics.mean(d):.3f}")
    tr = {}
    for (s0, e0, n0), (s1, e1, n1) in zip(rows, rows[1:]):
        tr.setdefault(f"{n0} -> {n1}", []).append((s1 - e0) / 1e3)
    print("transition | count | mean_gap_us | median_gap_us | max_gap_us")
    for k, g in tr.items():
        print(f"{k} | {len(g)} | {statistics.mean(g):.3f} | {statistics.median(g):.3f} | {max(g):.3f}")


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--summarize":
        summarize(sys.argv[2:])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--gaps":
        gaps(sys.argv[2:])
        return
    if len(sys.argv) > 1 and sys.argv[1] == "--copy-probe":
        copy_probe()
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="timed forward steps of the headline workload")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--precision", choices=["both", "fp32c", "fp64"], default="both",
                    help="paths to time (profiling runs time one)")
    ap.add_argument("--size", type=int, choices=[2048, 4096], default=HEAD_N,
                    help="grid side of the headline forward timing (profiling runs of the 2048 workload)")
    ap.add_argument("--c2-steps", type=int, default=200,
                    help="timed forward steps of the 2048x2048 workload beside the headline (0 skips it)")
    ap.add_argument("--adjoint-steps", type=int, default=4000,
                    help="steps of the 1024x1024 adjoint workload (0 skips it)")
    ap.add_argument("--adjoint-snapshots", type=int, default=64)
    ap.add_argument("--ensemble-steps", type=int, default=200,
                    help="timed steps of the 64-scenario ensemble workload (0 skips it)")
    ap.add_argument("--ensemble-full-steps", type=int, default=5000,
                    help="the ensemble at the workload's full step count, no warmup (0 skips it)")
    ap.add_argument("--large-adjoint-steps", type=int, default=200,
                    help="steps of the 4096x4096 adjoint (100 HBM snapshots; the workload's full length is 20000)")
    ap.add_argument("--precision-steps", type=int, default=10000,
                    help="steps of the FP32C-vs-FP64 precision report at 2048x2048 (0 skips it)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the whole grid lives in one MI355X's HBM); "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)

    n = args.size
    fp32c = fp64 = None
    if args.precision in ("both", "fp32c"):
        fp32c = time_gpu(L.FP32C, args.steps, args.warmup, n)
    if args.precision in ("both", "fp64"):
        fp64 = time_gpu(L.FP64, max(args.steps // 4, 10) if fp32c else args.steps,
                        max(args.warmup // 4, 2) if fp32c else args.warmup, n)
    head = fp32c if fp32c else fp64
    dtype = "f32 (compensated state, f64 depth/hydrostatics/dt)" if fp32c else "f64"
    seed, slides = WORKLOADS[n]
    workload = ("large single-GPU case 4096x4096, forward step" if n == 4096
                else "precision comparison 2048x2048, forward step")
    out = {
        "metric": "Mcell-updates/s (FP32-comp forward; plain FP64 beside it) on one MI355X",
        "value": round(head["mcells"], 2),
        "unit": "Mcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["total_ms"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": f"synthetic (seeded shelf + bumps + {slides} slides)",
        "config": {"workload": workload, "nx": n, "ny": n, "slides": slides, "friction_deg": 10.0, "cfl": 0.5},
        "fp64_mcells": round(fp64["mcells"], 2) if fp64 else None,
        "speedup_vs_fp64": round(fp32c["mcells"] / fp64["mcells"], 3) if (fp32c and fp64) else None,
        "fp64_kernel_ms": round(fp64["kernel_ms"], 5) if fp64 else None,
        "fp64_achieved_gbs": round(fp64["achieved_gbs"], 1) if fp64 else None,
        "roofline": {"bound": "hbm", "achieved": round(head["achieved_gbs"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(head["achieved_gbs"] / HBM_PEAK_GBS, 4),
                     "traffic": TRAFFIC_BYTES_FP32C.get(n) if head is fp32c else None,
                     "kernel_ms": round(head["kernel_ms"], 5), "bytes_per_cell": BYTES_PER_CELL},
    }
    if fp32c:
        nt, ngen, split = fp32c["classes"]
        out["tile_classes"] = {"class_tiles": nt, "general_tiles": ngen, "split_launch": split}
    if args.c2_steps > 0 and n != 2048:
        c2 = {}
        if fp32c:
            r = time_gpu(L.FP32C, args.c2_steps, args.warmup, 2048)
            c2["fp32c"] = {"mcells": round(r["mcells"], 2), "kernel_ms": round(r["kernel_ms"], 5),
                           "achieved_gbs": round(r["achieved_gbs"], 1),
                           "frac": round(r["achieved_gbs"] / HBM_PEAK_GBS, 4), "traffic": TRAFFIC_BYTES_FP32C[2048],
                           "general_tiles": r["classes"][1]}
        if fp64:
            r = time_gpu(L.FP64, max(args.c2_steps // 4, 10), 5, 2048)
            c2["fp64"] = {"mcells": round(r["mcells"], 2), "kernel_ms": round(r["kernel_ms"], 5),
                          "achieved_gbs": round(r["achieved_gbs"], 1)}
        out["c2_2048"] = c2
    if args.adjoint_steps > 0:
        adj = {}
        if fp32c:
            adj["fp32c"] = time_adjoint(L.FP32C, args.adjoint_steps, args.adjoint_snapshots)
            # every state kept in HBM (1024^2 x 4000 steps: 201 GB of the 288 GB)
            try:
                adj["fp32c_all_states"] = time_adjoint(L.FP32C, args.adjoint_steps, args.adjoint_steps + 1)
            except L.LhyseaError as e:
                adj["fp32c_all_states"] = {"error": str(e)}
        if fp64:
            adj["fp64"] = time_adjoint(L.FP64, args.adjoint_steps, args.adjoint_snapshots)
        out["adjoint"] = adj
    if args.large_adjoint_steps > 0 and fp32c:
        out["adjoint_4096"] = time_large_adjoint(args.large_adjoint_steps)
    if args.ensemble_steps > 0 and fp32c:
        out["ensemble"] = time_ensemble(args.ensemble_steps)
    if args.ensemble_full_steps > 0 and fp32c:
        out["ensemble_full"] = time_ensemble(args.ensemble_full_steps, warmup=0)
    if args.precision_steps > 0 and fp32c and fp64:
        out["precision_2048"] = [precision_report(args.precision_steps, a) for a in (10.0, 2.0)]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
