"""The library's kernel sources run on the CPU under an emulation of HIP.

tests/native/emu holds a host-side stand-in for <hip/hip_runtime.h> and its
runtime: the .hip sources compile with g++ into a library with the same C ABI,
whose kernels run workgroup by workgroup in index order (and, in a second
pass, in reverse index order), the lanes of a workgroup as cooperative
fibers.  Those orders are the most adverse ones for the protocols between the
workgroups of one launch (the general-tile launch that closes a step, the
class lists written by one step and read by the next): a workgroup starts
only after every one before it in the order has finished, the way a
workgroup that the GPU schedules late would.  A general-tile launch whose late
workgroups read the next step's tile list instead of leaving fails the FP32C
check on the 64 x 48 grid (whose general set grows) and the FP64 check by
metres of eta.

Checks, bit for bit between paths of the library on the same inputs:
  * FP32C split step (sediment-free and general launches) == every tile
    through the general tile function == the batched kernel, on a ragged grid
    of several class tiles with slides that cross tile edges;
  * FP64 split step (forced on and off) == the four-pass kernels;
  * run_until with the split forced on == the all-tiles launch, both precisions;
  * a second upload into a handle whose buffers hold sediment everywhere ==
    a fresh handle, both precisions (the sediment-free launch rewrites the
    layer-2 planes until its own writes are two steps back);
and against the FP64 oracle (C restatement): the emulated FP64 step within
1e-10 relative L2 over the state, and the discrete adjoint (FP64 four-pass and
FP32C fused reverse steps) against the oracle's central differences with the
dt sequence frozen (1e-4 and 2e-2, as on the GPU), under the Coulomb law and
under the Pouliquen law (dL/dangle of each of its four angles), with binomial
checkpoints (restores and class rebuilds in workgroup order) equal to keeping
every state, and the FP32C fused reverse kernel equal to the four-pass reverse
step to rounding (1e-6 of the largest gradient).

The emulated library is built with -ffp-contract=off, so its arithmetic is not
that of the gfx950 code object; the GPU tests repeat these checks on the
device.  The cases run in a child process so the emulated library never shares
a process with the HIP one.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "landslidehysea_amd", "csrc")
SOURCES = ["lhysea_capi.hip", "lhysea_batch.hip", "lhysea_fp64.hip", "lhysea_fp32c.hip", "lhysea_adjoint.hip",
           "lhysea_noalloc.hip", "host/lhysea_output.cpp", "host/lhysea_netcdf.cpp"]


def build_emulated(out: str):
    cmd = ["g++", "-std=c++20", "-O1", "-fPIC", "-shared", "-ffp-contract=off", "-w",
           "-I", os.path.join(HERE, "native", "emu"), "-x", "c++"]
    cmd += [os.path.join(CSRC, s) for s in SOURCES]
    cmd += ["-x", "none", os.path.join(HERE, "native", "emu", "emu_rt.cpp"), "-o", out]
    subprocess.run(cmd, check=True)


def adjoint_case(L, ob, case, gauge_cells, oracle_params, prec, pouliquen):
    """Gauge misfit over 12 steps of a flowing 24 x 18 slide against a run from
    a 10% thicker slide: the adjoint's dL/dW0 at smooth cells and dL/dangle
    against central differences of the oracle with the dt sequence frozen."""
    from dataclasses import replace
    nsteps = 12
    sc, W, H, prm = case(24, 18, angle=2.0, semi_axes=(5.0, 4.0), precision=prec)
    if pouliquen:
        prm = replace(prm, friction_law=1, angle_deg=(2.0, 6.0, 3.0, 8.0))
    gauges = gauge_cells(sc, 4, seed=3)
    po = oracle_params(ob, prm)
    Wt = W.copy()
    Wt[3] *= 1.1
    Wt[0] = np.maximum(0.0, (W[0] + W[3]) - Wt[3])
    obs = ob.run(po, Wt, H, nsteps, gauges=gauges)["gauges"]
    dts = ob.run(po, W.copy(), H, nsteps, gauges=gauges)["dt_hist"]
    out = []
    for nsnap in (4, nsteps + 1):
        with L.Simulation(prm) as sim:
            sim.upload(W, H)
            out.append(sim.adjoint(nsteps, nsnap, gauges, obs))
    r = out[0]
    fourpass_rel = None
    if prec == L.FP32C:
        # the four-pass form of the reverse step against the fused kernel
        with L.Simulation(prm) as sim:
            sim.set_adjoint_kernel(L.KERNEL_FOURPASS)
            sim.upload(W, H)
            r4 = sim.adjoint(nsteps, 4, gauges, obs)
        gmax = np.abs(r["grad_W0"]).max()
        fourpass_rel = max(float(np.abs(r4["grad_W0"] - r["grad_W0"]).max() / gmax),
                           float(np.abs(r4["grad_angles"] - r["grad_angles"]).max() / np.abs(r["grad_angles"]).max()))

    def loss(Wx, p=po):
        g = ob.run_fixed_dt(p, Wx.copy(), H, dts, gauges)
        return float(((g - obs) ** 2).sum())

    L0 = loss(W)
    rels, smooth = [], 0
    rng = np.random.default_rng(0)
    for c in rng.choice(np.flatnonzero(W[3].ravel() > 5.0), 5, replace=False):
        j, i = divmod(int(c), prm.nx)
        for k in (0, 3):
            eps = 1e-5 * max(1.0, W[k, j, i])
            Wp, Wm = W.copy(), W.copy()
            Wp[k, j, i] += eps
            Wm[k, j, i] -= eps
            lp, lm = loss(Wp), loss(Wm)
            fp, fm, fc = (lp - L0) / eps, (L0 - lm) / eps, (lp - lm) / (2 * eps)
            if abs(fp - fm) <= 1e-3 * abs(fc):
                smooth += 1
                rels.append(abs(r["grad_W0"][k, j, i] - fc) / abs(fc))
    # dL/dangle: Coulomb perturbs the four angles together (only angle1 enters),
    # Pouliquen each angle alone
    names = ("angle1", "angle2", "angle3", "angle4")
    fa = []
    for which in ((names,) if not pouliquen else tuple((a,) for a in names)):
        pp, pm = oracle_params(ob, prm), oracle_params(ob, prm)
        for q, sgn in ((pp, 1), (pm, -1)):
            for a in which:
                setattr(q, a, getattr(q, a) + sgn * 1e-6)
        fa.append((loss(W, pp) - loss(W, pm)) / 2e-6)
    ga = [r["grad_angle1"]] if not pouliquen else [float(x) for x in r["grad_angles"]]
    big = max(abs(x) for x in fa)
    angle_rel = max(abs(g - f) / abs(f) for g, f in zip(ga, fa) if abs(f) >= 1e-3 * big)
    return {"loss_rel": abs(r["loss"] - L0) / L0, "smooth": smooth, "max_rel": max(rels) if rels else None,
            "angle_rel": angle_rel, "angles_fd": fa, "angles_adj": ga, "advances": r["advances"],
            "fourpass_rel": fourpass_rel,
            "checkpoints_equal_all_states": bool(np.array_equal(out[0]["grad_W0"], out[1]["grad_W0"])
                                                 and np.array_equal(out[0]["grad_angles"], out[1]["grad_angles"]))}


def run_cases(lib: str) -> dict:
    """Child process: every case on the emulated library; returns the verdicts."""
    sys.path.insert(0, ROOT)
    from landslidehysea_amd import solver as L
    from landslidehysea_amd.scenario import gauge_cells
    from tests.cases import case, oracle_params, rel_l2
    from oracle import binding as ob

    L.LIB_PATH = lib
    L._lib = None
    res = {}

    def single(W, H, prm, steps, gauges, split, kernel=L.KERNEL_FUSED):
        with L.Simulation(prm, kernel=kernel) as sim:
            if split is not None:
                sim.set_tile_split(split)
            sim.upload(W, H)
            sim.track_eta_max(True)
            sim.set_gauges(gauges, steps)
            g0 = sim.tile_classes()
            sim.step(steps)
            return {"W": sim.download(), "time": sim.time(), "dt": sim.dt_history(steps), "gauges": sim.gauges(),
                    "emax": sim.eta_max(), "classes": (g0, sim.tile_classes())}

    def same(a, b):
        return bool(np.array_equal(a["W"], b["W"]) and a["time"] == b["time"] and np.array_equal(a["dt"], b["dt"])
                    and np.array_equal(a["gauges"], b["gauges"]) and np.array_equal(a["emax"][0], b["emax"][0])
                    and np.array_equal(a["emax"][1], b["emax"][1]))

    # FP32C: 157 x 61 cells = 3 x 5 class tiles of 62 x 14, slides across tile
    # edges; 64 x 48 cells = 2 x 4 class tiles, where the general set grows
    for name, nx, ny, seed, angle, axes, steps in (("a", 157, 61, 5, 1.0, (12.0, 7.0), 12),
                                                  ("b", 64, 48, 1, 2.0, (12.0, 9.0), 8)):
        sc, W, H, prm = case(nx, ny, seed=seed, angle=angle, n_slides=3 if name == "a" else 1, semi_axes=axes,
                             thickness=(30.0, 60.0) if name == "a" else (40.0, 40.0), precision=L.FP32C,
                             borde=(1.0, -1.0, 1.0, -1.0) if name == "a" else (-1.0, -1.0, -1.0, -1.0))
        gauges = gauge_cells(sc, 6, seed=seed)
        with L.Batch([prm]) as b:
            b.upload(0, W, H)
            b.track_eta_max(0, True)
            b.set_gauges(0, gauges, steps)
            b.initial_dt()
            b.step(steps)
            ref = {"W": b.download(0), "time": b.time(0), "emax": b.eta_max(0), "dt": b.dt_history(0, steps),
                   "gauges": b.gauges(0)}
        for split in (1, 0):
            a = single(W, H, prm, steps, gauges, split)
            (nt, g0, _), (_, g1, sp) = a["classes"]
            res[f"fp32c_{name}_split{split}"] = {"same": same(a, ref), "tiles": nt, "general": [g0, g1],
                                                 "split": sp, "steps": a["time"][2], "want_steps": steps}

    # a second upload into a handle whose buffers hold sediment everywhere:
    # the first steps after it must rewrite the layer-2 planes of the
    # sediment-free tiles (the other buffer holds the old run's state)
    for prec in (L.FP32C, L.FP64):
        sc, W, H, prm = case(157, 61, seed=5, angle=1.0, n_slides=3, semi_axes=(12.0, 7.0),
                             thickness=(30.0, 60.0), precision=prec)
        Wsed = W.copy()
        Wsed[3] = np.maximum(Wsed[3], 0.01)
        outs = []
        for reuse in (True, False):
            with L.Simulation(prm) as sim:
                sim.set_tile_split(1)
                if reuse:
                    sim.upload(Wsed, H)
                    sim.step(3)
                sim.upload(W, H)
                sim.step(6)
                outs.append((sim.download(), sim.time()))
        res[f"reupload_{prec}"] = {"same": bool(np.array_equal(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1])}

    # FP64: split forced on and off against the four-pass kernels
    sc, W, H, prm = case(131, 57, seed=6, angle=1.0, n_slides=3, semi_axes=(12.0, 7.0), thickness=(30.0, 60.0),
                         borde=(1.0, -1.0, 1.0, -1.0))
    gauges = gauge_cells(sc, 6, seed=6)
    ref = single(W, H, prm, steps, gauges, None, kernel=L.KERNEL_FOURPASS)
    for split in (1, 0):
        a = single(W, H, prm, steps, gauges, split)
        nt, g, sp = a["classes"][1]
        res[f"fp64_split{split}"] = {"same": same(a, ref), "tiles": nt, "general": g, "split": sp}
    Wo = W.copy()
    ob.run(oracle_params(ob, prm), Wo, H, steps)
    res["fp64_oracle"] = {"rel_l2": rel_l2(a["W"], Wo)}

    # run_until with the split forced on and off
    for prec in (L.FP32C, L.FP64):
        sc, W, H, prm = case(131, 57, seed=9, angle=1.0, n_slides=2, semi_axes=(12.0, 7.0), thickness=(30.0, 60.0),
                             precision=prec)
        out = {}
        for split in (1, 0):
            with L.Simulation(prm) as sim:
                sim.set_tile_split(split)
                sim.upload(W, H)
                n1 = sim.run_until(3.0)
                n2 = sim.run_until(6.0, 7)
                out[split] = (sim.download(), sim.time(), n1, n2)
        a, b = out[1], out[0]
        res[f"run_until_{prec}"] = {"same": bool(np.array_equal(a[0], b[0]) and a[1:] == b[1:]),
                                    "steps": [a[2], a[3]]}

    # discrete adjoint (FP32C: the fused reverse kernel; FP64: the four-pass
    # reverse step) against central differences of the oracle with the dt
    # sequence frozen, and binomial checkpointing against keeping every state;
    # under the Coulomb law and under the Pouliquen law (the gradients of its
    # four angles one by one)
    for prec in (L.FP64, L.FP32C):
        res[f"adjoint_{prec}"] = adjoint_case(L, ob, case, gauge_cells, oracle_params, prec, False)
        res[f"adjoint_pq_{prec}"] = adjoint_case(L, ob, case, gauge_cells, oracle_params, prec, True)
    return res


@pytest.fixture(scope="module")
def emulated_lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    lib = str(tmp_path_factory.mktemp("emu") / "liblhysea_emu.so")
    build_emulated(lib)
    return lib


def cases_in_order(lib, order):
    env = dict(os.environ, EMU_ORDER=order)
    out = subprocess.run([sys.executable, __file__, lib], capture_output=True, text=True, timeout=600, cwd=ROOT,
                         env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    res = json.loads(out.stdout.strip().splitlines()[-1])
    print(order, json.dumps(res, indent=1))
    return res


# every case with the workgroups of a launch in index order and in reverse
@pytest.fixture(scope="module", params=["forward", "reverse"])
def emulated(emulated_lib, request):
    return cases_in_order(emulated_lib, request.param)


def test_fp32c_split_step_in_workgroup_order(emulated):
    for name, tiles in (("a", 15), ("b", 8)):
        for split in (1, 0):
            r = emulated[f"fp32c_{name}_split{split}"]
            assert r["same"], r
            assert r["tiles"] == tiles and r["split"] == bool(split) and r["steps"] == r["want_steps"]
            assert all(0 < g < r["tiles"] for g in r["general"])
    # the general set grows during the 64 x 48 run: the step after lists more tiles
    assert emulated["fp32c_b_split1"]["general"][1] > emulated["fp32c_b_split1"]["general"][0]


def test_fp64_split_step_in_workgroup_order(emulated):
    for split in (1, 0):
        r = emulated[f"fp64_split{split}"]
        assert r["same"], r
        assert 0 < r["general"] < r["tiles"] and r["split"] == bool(split)


def test_reupload_rewrites_layer2_in_workgroup_order(emulated):
    for prec in (1, 0):
        assert emulated[f"reupload_{prec}"]["same"]


def test_emulated_fp64_step_matches_oracle(emulated):
    assert emulated["fp64_oracle"]["rel_l2"] <= 1e-10


def test_run_until_split_in_workgroup_order(emulated):
    for prec in (1, 0):
        r = emulated[f"run_until_{prec}"]
        assert r["same"] and r["steps"][0] > 5 and r["steps"][1] == 7, r


@pytest.mark.parametrize("law", ["", "pq_"])
@pytest.mark.parametrize("prec,tol", [(0, 1e-4), (1, 2e-2)])
def test_adjoint_in_workgroup_order_matches_oracle_fd(emulated, prec, tol, law):
    r = emulated[f"adjoint_{law}{prec}"]
    assert r["loss_rel"] <= (1e-9 if prec == 0 else 5e-2)
    assert r["smooth"] >= 6 and r["max_rel"] <= tol, r
    assert r["angle_rel"] <= tol, r
    assert r["advances"] > 0 and r["checkpoints_equal_all_states"], r
    if law:
        # every angle of the law moves the loss on this flowing slide
        assert len(r["angles_fd"]) == 4 and min(abs(x) for x in r["angles_fd"]) > 0, r
    if prec == 1:
        # FP32C: the fused reverse kernel equals the four-pass reverse step to rounding
        assert r["fourpass_rel"] <= 1e-6, r


if __name__ == "__main__":
    print(json.dumps(run_cases(sys.argv[1])))
