"""Failure detection: the health word of the step kernels and forced allocation failures.

The reference has no NaN checks and returns 1 when a device allocation fails
(src/GPU/ShallowWater.cu:49-57, :124).  Here every cell update of the step
kernels checks its new state and raises LHYSEA_HEALTH_NONFINITE (a NaN or
infinity) or LHYSEA_HEALTH_NEGATIVE_DEPTH (a layer depth below -eps_h after
the flux update, before the clip to zero) in the control block, with the
first step that raised one; lhysea_step and lhysea_run_until return
LHYSEA_ERR_NUMERIC once they see it.  The faults are injected from the host:

  * healthy runs (FP64 fused and four-pass, FP32C split and all-tiles) leave
    the word clear;
  * a NaN written into one wet cell's discharge is flagged at step 0; the next
    lhysea_step and lhysea_run_until return LHYSEA_ERR_NUMERIC; a new upload
    clears the word;
  * a depth of -2 eps_h on dry land (no wet neighbour, so every edge of the
    cell is dry and the update leaves the depth as it is) raises only the
    negative-depth bit, and the clip leaves a valid state;
  * in a batch, the scenario holding the NaN is flagged and the others are
    not, and equal the same scenarios stepped without it, bit for bit;
  * lhysea_step_device returns LHYSEA_ERR_NUMERIC with its outputs set;
  * the one-shot driver returns LHYSEA_ERR_NUMERIC;
  * lhysea_create returns LHYSEA_ERR_GPU_MEMORY when the device allocation
    fails: under the CPU emulation of HIP by an allocation cap
    (EMU_MEM_CAP), on the GPU with a grid (131072^2 cells, 0.8 TB of state)
    beyond its memory.

The same cases run on the CPU, on the library's kernel sources compiled
under the emulation of HIP (tests/native/emu, workgroups in reverse index
order), and on the MI355X.  None of the injected values steers an index
or a loop bound in a kernel: they only flow through the arithmetic.
"""
from __future__ import annotations

import json
import os
import shutil
import subprocess
import sys
from dataclasses import replace

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
NONFINITE, NEGATIVE = 1, 2


def wet_cell_near_slide(W):
    """A wet cell (h1 > 1 m) next to the slide's edge, row-major (j, i)."""
    j, i = np.argwhere((W[3] > 0.0))[0]
    for dj in range(0, 40):
        jj = j - dj
        if jj >= 0 and W[0, jj, i] > 1.0 and W[3, jj, i] == 0.0:
            return int(jj), int(i)
    raise AssertionError("no wet cell found")


def dry_island_cell(W):
    """A dry cell (h1 = h2 = 0) whose 8 neighbours are dry too."""
    dry = (W[0] == 0.0) & (W[3] == 0.0)
    inner = dry[1:-1, 1:-1] & dry[:-2, 1:-1] & dry[2:, 1:-1] & dry[1:-1, :-2] & dry[1:-1, 2:] \
        & dry[:-2, :-2] & dry[:-2, 2:] & dry[2:, :-2] & dry[2:, 2:]
    j, i = np.argwhere(inner)[0]
    return int(j) + 1, int(i) + 1


def health_cases(L, device_buffers) -> dict:
    """Every case on the library bound as L; device_buffers(arrays) returns
    (device addresses, read-back function, keep-alive) for lhysea_step_device."""
    from tests.cases import case
    res = {}
    sc, W, H, p64 = case(64, 48, angle=2.0, semi_axes=(12.0, 9.0))
    wj, wi = wet_cell_near_slide(W)
    dj, di = dry_island_cell(W)
    res["cells"] = [wj, wi, dj, di]
    Wnan = W.copy()
    Wnan[1, wj, wi] = np.nan
    Wneg = W.copy()
    Wneg[0, dj, di] = -2.0 * p64.eps_h

    variants = [("fp64_fused", p64, L.KERNEL_FUSED, None), ("fp64_fourpass", p64, L.KERNEL_FOURPASS, None),
                ("fp32c_split", replace(p64, precision=L.FP32C), L.KERNEL_FUSED, 1),
                ("fp32c_all", replace(p64, precision=L.FP32C), L.KERNEL_FUSED, 0)]
    for name, prm, kernel, split in variants:
        r = {}
        with L.Simulation(prm, kernel=kernel) as sim:
            if split is not None:
                sim.set_tile_split(split)
            sim.upload(W, H)
            sim.step(20)
            r["healthy"] = sim.health()
            sim.upload(Wnan, H)
            sim.step(5)
            r["nan"] = sim.health()
            try:
                sim.step(1)
                r["next_step"] = 0
            except L.LhyseaError as e:
                r["next_step"] = e.code
            try:
                sim.run_until(1e9, 50)
                r["run_until"] = 0
            except L.LhyseaError as e:
                r["run_until"] = e.code
            sim.upload(W, H)
            r["after_upload"] = sim.health()
            sim.run_until(1e9, 10)
            sim.upload(Wneg, H)
            sim.step(3)
            r["negative"] = sim.health()
            Wr = sim.download()
            r["negative_clipped"] = bool(Wr[0, dj, di] == 0.0 and np.isfinite(Wr).all())
        res[name] = r

    # batch: NaN in scenario 1 of 3
    p32 = replace(p64, precision=L.FP32C)
    outs = {}
    for tag, bad in (("with_nan", True), ("clean", False)):
        with L.Batch([p32] * 3) as b:
            for sid in range(3):
                b.upload(sid, Wnan if (bad and sid == 1) else W, H)
            b.step(8)
            outs[tag] = ([b.health(sid) for sid in range(3)], [b.download(sid) for sid in (0, 2)])
    res["batch"] = {"health": outs["with_nan"][0], "clean_health": outs["clean"][0],
                    "others_equal": bool(all(np.array_equal(a, c) for a, c in
                                             zip(outs["with_nan"][1], outs["clean"][1])))}

    # no-allocation entry point
    r = {}
    for prec in (L.FP64, L.FP32C):
        prm = replace(p64, precision=prec)
        (hd, wd, wk), _, keep = device_buffers(prm, H, Wnan)
        try:
            L.step_device(prm, hd, wd, wk, 4)
            r[str(prec)] = 0
        except L.LhyseaError as e:
            r[str(prec)] = e.code
        del keep
    res["step_device"] = r

    # one-shot driver
    try:
        L.shallow_water(p64, Wnan, H, t_end=1e9)
        res["shallow_water"] = 0
    except L.LhyseaError as e:
        res["shallow_water"] = e.code
    return res


def check(res):
    for name in ("fp64_fused", "fp64_fourpass", "fp32c_split", "fp32c_all"):
        r = res[name]
        assert r["healthy"] == [0, -1] or r["healthy"] == (0, -1), (name, r)
        assert r["nan"][0] & NONFINITE and r["nan"][1] == 0, (name, r)
        assert r["next_step"] == 7 and r["run_until"] == 7, (name, r)
        assert tuple(r["after_upload"]) == (0, -1), (name, r)
        assert tuple(r["negative"]) == (NEGATIVE, 0), (name, r)
        assert r["negative_clipped"], (name, r)
    b = res["batch"]
    assert [tuple(h) for h in b["clean_health"]] == [(0, -1)] * 3, b
    assert tuple(b["health"][0]) == (0, -1) and tuple(b["health"][2]) == (0, -1), b
    assert b["health"][1][0] & NONFINITE and b["health"][1][1] == 0, b
    assert b["others_equal"], b
    assert res["step_device"] == {"0": 7, "1": 7}, res["step_device"]
    assert res["shallow_water"] == 7


def host_buffers(prm, H, W):
    """Device memory is host memory under the emulation: numpy buffers."""
    from landslidehysea_amd import solver as L
    Hd = np.ascontiguousarray(H)
    if prm.precision == L.FP64:
        Wd = np.ascontiguousarray(W.reshape(6, -1)).copy()
    else:
        hi = W.astype(np.float32)
        lo = (W - hi.astype(np.float64)).astype(np.float32)
        Wd = np.ascontiguousarray(np.concatenate([hi.reshape(6, -1), lo.reshape(6, -1)]))
    Wk = np.empty_like(Wd)
    return (Hd.ctypes.data, Wd.ctypes.data, Wk.ctypes.data), None, (Hd, Wd, Wk)


def emulated_child(lib: str, cap: bool) -> dict:
    sys.path.insert(0, ROOT)
    from landslidehysea_amd import solver as L
    from tests.cases import case
    L.LIB_PATH = lib
    L._lib = None
    if cap:
        sc, W, H, prm = case(64, 48)
        out = {}
        for prec in (L.FP64, L.FP32C):
            try:
                with L.Simulation(replace(prm, precision=prec)):
                    out[str(prec)] = 0
            except L.LhyseaError as e:
                out[str(prec)] = e.code
        return out
    return health_cases(L, host_buffers)


@pytest.fixture(scope="module")
def emulated_lib(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    from tests.test_emulated_kernels import build_emulated
    lib = str(tmp_path_factory.mktemp("emu_health") / "liblhysea_emu.so")
    build_emulated(lib)
    return lib


def run_child(lib, env_extra, cap=False):
    env = dict(os.environ, **env_extra)
    out = subprocess.run([sys.executable, __file__, lib, "cap" if cap else "cases"], capture_output=True, text=True,
                         timeout=600, cwd=ROOT, env=env)
    assert out.returncode == 0, out.stdout + out.stderr
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_health_word_under_emulation(emulated_lib):
    # workgroups from the last one down (the protocols between workgroups are
    # checked in both orders by test_emulated_kernels.py)
    res = run_child(emulated_lib, {"EMU_ORDER": "reverse"})
    print(json.dumps(res))
    check(res)


def test_create_reports_device_memory_exhaustion_under_emulation(emulated_lib):
    # 64 x 48 FP64 needs 6 * 3072 * 8 = 147 kB per state buffer: a 100 kB cap refuses it
    res = run_child(emulated_lib, {"EMU_MEM_CAP": "100000"}, cap=True)
    assert res == {"0": 1, "1": 1}, res


@pytest.mark.gpu
def test_health_word_on_gpu():
    import torch
    from landslidehysea_amd import solver as L

    def torch_buffers(prm, H, W):
        dev = torch.device("cuda:0")
        (_, _, _), _, (Hh, Wh, _) = host_buffers(prm, H, W)
        Hd = torch.from_numpy(Hh).to(dev)
        Wd = torch.from_numpy(Wh).to(dev)
        Wk = torch.empty_like(Wd)
        torch.cuda.synchronize()
        return (Hd.data_ptr(), Wd.data_ptr(), Wk.data_ptr()), None, (Hd, Wd, Wk)

    print("health cases on the GPU: NaN / negative depth injected into a 64x48 case", flush=True)
    res = health_cases(L, torch_buffers)
    torch.cuda.synchronize()
    print(json.dumps(res), flush=True)
    check(res)


@pytest.mark.gpu
def test_create_reports_device_memory_exhaustion_on_gpu():
    from landslidehysea_amd import solver as L
    from tests.cases import case
    sc, W, H, prm = case(64, 48)
    for prec in (L.FP64, L.FP32C):
        big = replace(prm, nx=131072, ny=131072, precision=prec)
        with pytest.raises(L.LhyseaError) as e:
            L.Simulation(big)
        print(f"precision {prec}: {e.value}", flush=True)
        assert e.value.code == L.ERR_GPU_MEMORY
    with L.Simulation(prm) as sim:              # the device is usable afterwards
        sim.upload(W, H)
        sim.step(5)
        assert sim.health() == (0, -1)


if __name__ == "__main__":
    print(json.dumps(emulated_child(sys.argv[1], sys.argv[2] == "cap")))
