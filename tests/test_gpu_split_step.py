"""FP32C step by tile class (GPU): tiles whose cells and halo hold no
sediment take the reduced kernel, the others the general tile function.

Bar: bit-exact.  The single-simulation step, with the split launch forced on
and forced off, must reproduce bit for bit the batched kernel (which runs the
general tile function on every tile) on grids of many tiles where slides
cross tile boundaries during the run, so tiles change class from step to
step.  They are checked on state, time, dt history, gauges and eta max.  The adjoint's re-advances after a snapshot
restore rebuild the classes; test_gpu_adjoint.py checks them bitwise against
the run that keeps every state.
"""
import numpy as np
import pytest

from tests.cases import case
from landslidehysea_amd import solver as L
from landslidehysea_amd.scenario import gauge_cells

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run_single(W, H, prm, steps, gauges, split):
    with L.Simulation(prm) as sim:
        sim.set_tile_split(split)
        sim.upload(W, H)
        sim.track_eta_max(True)
        sim.set_gauges(gauges, steps)
        classes0 = sim.tile_classes()
        sim.step(steps)
        return {"W": sim.download(), "time": sim.time(), "emax": sim.eta_max(), "dt": sim.dt_history(steps),
                "gauges": sim.gauges(), "classes": (classes0, sim.tile_classes())}


def run_batch(W, H, prm, steps, gauges):
    with L.Batch([prm]) as b:
        b.upload(0, W, H)
        b.track_eta_max(0, True)
        b.set_gauges(0, gauges, steps)
        b.initial_dt()
        b.step(steps)
        return {"W": b.download(0), "time": b.time(0), "emax": b.eta_max(0), "dt": b.dt_history(0, steps),
                "gauges": b.gauges(0)}


@pytest.mark.parametrize("nx,ny,seed,angle,steps", [(400, 300, 5, 1.0, 240), (257, 333, 6, 3.0, 200),
                                                      (1024, 512, 7, 1.0, 60)])
def test_split_step_equals_general_kernel(nx, ny, seed, angle, steps):
    sc, W, H, prm = case(nx, ny, seed=seed, angle=angle, n_slides=3, semi_axes=(14.0, 9.0), thickness=(30.0, 60.0),
                         precision=L.FP32C, borde=(1.0, -1.0, 1.0, -1.0))
    gauges = gauge_cells(sc, 8, seed=seed)
    tiles = ((nx + 61) // 62) * ((ny + 13) // 14)
    say(f"split step {nx}x{ny} seed={seed} steps={steps} tiles={tiles} sediment cells={(W[3] > 0).sum()}")
    b = run_batch(W, H, prm, steps, gauges)
    for split in (1, 0):
        a = run_single(W, H, prm, steps, gauges, split)
        (nt, g0, _), (_, g1, sp) = a["classes"]
        moved = int(((a["W"][3] > 0) & ~(W[3] > 0)).sum())
        say(f"split={split}: general tiles {g0} -> {g1} of {nt}, launches split {sp}, cells that gained "
            f"sediment {moved}, time {a['time']}")
        assert moved > 0 and 0 < g0 < nt and 0 < g1 < nt and sp == bool(split)
        assert np.array_equal(a["W"], b["W"])
        assert a["time"] == b["time"]
        assert np.array_equal(a["dt"], b["dt"])
        assert np.array_equal(a["gauges"], b["gauges"])
        assert np.array_equal(a["emax"][0], b["emax"][0]) and np.array_equal(a["emax"][1], b["emax"][1])


def test_split_step_without_sediment_and_all_sediment():
    """No tile general (lake over a bare bed with a water hump) and every tile
    general (sediment everywhere): both equal the general kernel bitwise."""
    nx, ny, steps = 300, 200, 80
    sc, W, H, prm = case(nx, ny, seed=8, angle=2.0, n_slides=1, semi_axes=(10.0, 8.0), precision=L.FP32C)
    W0 = W.copy()
    W0[3:] = 0.0
    W0[0] = np.maximum(W0[0], 0.0)
    jj, ii = np.mgrid[0:ny, 0:nx]
    W0[0] += 0.5 * np.exp(-(((ii - 150) / 12.0) ** 2 + ((jj - 100) / 9.0) ** 2)) * (W0[0] > 0)
    W1 = W.copy()
    W1[3] = np.maximum(W1[3], 0.01)
    gauges = gauge_cells(sc, 4, seed=8)
    for name, Ws, general in (("no sediment", W0, 0), ("sediment everywhere", W1, None)):
        b = run_batch(Ws, H, prm, steps, gauges)
        for split in (1, 0, -1):
            say("split step", name, "mode", split)
            a = run_single(Ws, H, prm, steps, gauges, split)
            (nt, g0, _), _ = a["classes"]
            assert g0 == (nt if general is None else general)
            assert np.array_equal(a["W"], b["W"]) and a["time"] == b["time"]
            assert np.array_equal(a["gauges"], b["gauges"])


@pytest.mark.parametrize("nx,ny,seed,angle,steps", [(400, 300, 5, 1.0, 160), (1024, 512, 7, 1.0, 40)])
def test_fp64_split_step_equals_fourpass(nx, ny, seed, angle, steps):
    """FP64: the split launch (forced on and off) equals the four-pass kernels
    (the reference's structure, HBM accumulators) bit for bit."""
    sc, W, H, prm = case(nx, ny, seed=seed, angle=angle, n_slides=3, semi_axes=(14.0, 9.0), thickness=(30.0, 60.0),
                         precision=L.FP64, borde=(1.0, -1.0, 1.0, -1.0))
    gauges = gauge_cells(sc, 8, seed=seed)

    def run(kernel, split):
        with L.Simulation(prm, kernel=kernel) as sim:
            if split is not None:
                sim.set_tile_split(split)
            sim.upload(W, H)
            sim.track_eta_max(True)
            sim.set_gauges(gauges, steps)
            sim.step(steps)
            return {"W": sim.download(), "time": sim.time(), "dt": sim.dt_history(steps), "gauges": sim.gauges(),
                    "emax": sim.eta_max(), "classes": sim.tile_classes()}

    ref = run(L.KERNEL_FOURPASS, None)
    for split in (1, 0):
        a = run(L.KERNEL_FUSED, split)
        nt, g, sp = a["classes"]
        say(f"fp64 {nx}x{ny} split={split}: general tiles {g} of {nt}, launches split {sp}")
        assert 0 < g < nt and sp == bool(split)
        assert np.array_equal(a["W"], ref["W"]) and a["time"] == ref["time"]
        assert np.array_equal(a["dt"], ref["dt"]) and np.array_equal(a["gauges"], ref["gauges"])
        assert np.array_equal(a["emax"][0], ref["emax"][0]) and np.array_equal(a["emax"][1], ref["emax"][1])


@pytest.mark.parametrize("precision", [L.FP32C, L.FP64])
def test_split_step_run_until(precision):
    """run_until (device-side stop inside launched chunks) with the split launch
    forced on equals the all-tiles launch bit for bit, in both precisions."""
    sc, W, H, prm = case(400, 300, seed=9, angle=1.0, n_slides=2, semi_axes=(14.0, 9.0), thickness=(30.0, 60.0),
                         precision=precision)
    out = {}
    for split in (1, 0):
        with L.Simulation(prm) as sim:
            sim.set_tile_split(split)
            sim.upload(W, H)
            n1 = sim.run_until(2.5)
            n2 = sim.run_until(4.0, 37)
            out[split] = (sim.download(), sim.time(), n1, n2)
    a, b = out[1], out[0]
    say("run_until steps", a[2], a[3], "time", a[1])
    assert a[2] == b[2] and a[3] == b[3] and a[1] == b[1]
    assert np.array_equal(a[0], b[0])


@pytest.mark.parametrize("precision", [L.FP32C, L.FP64])
def test_reupload_over_sediment_equals_fresh_handle(precision):
    """The sediment-free launch leaves the layer-2 planes of a tile unwritten
    once its own zeros are in the output buffer (the tile was sediment-free two
    steps back, after the last classification).  A second upload into a handle
    whose buffers hold sediment everywhere must still match a fresh handle bit
    for bit, and so must a long run where slides leave tiles behind."""
    sc, W, H, prm = case(400, 300, seed=5, angle=1.0, n_slides=3, semi_axes=(14.0, 9.0), thickness=(30.0, 60.0),
                         precision=precision, borde=(1.0, -1.0, 1.0, -1.0))
    Wsed = W.copy()
    Wsed[3] = np.maximum(Wsed[3], 0.01)
    outs = []
    for reuse in (True, False):
        with L.Simulation(prm) as sim:
            sim.set_tile_split(1)
            if reuse:
                sim.upload(Wsed, H)
                sim.step(5)
            sim.upload(W, H)
            sim.step(120)
            outs.append((sim.download(), sim.time()))
    say(f"reupload precision={precision}: time {outs[0][1]}")
    assert np.array_equal(outs[0][0], outs[1][0]) and outs[0][1] == outs[1][1]
