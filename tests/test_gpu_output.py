"""lhysea_shallow_water_output on the GPU: shallowWater with its output files.

The run's records must be the states a caller of the handle API sees at the
reference's save points (src/GPU/ShallowWater.cu:269-334: at the top of the
loop, whenever t >= the next save time, which then advances by
tiempo_guardar), written as the output layout prescribes (checked on the CPU
in test_output.py): NetCDF records bit for bit equal to the float32 of those
states, the record times, max_height from eta max, the final state, and the
gauge text lines; per-step stepping (verbose) and stepping to each save time
(run_until) give the same files; FP64 and FP32C.
"""
from dataclasses import replace

import numpy as np
import pytest

from tests.cases import case
from landslidehysea_amd import solver as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def handle_run(prm, W, H, t_end, save):
    """The reference's loop through the handle API: saves, then one step."""
    recs = []
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.track_eta_max(True)
        sim.initial_dt()
        t, ns = 0.0, 0.0
        while t < t_end:
            if save >= 0.0 and t >= ns:
                recs.append((sim.download(), t))
                ns += save
            sim.step(1)
            t = sim.time()[0]
        em, _ = sim.eta_max()
        return recs, em, sim.download(), sim.time()[2]


@pytest.mark.parametrize("precision", [L.FP64, L.FP32C])
def test_netcdf_run_equals_handle_stepping(tmp_path, precision):
    from scipy.io import netcdf_file
    sc, W, H, prm = case(64, 48, angle=2.0, semi_axes=(12.0, 9.0))
    prm = replace(prm, precision=precision)
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        dt0 = sim.initial_dt()
    save, t_end = 7.3 * dt0, 60.5 * dt0
    hm = -0.5
    recs, em, Wf, nsteps = handle_run(prm, W, H, t_end, save)
    say(f"precision {precision}: {nsteps} steps, {len(recs)} records")
    files = {}
    for verbose in (0, 1):
        prefix = str(tmp_path / f"v{verbose}")
        Wo, secs, n = L.shallow_water_output(prm, W, H, t_end, prefix, verbose=bool(verbose), save_interval=save,
                                             Hmin=hm, Q=2.0, T=3.0, sim_time=t_end)
        assert n == nsteps and np.array_equal(Wo, Wf)
        with netcdf_file(prefix + "_eta1.nc", "r", mmap=False) as f:
            eta = f.variables["eta1"][:].copy()
            times = f.variables["time"][:].copy()
            mh = f.variables["max_height"][:].copy()
        with netcdf_file(prefix + "_q2x.nc", "r", mmap=False) as f:
            q2x = f.variables["q2x"][:].copy()
        files[verbose] = (eta, times, mh, q2x)
        assert eta.shape[0] == len(recs)
        for r, (Wr, t) in enumerate(recs):
            assert np.array_equal(eta[r], ((Wr[0] + Wr[3] - H - hm) * prm.H).astype(np.float32)), r
            assert np.array_equal(q2x[r], (Wr[4] * 2.0).astype(np.float32)), r
            assert times[r] == np.float32(t * 3.0)
        assert np.array_equal(mh, ((em - hm) * prm.H).astype(np.float32))
    for a, b in zip(files[0], files[1]):
        assert np.array_equal(a, b)


def test_gauge_run_equals_handle_stepping(tmp_path):
    sc, W, H, prm = case(64, 48, angle=2.0, semi_axes=(12.0, 9.0))
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        dt0 = sim.initial_dt()
    save, t_end = 5.0 * dt0, 40.5 * dt0
    gauges = np.array([5, -1, 64 * 20 + 30, 64 * 48 - 1], dtype=np.int32)
    recs, em, Wf, nsteps = handle_run(prm, W, H, t_end, save)
    prefix = str(tmp_path / "g")
    Wo, secs, n = L.shallow_water_output(prm, W, H, t_end, prefix, save_interval=save, gauges=gauges)
    assert n == nsteps and np.array_equal(Wo, Wf)
    lines = open(prefix + "_eta_puntos.txt").read().splitlines()
    assert len(lines) == len(recs)
    Hf = H.ravel()
    for line, (Wr, t) in zip(lines, recs):
        h1 = Wr[0].ravel()
        want = ["%e" % t] + [("%.8e" % ((h1[g] - Hf[g]) * prm.H)) if g >= 0 else "-999" for g in gauges]
        assert line.split() == want
