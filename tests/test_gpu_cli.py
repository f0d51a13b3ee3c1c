"""L-HySEA.exe end to end on the GPU: the reference's command-line contract
(data.dat, .bat/.ini in, six NetCDF files or the gauge text out;
src/GPU/main.cxx:62-182, src/GPU/ShallowWater.cu:211-334, 462-474) run for
real, checked against the library stepped directly from Python with the
parameters the executable parsed (its --dump output).

NetCDF stores float32, so frames are compared to 1e-4 m (the float32 spacing
of the values is ~3e-5 m); saved times and frame counts exactly.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from landslidehysea_amd import scenario as S
from landslidehysea_amd import solver as L
from landslidehysea_amd.case_io import write_case

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "landslidehysea_amd", "L-HySEA.exe")


def say(*a):
    print(*a, flush=True)


def cli(args, cwd):
    if not os.path.exists(EXE):
        pytest.fail("L-HySEA.exe is not built (run __graft_entry__.build())")
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


def params_of(dump, precision=L.FP64):
    law = dump.get("friction_law", 0)
    return L.Params(nx=dump["nx"], ny=dump["ny"], dx=dump["dx"], dy=dump["dy"], borde=tuple(dump["borde"]),
                    cfl=dump["cfl"], r=dump["r"],
                    angle_deg=tuple(dump["angles_deg"]) if law == 1 else (dump["angle1_deg"],), mfc=dump["mfc"],
                    mf0=dump["mf0"], mfs=dump["mfs"], vmax1=dump["vmax1"], vmax2=dump["vmax2"], g=dump["g"],
                    eps_h=dump["eps_h"], L=dump["L"], H=dump["H"], friction_law=law, precision=precision)


def library_frames(prm, W, H, t_end, t_save):
    """The executable's loop (save when t >= next save time, then one step)
    through the Python mirror of the C-ABI."""
    frames, times = [], []
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.track_eta_max(True)
        sim.initial_dt()
        t, nxt = 0.0, 0.0
        while t < t_end:
            if t >= nxt:
                frames.append(sim.download())
                times.append(t)
                nxt += t_save
            sim.step(1)
            t = sim.time()[0]
        return frames, times, sim.eta_max()[0], sim.download()


# --quiet steps to each save time in one call (run_until); without it the
# executable steps one at a time and prints a line per step; a save interval
# shorter than dt saves after every step
# --pouliquen: the Pouliquen law with four angles read from data.dat
@pytest.mark.parametrize("precision,t_end,t_save,quiet,pouliquen",
                         [("fp64", 12.0, 3.0, True, False), ("fp32c", 12.0, 3.0, True, False),
                          ("fp64", 2.0, 0.05, True, False), ("fp32c", 12.0, 3.0, False, False),
                          ("fp64", 12.0, 3.0, True, True), ("fp32c", 12.0, 3.0, True, True)])
def test_cli_netcdf_run_matches_library(tmp_path, precision, t_end, t_save, quiet, pouliquen):
    from scipy.io import netcdf_file
    sc = S.make_scenario(64, 48, seed=1, semi_axes=(12.0, 9.0))
    law = ["--pouliquen"] if pouliquen else []
    write_case(str(tmp_path), sc, prefix="res", t_end=t_end, t_save=t_save,
               angle_deg=(2.0, 6.0, 3.0, 8.0) if pouliquen else 2.0)
    dump = json.loads(cli(["data.dat", "--dump"] + law, tmp_path).stdout)
    assert dump["friction_law"] == (1 if pouliquen else 0)
    r = cli(["data.dat", "--precision", precision] + (["--quiet"] if quiet else []) + law, tmp_path)
    say(f"L-HySEA.exe ({precision}) rc={r.returncode}: {r.stdout.strip().splitlines()[-1] if r.stdout else ''}")
    assert r.returncode == 0, r.stderr
    H, _ = S.hmin_shift(sc.H)
    prm = params_of(dump, L.FP64 if precision == "fp64" else L.FP32C)
    frames, times, emax, _ = library_frames(prm, np.ascontiguousarray(sc.state), np.ascontiguousarray(H),
                                            dump["tiempo_tot"], dump["tiempo_guardar"])
    with netcdf_file(str(tmp_path / "res_eta1.nc"), "r", mmap=False) as f:
        eta = f.variables["eta1"][:].copy()
        t_nc = f.variables["time"][:].copy()
        mh = f.variables["max_height"][:].copy()
        assert f.friction_law == (b"Pouliquen" if pouliquen else b"Coulomb")
    with netcdf_file(str(tmp_path / "res_q2x.nc"), "r", mmap=False) as f:
        q2x = f.variables["q2x"][:].copy()
    say(f"  frames {len(t_nc)} at {t_nc.tolist()}; library {len(frames)} at {[round(t, 4) for t in times]}")
    assert len(t_nc) == len(frames) >= 4
    np.testing.assert_array_equal(t_nc, np.asarray(times, dtype=np.float32))
    # the files hold eta in the original datum: h1 + h2 - H with the unshifted
    # depth (the library steps in the Hmin-shifted frame)
    for k, W in enumerate(frames):
        np.testing.assert_allclose(eta[k], (W[0] + W[3] - sc.H).astype(np.float32), rtol=0, atol=1e-4)
        np.testing.assert_allclose(q2x[k], W[4].astype(np.float32), rtol=1e-5, atol=1e-6)
    assert np.abs(q2x[-1]).max() > 0                 # the slide moves
    np.testing.assert_allclose(mh, (emax + H - sc.H).astype(np.float32), rtol=0, atol=1e-4)


def test_cli_gauge_text_run_matches_library(tmp_path):
    sc = S.make_scenario(64, 48, seed=1, semi_axes=(12.0, 9.0))

    # gauge positions in the reference's fixed lon/lat projection
    # (src/GPU/Problema.cxx:326-327), aimed at two wet cells
    def lonlat(i, j):
        return i * sc.dx * 0.6 / 63259.0 - 67.7, j * sc.dy * 0.5 / 55604.0 + 18.3
    cells = [(40, 30), (20, 10)]
    with open(tmp_path / "gauges.txt", "w") as f:
        f.write(f"{len(cells)}\n")
        for (i, j) in cells:
            lo, la = lonlat(i, j)
            f.write(f"{lo!r} {la!r}\n")
    write_case(str(tmp_path), sc, prefix="g", t_end=6.0, t_save=1.0, angle_deg=2.0, gauges_file="gauges.txt")
    dump = json.loads(cli(["data.dat", "--dump"], tmp_path).stdout)
    r = cli(["data.dat", "--quiet"], tmp_path)
    assert r.returncode == 0, r.stderr
    rows = [list(map(float, ln.split())) for ln in (tmp_path / "g_eta_puntos.txt").read_text().splitlines()
            if ln.strip()]
    H, _ = S.hmin_shift(sc.H)
    frames, times, _, _ = library_frames(params_of(dump), np.ascontiguousarray(sc.state), np.ascontiguousarray(H),
                                         dump["tiempo_tot"], dump["tiempo_guardar"])
    say(f"gauge text: {len(rows)} rows, library {len(frames)} frames")
    assert len(rows) == len(frames) >= 4
    for row, W, t in zip(rows, frames, times):
        assert row[0] == pytest.approx(t, rel=1e-6, abs=1e-6)
        for k, (i, j) in enumerate(cells):
            # the reference's gauge value (h1 - H - Hmin) * H_scale: h1 - H with the unshifted depth
            assert row[1 + k] == pytest.approx(W[0, j, i] - sc.H[j, i], abs=1e-6)
