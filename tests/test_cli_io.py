"""Host front end (L-HySEA.exe) on the CPU: data.dat grammar, .bat/.ini readers,
normalisation, the NetCDF writer (read back with scipy) and failure paths.

The Port Valdez data.dat (tests/golden/pvaldez_data.dat, the reference's only
sample case) has 22 lines and no gauge-flag line; the stream grammar reads
"0.9" as the int 0 and then ".9" as the CFL number (src/GPU/Problema.cxx:
232-247).  Its .bat/.ini files are missing from the reference, so synthetic
ones of a small size are written next to it.
"""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from landslidehysea_amd import scenario as S
from landslidehysea_amd.case_io import write_case, write_bat, write_ini

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "landslidehysea_amd", "L-HySEA.exe")
GOLD = os.path.join(ROOT, "tests", "golden")


def run(args, cwd):
    if not os.path.exists(EXE):
        pytest.fail("L-HySEA.exe is not built (run __graft_entry__.build())")
    return subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=120)


def test_pvaldez_stream_grammar(tmp_path):
    shutil.copy(os.path.join(GOLD, "pvaldez_data.dat"), tmp_path / "data.dat")
    sc = S.make_scenario(24, 16, seed=5, semi_axes=(4.0, 3.0))
    write_bat(str(tmp_path / "batimetria.bat"), sc)
    write_ini(str(tmp_path / "HPV_slide.ini"), sc)
    r = run(["data.dat", "--dump"], tmp_path)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    L, H = 1000.0, 100.0
    Q = math.sqrt(9.81 * H ** 3)
    T = L * H / Q
    fac = (Q / H) * math.sqrt(L) / H ** (7.0 / 6.0)
    assert d["nombre_bati"] == "HPV_slide" and d["prefijo"] == "HPV_slide"
    assert d["leer_fichero_puntos"] == 0
    assert d["cfl"] == 0.9 and d["r"] == 0.44 and d["angle1_deg"] == pytest.approx(13.0, rel=1e-15)
    assert d["normalizar"] == 1 and d["L"] == L and d["H"] == H and d["g"] == 1.0
    assert d["Q"] == pytest.approx(Q, rel=1e-15) and d["T"] == pytest.approx(T, rel=1e-15)
    assert d["mfc"] == pytest.approx(8e-2 * L, rel=1e-15)
    assert d["mf0"] == pytest.approx(2e-3 * fac, rel=1e-14)
    assert d["mfs"] == pytest.approx(5e-2 * fac, rel=1e-14)
    assert d["vmax1"] == pytest.approx(120.0 / (Q / H), rel=1e-15)
    assert d["vmax2"] == pytest.approx(180.0 / (Q / H), rel=1e-15)
    assert d["tiempo_tot"] == pytest.approx(1.0 / T, rel=1e-15)
    assert d["tiempo_guardar"] == pytest.approx(0.5 / T, rel=1e-15)
    assert d["borde"] == [1.0, 1.0, 1.0, 1.0]
    assert d["eps_h"] == pytest.approx(5e-3 / H, rel=1e-15)
    assert (d["nx"], d["ny"]) == (24, 16)
    assert d["dx"] == pytest.approx(sc.dx / L, rel=1e-12)
    assert d["Hmin"] == pytest.approx(min(0.0, float(sc.H.min()) / H), rel=1e-12)


def test_missing_input_files_fail(tmp_path):
    shutil.copy(os.path.join(GOLD, "pvaldez_data.dat"), tmp_path / "data.dat")
    r = run(["data.dat", "--dump"], tmp_path)
    assert r.returncode == 1
    assert "batimetria.bat" in r.stderr
    r = run(["absent.dat"], tmp_path)
    assert r.returncode == 1
    r = subprocess.run([EXE], capture_output=True, text=True)
    assert r.returncode == 1 and "Uso" in r.stderr


def test_netcdf_output_reads_back(tmp_path):
    from scipy.io import netcdf_file
    sc = S.make_scenario(30, 20, seed=9, semi_axes=(6.0, 4.0))
    path = write_case(str(tmp_path), sc, name="synthetic_bati", prefix="res", t_save=0.0, angle_deg=12.0)
    r = run([os.path.basename(path), "--dry-run"], tmp_path)
    assert r.returncode == 0, r.stderr
    names = ["eta1", "q1x", "q1y", "eta2", "q2x", "q2y"]
    for v in names:
        assert (tmp_path / f"res_{v}.nc").exists()
    H = sc.H
    hmin = min(0.0, float(H.min()))
    with netcdf_file(str(tmp_path / "res_eta1.nc"), "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions["lon"] == 30 and f.dimensions["lat"] == 20 and f.dimensions["time"] is None
        eta = f.variables["eta1"][:]
        assert eta.shape == (1, 20, 30)
        # eta1 = (h1 + h2 - (H - Hmin) - Hmin) * H_scale = h1 + h2 - H
        np.testing.assert_allclose(eta[0], (sc.h1 + sc.h2 - H).astype(np.float32), rtol=0, atol=1e-4)
        np.testing.assert_allclose(f.variables["bathymetry"][:], H.astype(np.float32), rtol=1e-6)
        np.testing.assert_allclose(f.variables["max_height"][:], (sc.h1 + sc.h2 - H).astype(np.float32),
                                   rtol=0, atol=1e-4)
        lon = f.variables["lon"][:]
        np.testing.assert_allclose(lon, (np.arange(30) + 0.5) * sc.dx, rtol=1e-6)
        assert f.variables["time"][:].tolist() == [0.0]
        assert f.variables["eta1"].units == b"meters"
        assert f.friction_law == b"Coulomb"
        assert float(f.CFL) == pytest.approx(0.5)
        assert f.grid_name == b"synthetic_bati"
        assert f.angle_of_repose == b"12.0000"
    with netcdf_file(str(tmp_path / "res_eta2.nc"), "r", mmap=False) as f:
        np.testing.assert_allclose(f.variables["eta2"][0], sc.h2.astype(np.float32), rtol=1e-6)
        assert "bathymetry" not in f.variables
    with netcdf_file(str(tmp_path / "res_q1x.nc"), "r", mmap=False) as f:
        assert f.variables["q1x"].long_name == b"Mass flow of water along x"
        assert not f.variables["q1x"][0].any()
    assert hmin <= 0.0


def test_gauge_text_output(tmp_path):
    sc = S.make_scenario(30, 20, seed=9, semi_axes=(6.0, 4.0))
    # the reference projects lon/lat with fixed Port Valdez constants
    # (src/GPU/Problema.cxx:326-327); invert them to aim at two cells
    def lonlat(i, j):
        x, y = i * sc.dx, j * sc.dy
        return x * 0.6 / 63259.0 - 67.7, y * 0.5 / 55604.0 + 18.3
    with open(tmp_path / "gauges.txt", "w") as f:
        f.write("3\n")
        for (i, j) in ((3, 4), (10, 12)):
            lo, la = lonlat(i, j)
            f.write(f"{lo!r} {la!r}\n")
        f.write("10.0 10.0\n")          # outside the grid
    path = write_case(str(tmp_path), sc, prefix="g", t_save=0.0, gauges_file="gauges.txt")
    r = run([os.path.basename(path), "--dump"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert json.loads(r.stdout)["gauges"] == 3
    r = run([os.path.basename(path), "--dry-run"], tmp_path)
    assert r.returncode == 0, r.stderr
    line = (tmp_path / "g_eta_puntos.txt").read_text().split()
    assert len(line) == 4 and line[3] == "-999"
    H = sc.H
    for k, (i, j) in enumerate(((3, 4), (10, 12))):
        # reference gauge value: (h1 - H - Hmin) * H_scale in the shifted frame = h1 - H
        assert float(line[1 + k]) == pytest.approx(sc.h1[j, i] - H[j, i], abs=1e-6)


def test_pouliquen_grammar_and_netcdf_attributes(tmp_path):
    """--pouliquen reads four friction angles (src/GPU/Problema.cxx:247-255,
    the reference's build without COULOMB) and the fields after them stay in
    place; the NetCDF files name the law and its angles (src/GPU/netcdf.cu:236-244)."""
    from scipy.io import netcdf_file
    sc = S.make_scenario(30, 20, seed=9, semi_axes=(6.0, 4.0))
    path = write_case(str(tmp_path), sc, prefix="pq", t_save=0.0, angle_deg=(12.0, 25.0, 13.5, 30.0),
                      mfc=0.07, vmax=(3.0, 4.0))
    r = run([os.path.basename(path), "--dump", "--pouliquen"], tmp_path)
    assert r.returncode == 0, r.stderr
    d = json.loads(r.stdout)
    assert d["friction_law"] == 1
    assert d["angles_deg"] == pytest.approx([12.0, 25.0, 13.5, 30.0], rel=1e-14)
    assert d["mfc"] == 0.07 and d["vmax1"] == 3.0 and d["vmax2"] == 4.0 and d["prefijo"] == "pq"
    r = run([os.path.basename(path), "--dry-run", "--pouliquen"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "Angulos de reposo: 12, 25, 13.5, 30" in r.stdout
    with netcdf_file(str(tmp_path / "pq_eta1.nc"), "r", mmap=False) as f:
        assert f.friction_law == b"Pouliquen"
        assert f.angle_of_repose == b"12.0000, 25.0000, 13.5000, 30.0000"
    # the same file read as Coulomb takes the second angle for mfc: the grammar is positional
    r = run([os.path.basename(path), "--dump"], tmp_path)
    d = json.loads(r.stdout)
    assert d["friction_law"] == 0 and d["mfc"] == 25.0
