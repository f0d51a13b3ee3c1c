"""Output files through the C-ABI (lhysea_output_open / write / close; host code, no GPU).

The layout follows the reference's writer (src/GPU/netcdf.cu:29-411) and
what shallowWater puts in it (src/GPU/ShallowWater.cu:276-331, 462-474):
eta1 = (h1 + h2 - H - Hmin) * H_scale, eta2 = h2 * H_scale, discharges * Q,
times * T, the bathymetry (H + Hmin) * H_scale, max_height
(eta_max - Hmin) * H_scale, and the gauge text line "%e" of the time then
" %.8e" of (h1 - H - Hmin) * H_scale per gauge (-999 outside the grid).
Values are chosen so that every scale and shift changes the result; the
files are read back with scipy's NetCDF reader.
"""
import numpy as np
import pytest

from landslidehysea_amd import solver as L

NX, NY = 7, 5
SCALES = dict(Hmin=-0.25, Q=3.0, T=5.0, xmin=2.0, ymin=-1.0, sim_time=12.0)
HSCALE, LSCALE = 40.0, 100.0


def setup():
    rng = np.random.default_rng(3)
    p = L.Params(nx=NX, ny=NY, dx=0.5, dy=0.25, H=HSCALE, L=LSCALE, angle_deg=(13.0,), vmax1=0.3, vmax2=0.2,
                 precision=L.FP64)
    H = 1.0 + rng.random((NY, NX))
    W = np.zeros((6, NY, NX))
    W[0] = rng.random((NY, NX))
    W[1], W[2] = rng.standard_normal((2, NY, NX))
    W[3] = 0.5 * rng.random((NY, NX))
    W[4], W[5] = rng.standard_normal((2, NY, NX))
    return p, H, W


def test_netcdf_records_and_attributes(tmp_path):
    from scipy.io import netcdf_file
    p, H, W = setup()
    prefix = str(tmp_path / "run")
    out = L.Output(p, H, prefix, grid_name="grid.bat", save_interval=0.5, **SCALES)
    W2 = W.copy()
    W2[0] += 0.125
    out.write(W, 0.0)
    out.write(W2, 0.75)
    em = W2[0] + W2[3] - H
    out.close(em)
    hm, Q, T = SCALES["Hmin"], SCALES["Q"], SCALES["T"]
    with netcdf_file(prefix + "_eta1.nc", "r", mmap=False) as f:
        eta = f.variables["eta1"][:]
        assert eta.shape == (2, NY, NX)
        for r, Wr in enumerate((W, W2)):
            np.testing.assert_array_equal(eta[r], ((Wr[0] + Wr[3] - H - hm) * HSCALE).astype(np.float32))
        np.testing.assert_array_equal(f.variables["time"][:], np.float32([0.0, 0.75 * T]))
        np.testing.assert_array_equal(f.variables["bathymetry"][:], ((H + hm) * HSCALE).astype(np.float32))
        np.testing.assert_array_equal(f.variables["max_height"][:], ((em - hm) * HSCALE).astype(np.float32))
        np.testing.assert_allclose(f.variables["lon"][:], (SCALES["xmin"] + (np.arange(NX) + 0.5) * 0.5) * LSCALE,
                                   rtol=1e-6)
        np.testing.assert_allclose(f.variables["lat"][:], (SCALES["ymin"] + (np.arange(NY) + 0.5) * 0.25) * LSCALE,
                                   rtol=1e-6)
        assert f.grid_name == b"grid.bat" and f.friction_law == b"Coulomb"
        assert f.angle_of_repose == b"13.0000"
        assert float(f.simulation_time) == pytest.approx(SCALES["sim_time"] * T)
        assert float(f.max_speed_water) == pytest.approx(0.3 * Q / HSCALE, rel=1e-6)
        assert float(f.friction_between_layers) == pytest.approx(p.mfc / LSCALE, rel=1e-6)
    for k, name, scale in ((1, "q1x", Q), (2, "q1y", Q), (3, "eta2", HSCALE), (4, "q2x", Q), (5, "q2y", Q)):
        with netcdf_file(f"{prefix}_{name}.nc", "r", mmap=False) as f:
            v = f.variables[name][:]
            np.testing.assert_array_equal(v[0], (W[k] * scale).astype(np.float32))
            np.testing.assert_array_equal(v[1], (W2[k] * scale).astype(np.float32))
            assert "bathymetry" not in f.variables and "max_height" not in f.variables


def test_pouliquen_attributes(tmp_path):
    from scipy.io import netcdf_file
    from dataclasses import replace
    p, H, W = setup()
    p = replace(p, friction_law=1, angle_deg=(12.0, 25.0, 13.5, 30.0))
    out = L.Output(p, H, str(tmp_path / "pq"), **SCALES)
    out.write(W, 0.0)
    out.close(None)
    with netcdf_file(str(tmp_path / "pq_eta1.nc"), "r", mmap=False) as f:
        assert f.friction_law == b"Pouliquen"
        assert f.angle_of_repose == b"12.0000, 25.0000, 13.5000, 30.0000"
        # no eta max given: max_height keeps the fill value
        assert np.all(f.variables["max_height"][:] == 0.0) or np.all(f.variables["max_height"][:] == -1e30)


def test_gauge_text(tmp_path):
    p, H, W = setup()
    gauges = np.array([3, -1, NX * NY - 1], dtype=np.int32)
    prefix = str(tmp_path / "g")
    out = L.Output(p, H, prefix, gauges=gauges, **SCALES)
    out.write(W, 0.0)
    out.write(W, 1.5)
    out.close(None)
    lines = open(prefix + "_eta_puntos.txt").read().splitlines()
    assert len(lines) == 2
    hm = SCALES["Hmin"]
    h1, Hf = W[0].ravel(), H.ravel()
    for line, t in zip(lines, (0.0, 1.5)):
        f = line.split()
        assert f[0] == "%e" % (t * SCALES["T"])
        assert f[1] == "%.8e" % ((h1[3] - Hf[3] - hm) * HSCALE)
        assert f[2] == "-999"
        assert f[3] == "%.8e" % ((h1[-1] - Hf[-1] - hm) * HSCALE)


def test_output_errors(tmp_path):
    p, H, W = setup()
    with pytest.raises(L.LhyseaError) as e:
        L.Output(p, H, str(tmp_path / "g"), gauges=np.array([NX * NY], dtype=np.int32))
    assert e.value.code == L.ERR_ARGUMENT
    with pytest.raises(L.LhyseaError) as e:
        L.Output(p, H, str(tmp_path / "absent_dir" / "x"))
    assert e.value.code == 6
