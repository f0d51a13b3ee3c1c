"""GPU FP64 step vs the FP64 oracle, through the C-ABI.

Bar: relative L2 <= 1e-10 on eta, h2 and the four
discharges; mass conserved to round-off with walls.  The FP64 kernels are
compiled without contraction and keep the reference's evaluation order, so
most cases agree to the last bit; the tolerance below is the stated bar.
"""
import numpy as np
import pytest

from tests.cases import case, oracle_params, rel_l2
from landslidehysea_amd import solver as L

pytestmark = pytest.mark.gpu
TOL = 1e-10


def say(*a):
    print(*a, flush=True)


def compare(Wg, Wo, H, tol=TOL):
    eta_g = Wg[0] + Wg[3] - H
    eta_o = Wo[0] + Wo[3] - H
    errs = {"eta": rel_l2(eta_g, eta_o), "h2": rel_l2(Wg[3], Wo[3])}
    for k, name in ((1, "q1x"), (2, "q1y"), (4, "q2x"), (5, "q2y")):
        # discharges can be identically zero in a static case: compare against the state scale
        den = max(np.linalg.norm(Wo[k]), 1e-300)
        errs[name] = float(np.linalg.norm(Wg[k] - Wo[k]) / den) if np.linalg.norm(Wo[k]) > 0 else \
            float(np.abs(Wg[k]).max())
    bad = {k: v for k, v in errs.items() if not (v <= tol)}
    return errs, bad


def run_pair(oracle, nx, ny, steps, kernel=L.KERNEL_FUSED, **kw):
    sc, W, H, prm = case(nx, ny, **kw)
    say(f"case {nx}x{ny} steps={steps} kernel={kernel} {kw}")
    po = oracle_params(oracle, prm)
    Wo = W.copy()
    res = oracle.run(po, Wo, H, steps)
    with L.Simulation(prm, kernel=kernel) as sim:
        sim.upload(W, H)
        dt0 = sim.initial_dt()
        sim.step(steps)
        Wg = sim.download()
        t, dtn, n = sim.time()
        hist = sim.dt_history(steps)
    return W, H, Wo, Wg, res, (dt0, t, dtn, n, hist)


@pytest.mark.parametrize("kernel", [L.KERNEL_FUSED, L.KERNEL_FOURPASS])
def test_initial_dt_matches_oracle(oracle, kernel):
    for (nx, ny) in ((64, 48), (37, 23)):
        sc, W, H, prm = case(nx, ny, angle=2.0)
        po = oracle_params(oracle, prm)
        d_o = oracle.initial_dt(po, W, H)
        with L.Simulation(prm, kernel=kernel) as sim:
            sim.upload(W, H)
            d_g = sim.initial_dt()
        say(f"initial dt {nx}x{ny}: gpu {d_g!r} oracle {d_o!r}")
        assert abs(d_g - d_o) <= 1e-13 * d_o


@pytest.mark.parametrize("kernel", [L.KERNEL_FUSED, L.KERNEL_FOURPASS])
@pytest.mark.parametrize("nx,ny,steps,angle,borde", [
    (64, 48, 60, 2.0, (-1.0, -1.0, -1.0, -1.0)),     # moving slide, walls
    (37, 23, 60, 2.0, (1.0, 1.0, 1.0, 1.0)),         # odd sizes, open boundaries
    (61, 46, 40, 10.0, (0.0, -1.0, 1.0, -1.0)),      # mixed boundaries, odd x
])
def test_small_grids_match_oracle(oracle, kernel, nx, ny, steps, angle, borde):
    W, H, Wo, Wg, res, (dt0, t, dtn, n, hist) = run_pair(oracle, nx, ny, steps, kernel=kernel, angle=angle,
                                                         borde=borde, semi_axes=(12.0, 9.0))
    errs, bad = compare(Wg, Wo, H)
    say("rel L2:", errs, "max|dW|:", float(np.abs(Wg - Wo).max()))
    assert not bad, bad
    assert n == steps
    assert abs(t - res["t"]) <= 1e-12 * res["t"]
    assert np.allclose(hist, res["dt_hist"], rtol=1e-12, atol=0)
    assert abs(dtn - res["dt_next"]) <= 1e-10 * res["dt_next"]


def test_fused_equals_fourpass_bitwise(oracle):
    """Both GPU paths evaluate the same edges in the same class order."""
    sc, W, H, prm = case(70, 52, angle=2.0, semi_axes=(14.0, 10.0))
    outs = []
    for kernel in (L.KERNEL_FUSED, L.KERNEL_FOURPASS):
        with L.Simulation(prm, kernel=kernel) as sim:
            sim.upload(W, H)
            sim.step(50)
            outs.append(sim.download())
    say("max |fused - fourpass|:", float(np.abs(outs[0] - outs[1]).max()))
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("angle", [10.0, 2.0])
def test_config1_400x300_matches_oracle(oracle, angle):
    """400x300 plane shelf + 8 bumps + paraboloid cap (literal 10 deg, and a moving 2 deg variant)."""
    steps = 300
    W, H, Wo, Wg, res, (dt0, t, dtn, n, hist) = run_pair(oracle, 400, 300, steps, angle=angle)
    errs, bad = compare(Wg, Wo, H)
    say("rel L2:", errs)
    assert not bad, bad
    # mass conservation with walls (no h < 0 clipping in these runs)
    for k in (0, 3):
        m0, m1 = W[k].sum(), Wg[k].sum()
        assert abs(m1 - m0) <= 1e-12 * m0


def test_config1_literal_2000_steps_six_gauges(oracle):
    """The 400x300 plain-FP64 workload as stated: 10 deg, 2000 steps, 6 seeded
    gauges, eta max tracked; state, gauge series, eta max, dt history and time
    against the oracle."""
    from landslidehysea_amd.scenario import gauge_cells
    steps = 2000
    sc, W, H, prm = case(400, 300, angle=10.0)
    gauges = gauge_cells(sc, 6, seed=7)
    say(f"config 1 literal: 400x300, {steps} steps, gauges {gauges.tolist()}")
    po = oracle_params(oracle, prm)
    Wo = W.copy()
    em_o = np.stack([(W[0] + W[3] - H).ravel(), np.zeros(W[0].size)], axis=1).copy()
    res = oracle.run(po, Wo, H, steps, eta_max=em_o, gauges=gauges)
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.track_eta_max(True)
        sim.set_gauges(gauges, steps)
        sim.step(steps)
        Wg = sim.download()
        ev, et = sim.eta_max()
        gs = sim.gauges()
        t, dtn, n = sim.time()
        hist = sim.dt_history(steps)
    errs, bad = compare(Wg, Wo, H)
    rest = float((W[0] + W[3] - H)[W[0] > 1.0].mean())
    g_err = rel_l2(gs - rest, res["gauges"] - rest)
    say("rel L2:", errs, "gauge signal rel L2", g_err, "eta max", rel_l2(ev.ravel(), em_o[:, 0]))
    assert not bad, bad
    assert gs.shape == (steps, 6) and g_err <= TOL
    assert rel_l2(ev.ravel(), em_o[:, 0]) <= TOL
    assert np.allclose(et.ravel(), em_o[:, 1], rtol=1e-12, atol=0)
    assert n == steps and abs(t - res["t"]) <= 1e-12 * res["t"]
    assert np.allclose(hist, res["dt_hist"], rtol=1e-12, atol=0)


def test_eta_max_and_gauges_match_oracle(oracle):
    sc, W, H, prm = case(96, 64, angle=2.0, semi_axes=(20.0, 14.0))
    from landslidehysea_amd.scenario import gauge_cells
    gauges = gauge_cells(sc, 6, seed=7)
    steps = 80
    po = oracle_params(oracle, prm)
    Wo = W.copy()
    em_o = np.stack([(W[0] + W[3] - H).ravel(), np.zeros(W[0].size)], axis=1).copy()
    res = oracle.run(po, Wo, H, steps, eta_max=em_o, gauges=gauges)
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.track_eta_max(True)
        sim.set_gauges(gauges, steps)
        sim.step(steps)
        ev, et = sim.eta_max()
        gs = sim.gauges()
    say("eta max rel L2", rel_l2(ev.ravel(), em_o[:, 0]), "gauge rel L2", rel_l2(gs, res["gauges"]))
    assert rel_l2(ev.ravel(), em_o[:, 0]) <= TOL
    assert np.allclose(et.ravel(), em_o[:, 1], rtol=1e-12, atol=0)
    assert gs.shape == res["gauges"].shape
    assert rel_l2(gs, res["gauges"]) <= TOL


def test_lake_at_rest_is_exact():
    """Water at rest over bathymetry with emerged land and no sediment stays exactly at rest.

    Depths and the shifted sea level are dyadic (multiples of 1/64), so every
    free-surface difference is exactly zero and the step must return the
    input bit for bit."""
    sc, W, H, prm = case(80, 60)
    H = np.ascontiguousarray(np.round(H * 64.0) / 64.0)
    sea = np.round(-float(sc.H.min()) * 64.0) / 64.0      # shifted sea level
    W[:] = 0.0
    W[0] = np.maximum(0.0, H - sea)
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.step(100)
        Wg = sim.download()
    assert np.array_equal(Wg, W)


def test_shallow_water_one_shot_matches_stepping(oracle):
    sc, W, H, prm = case(48, 40, angle=2.0, semi_axes=(10.0, 8.0))
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.step(25)
        t25 = sim.time()[0]
    Wf, em, et, secs, n = L.shallow_water(prm, W, H, t_end=t25 * 0.999999)
    say("one-shot steps", n, "t_end", t25)
    assert n == 25
    po = oracle_params(oracle, prm)
    Wo = W.copy()
    oracle.run(po, Wo, H, 25)
    errs, bad = compare(Wf.reshape(6, prm.ny, prm.nx), Wo, H)
    assert not bad, bad


@pytest.mark.parametrize("kernel", [L.KERNEL_FUSED, L.KERNEL_FOURPASS])
def test_pouliquen_law_matches_oracle(oracle, kernel):
    """Pouliquen friction (friction_law 1, src/GPU/Arista_kernel.cu:20-58;
    the oracle's law is pinned by closed forms in test_oracle_kat.py) on a
    flowing slide with four distinct angles: relative L2 <= 1e-10."""
    from dataclasses import replace
    steps = 60
    sc, W, H, prm = case(64, 48, angle=2.0, semi_axes=(12.0, 9.0))
    prm = replace(prm, friction_law=1, angle_deg=(2.0, 6.0, 3.0, 8.0))
    Wo = W.copy()
    oracle.run(oracle_params(oracle, prm), Wo, H, steps)
    with L.Simulation(prm, kernel=kernel) as sim:
        sim.upload(W, H)
        sim.step(steps)
        Wg = sim.download()
    errs, bad = compare(Wg, Wo, H)
    say("pouliquen rel L2:", errs)
    assert not bad, bad
    assert np.abs(Wo[4]).max() > 1e-3             # the slide moves
    Wc = W.copy()                                  # and the law matters: Coulomb at angle 2 differs
    oracle.run(oracle_params(oracle, replace(prm, friction_law=0)), Wc, H, steps)
    assert rel_l2(Wc[3], Wo[3]) > 1e-8
