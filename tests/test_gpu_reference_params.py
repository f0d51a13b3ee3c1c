"""The reference's own sample parameters and the speed caps, on the GPU, against the oracle.

* Port Valdez's data.dat parameters (the reference's only sample case:
  normalised units with g = 1, eps_h = 5e-5, CFL 0.9, r 0.44, Coulomb 13 deg,
  vmax 120 / 180 m/s, open boundaries; src/GPU/Problema.cxx:226-296) on a
  synthetic 21-degree shelf where the slide flows, with the sample's caps and
  with caps low enough to act (0.3 / 0.2 m/s): the FP64 fused and four-pass
  steps within relative L2 1e-10 of the oracle, dt history to 1e-12.
* The speed caps on fast fronts thinner than eps_h (a dam break of water, then
  of sediment, onto a dry bed; src/GPU/Volumen_kernel.cu:150-179): FP64 within
  1e-10; FP32C within 1e-6 on h, 1e-5 on q and 1e-6 (relative to max |q|) in
  the thin cells.  Scaling q by vmax/u without the h s(h) factor moves q by
  6e-5 to 9e-5 relative and the thin cells by 7e-5 to 1e-4 relative
  (test_oracle_plants.py checks that margin on the CPU).
"""
from dataclasses import replace

import numpy as np
import pytest

from tests.cases import oracle_params, pvaldez_case, rel_l2, thin_front_case
from landslidehysea_amd import solver as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("kernel", [L.KERNEL_FUSED, L.KERNEL_FOURPASS])
@pytest.mark.parametrize("vmax", [(120.0, 180.0), (0.3, 0.2)], ids=["sample_caps", "low_caps"])
def test_pvaldez_parameters_fp64_match_oracle(oracle, kernel, vmax):
    steps = 300
    sc, W, H, prm = pvaldez_case(96, 64, vmax=vmax)
    say(f"pvaldez parameters 96x64 steps={steps} vmax={vmax} kernel={kernel} eps_h={prm.eps_h} g={prm.g}")
    po = oracle_params(oracle, prm)
    Wo = W.copy()
    res = oracle.run(po, Wo, H, steps)
    with L.Simulation(prm, kernel=kernel) as sim:
        sim.upload(W, H)
        sim.step(steps)
        Wg = sim.download()
        t, dtn, n = sim.time()
        hist = sim.dt_history(steps)
    errs = {name: rel_l2(Wg[k], Wo[k]) for k, name in enumerate(("h1", "q1x", "q1y", "h2", "q2x", "q2y"))}
    errs["eta"] = rel_l2(Wg[0] + Wg[3] - H, Wo[0] + Wo[3] - H)
    say("rel L2:", errs)
    assert all(v <= 1e-10 for v in errs.values()), errs
    assert n == steps and abs(t - res["t"]) <= 1e-12 * res["t"]
    assert np.allclose(hist, res["dt_hist"], rtol=1e-12, atol=0)
    assert np.abs(Wo[3] - W[3]).max() > 0.1 * W[3].max()        # the slide flows at 13 deg


@pytest.mark.parametrize("precision", [L.FP64, L.FP32C], ids=["fp64", "fp32c"])
@pytest.mark.parametrize("layer", [1, 2])
def test_speed_caps_on_thin_fronts(oracle, layer, precision):
    W, H, prm, steps = thin_front_case(layer, precision=precision)
    say(f"thin front layer {layer} precision {precision} vmax {prm.vmax1, prm.vmax2} eps_h {prm.eps_h}")
    po = oracle_params(oracle, prm)
    Wo = W.copy()
    oracle.run(po, Wo, H, steps)
    Wu = W.copy()                                   # the caps matter in this case
    oracle.run(oracle_params(oracle, replace(prm, vmax1=0.0, vmax2=0.0)), Wu, H, steps)
    k = 0 if layer == 1 else 3
    assert rel_l2(Wu[k + 1], Wo[k + 1]) > 1e-2
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.step(steps)
        Wg = sim.download()
    thin = (Wo[k] > 0) & (Wo[k] < prm.eps_h)
    eh, eq = rel_l2(Wg[k], Wo[k]), rel_l2(Wg[k + 1], Wo[k + 1])
    ethin = float(np.abs(Wg[k + 1] - Wo[k + 1])[thin].max() / np.abs(Wo[k + 1]).max())
    say(f"h {eh:.3e} q {eq:.3e} thin cells {int(thin.sum())}: max|dq|/max|q| {ethin:.3e}")
    assert thin.sum() >= 6
    if precision == L.FP64:
        assert eh <= 1e-10 and eq <= 1e-10 and ethin <= 1e-10
    else:
        assert eh <= 1e-6 and eq <= 1e-5 and ethin <= 1e-6
