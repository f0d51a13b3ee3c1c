"""lhysea_run_until: the reference's `while (t < t_end)` loop
(src/GPU/ShallowWater.cu:269) stepped on the device in chunks.

Bar: bit-exact against a host loop that reads t back before every step, for
the FP64 fused and four-pass kernels and the FP32C kernel, with end times that
fall inside a chunk (launched steps past t_end must do nothing) and with a step
budget that stops first."""
import numpy as np
import pytest

from tests.cases import case
from landslidehysea_amd import solver as L

pytestmark = pytest.mark.gpu


def host_loop(prm, W, H, t_end, max_steps, kernel):
    with L.Simulation(prm, kernel=kernel) as sim:
        sim.upload(W, H)
        sim.initial_dt()
        n = 0
        while n < max_steps:
            t, _, _ = sim.time()
            if not (t < t_end):
                break
            sim.step(1)
            n += 1
        return sim.download(), sim.time(), n


@pytest.mark.parametrize("precision,kernel", [(L.FP64, L.KERNEL_FUSED), (L.FP64, L.KERNEL_FOURPASS),
                                              (L.FP32C, L.KERNEL_FUSED)])
def test_run_until_matches_host_loop(precision, kernel):
    sc, W, H, prm = case(50, 37, angle=2.0, semi_axes=(8.0, 6.0), precision=precision)
    with L.Simulation(prm, kernel=kernel) as sim:
        sim.upload(W, H)
        sim.initial_dt()
        sim.step(7)
        t7, dt7, _ = sim.time()
    for t_end, max_steps in ((t7 + 0.37 * dt7, 10**6), (t7 * 3.3, 10**6), (1e9, 23)):
        ref_W, ref_time, ref_n = host_loop(prm, W, H, t_end, max_steps, kernel)
        with L.Simulation(prm, kernel=kernel) as sim:
            sim.upload(W, H)
            n = sim.run_until(t_end, max_steps)
            got_W, got_time = sim.download(), sim.time()
            # stepping continues from the right buffer afterwards
            sim.step(2)
            after = sim.download()
        print(f"t_end {t_end:.4f} budget {max_steps}: steps {n} vs {ref_n}", flush=True)
        assert n == ref_n and got_time == ref_time
        assert np.array_equal(got_W, ref_W)
        with L.Simulation(prm, kernel=kernel) as sim:
            sim.upload(W, H)
            sim.step(ref_n + 2)
            assert np.array_equal(sim.download(), after)
