"""lhysea_step_device (steps without device allocation) against lhysea_step on a handle (GPU).

The no-allocation entry point runs the same fused kernels on caller-owned
device buffers with its scratch in a per-GPU static array; its results must
equal the handle's bit for bit: state, time and next dt, in both precisions.
Three calls in a row on one GPU: a 1024 x 1024 grid (the split launch, class
tables in the scratch array), then a smaller 300 x 200 grid (all tiles
general), then the first grid again on torch's current stream, so that what
a call leaves in the array cannot change the next call's result.  Device
memory comes from torch (allocated by the test, not by the entry point).
"""
import numpy as np
import pytest
import torch

from tests.cases import case
from landslidehysea_amd import solver as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def split32(W):
    hi = W.astype(np.float32)
    lo = (W - hi.astype(np.float64)).astype(np.float32)
    return np.ascontiguousarray(np.concatenate([hi.reshape(6, -1), lo.reshape(6, -1)]))


def merge32(X):
    return X[:6].astype(np.float64) + X[6:].astype(np.float64)


def by_handle(prm, W, H, steps):
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.step(steps)
        Wr = sim.download()
        t, dt, n = sim.time()
        split = sim.tile_classes()[2]
    return Wr, t, dt, split


def by_device_call(prm, W, H, steps, on_torch_stream):
    dev = torch.device("cuda:0")
    Hd = torch.from_numpy(np.ascontiguousarray(H)).to(dev)
    if prm.precision == L.FP64:
        Wd = torch.from_numpy(np.ascontiguousarray(W.reshape(6, -1))).to(dev)
    else:
        Wd = torch.from_numpy(split32(W)).to(dev)
    Wk = torch.empty_like(Wd)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream if on_torch_stream else None
    t, dt, in_work = L.step_device(prm, Hd.data_ptr(), Wd.data_ptr(), Wk.data_ptr(), steps, stream=stream)
    torch.cuda.synchronize()
    R = (Wk if in_work else Wd).cpu().numpy()
    Wr = R.reshape(6, prm.ny, prm.nx) if prm.precision == L.FP64 else merge32(R).reshape(6, prm.ny, prm.nx)
    return Wr, t, dt, in_work


@pytest.mark.parametrize("precision", [L.FP64, L.FP32C], ids=["fp64", "fp32c"])
def test_step_device_equals_handle_steps(precision):
    calls = [(1024, 1024, 9, 3, False), (300, 200, 6, 7, False), (1024, 1024, 9, 3, True)]
    for nx, ny, steps, seed, on_stream in calls:
        sc, W, H, prm = case(nx, ny, seed=seed, angle=2.0, n_slides=2, semi_axes=(nx / 20, ny / 20),
                             precision=precision)
        say(f"step_device precision={precision} {nx}x{ny} steps={steps} torch stream={on_stream}")
        Wr, tr, dtr, split = by_handle(prm, W, H, steps)
        Wd, td, dtd, in_work = by_device_call(prm, W, H, steps, on_stream)
        say(f"  handle split={split} result in work buffer={in_work} max|dW|={float(np.abs(Wd - Wr).max()):.3e}")
        assert in_work == steps % 2
        assert np.array_equal(Wd, Wr)
        assert td == tr and dtd == dtr


def test_step_device_rejects_oversized_grids():
    # 65536 class tiles at most: FP64 tiles are 30 x 14 cells
    prm = L.Params(nx=30 * 300, ny=14 * 220, dx=1.0, dy=1.0)
    dev = torch.device("cuda:0")
    buf = torch.zeros(1, dtype=torch.float64, device=dev)
    with pytest.raises(L.LhyseaError) as e:
        L.step_device(prm, buf.data_ptr(), buf.data_ptr() + 8, buf.data_ptr() + 16, 1)
    assert e.value.code == 3 and "65536" in str(e.value)
