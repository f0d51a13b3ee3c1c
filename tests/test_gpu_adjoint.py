"""Discrete adjoint on the GPU vs central finite differences of the FP64 oracle.

The loss is the gauge misfit L = sum_n sum_g (eta_g^n - obs_g^n)^2 with
observations from an oracle run of a thicker slide.  Finite differences use
the oracle with the dt sequence frozen at the unperturbed run (the adjoint
differentiates with dt and every branch held at the forward choices).

Stated bars: FP64 adjoint within 1e-4 relative of the FD gradient; the FP32
adjoint of the compensated path within 2e-2 (its forward already differs
from FP64 by FP32 round-off amplified through the branchy slide dynamics).
Finite-difference cells are chosen inside the slide (h2 > 5 m): at h = 0 the
model is not differentiable (sqrt(h) in the Roe average).  The reference's
scheme is branchy (Coulomb stop, wet/dry, positivity limiter), so the loss has
kinks and jumps: where the one-sided differences agree the adjoint is
checked against the central difference; at a branch point it must equal the
one-sided derivative of one side (the side whose branches the forward took).
"""
import ctypes as C
import math
from dataclasses import replace

import numpy as np
import pytest
import torch

from tests.cases import case, oracle_params
from landslidehysea_amd import solver as L
from landslidehysea_amd.autograd import GaugeMisfit, MisfitSpec
from landslidehysea_amd.scenario import gauge_cells

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


# the Pouliquen law's four angles (stop / start pairs) on the flowing slide
PQ_ANGLES = (2.0, 6.0, 3.0, 8.0)


def setup(oracle, precision, nx=48, ny=36, N=40, pouliquen=False):
    sc, W, H, prm = case(nx, ny, angle=2.0, semi_axes=(10.0, 7.0), precision=precision)
    if pouliquen:
        prm = replace(prm, friction_law=1, angle_deg=PQ_ANGLES)
    gauges = gauge_cells(sc, 6, seed=3)
    po = oracle_params(oracle, prm)
    Wp = W.copy()
    Wp[3] *= 1.1
    Wp[0] = np.maximum(0.0, (W[0] + W[3]) - Wp[3])
    obs = oracle.run(po, Wp, H, N, gauges=gauges)["gauges"]
    r0 = oracle.run(po, W.copy(), H, N, gauges=gauges)
    return sc, W, H, prm, po, gauges, obs, r0["dt_hist"]


def fd_loss(oracle, po, W, H, dts, gauges, obs):
    g = oracle.run_fixed_dt(po, W.copy(), H, dts, gauges)
    return float(((g - obs) ** 2).sum())


@pytest.mark.parametrize("precision,tol", [(L.FP64, 1e-4), (L.FP32C, 2e-2)])
def test_adjoint_matches_oracle_fd(oracle, precision, tol):
    N = 40
    sc, W, H, prm, po, gauges, obs, dts = setup(oracle, precision, N=N)
    say(f"adjoint precision={precision} 48x36 N={N} snapshots=8")
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        res = sim.adjoint(N, 8, gauges, obs)
    say("loss", res["loss"], "advances", res["advances"], "reverse steps", res["reverse_steps"])
    assert res["reverse_steps"] == N
    L0 = fd_loss(oracle, po, W, H, dts, gauges, obs)
    assert abs(res["loss"] - L0) <= (1e-9 if precision == L.FP64 else 5e-2) * L0
    rng = np.random.default_rng(5)
    cells = rng.choice(np.flatnonzero(W[3].ravel() > 5.0), 8, replace=False)
    nx = prm.nx
    L0 = fd_loss(oracle, po, W, H, dts, gauges, obs)
    smooth = 0
    for c in cells:
        j, i = divmod(int(c), nx)
        for k in (0, 3):
            eps = 1e-5 * max(1.0, W[k, j, i])
            Wp, Wm = W.copy(), W.copy()
            Wp[k, j, i] += eps
            Wm[k, j, i] -= eps
            lp = fd_loss(oracle, po, Wp, H, dts, gauges, obs)
            lm = fd_loss(oracle, po, Wm, H, dts, gauges, obs)
            fp, fm, fc = (lp - L0) / eps, (L0 - lm) / eps, (lp - lm) / (2 * eps)
            ad = res["grad_W0"][k, j, i]
            say(f"cell {c} field {k}: adjoint {ad:.6e} central {fc:.6e} one-sided {fp:.6e} {fm:.6e}")
            if abs(fp - fm) <= 1e-3 * abs(fc) + 1e-12:
                smooth += 1
                assert abs(ad - fc) <= tol * abs(fc) + 1e-12
            else:
                # a branch (Coulomb stop, wet/dry, limiter) flips on one side of this
                # state: the frozen-branch adjoint is the derivative of the other side
                assert min(abs(ad - fp), abs(ad - fm)) <= max(tol, 1e-3) * max(abs(fp), abs(fm)) + 1e-12
    assert smooth >= 8
    da = 1e-6
    pp, pm = oracle_params(oracle, prm), oracle_params(oracle, prm)
    for p, s in ((pp, 1), (pm, -1)):
        p.angle1 += s * da
        p.angle2 += s * da
        p.angle3 += s * da
        p.angle4 += s * da
    fd = (fd_loss(oracle, pp, W, H, dts, gauges, obs) - fd_loss(oracle, pm, W, H, dts, gauges, obs)) / (2 * da)
    say(f"angle: adjoint {res['grad_angle1']:.6e} fd {fd:.6e}")
    assert abs(res["grad_angle1"] - fd) <= tol * abs(fd)


def test_autograd_wrapper_gradients(oracle):
    """torch.autograd.Function over the C-ABI pair; gradients checked against oracle FD.

    Coupled mode (water at rest over the slide: h1 = surface - h2) must equal
    dL/dh2 - dL/dh1 of the uncoupled run where the cell is wet."""
    N = 30
    sc, W, H, prm, po, gauges, obs, dts = setup(oracle, L.FP64, N=N)
    obs = obs[:N]
    grads = {}
    for couple in (False, True):
        spec = MisfitSpec(prm, W, H, gauges, obs, N, 6, couple_water=couple)
        h2 = torch.tensor(W[3], dtype=torch.float64, requires_grad=True)
        ang = torch.tensor(math.radians(2.0), dtype=torch.float64, requires_grad=True)
        loss = GaugeMisfit.apply(h2, ang, spec)
        loss.backward()
        grads[couple] = (float(loss), h2.grad.numpy().copy(), float(ang.grad))
        say("autograd loss", float(loss), "coupled" if couple else "uncoupled")
    L0 = fd_loss(oracle, po, W, H, dts, gauges, obs)
    assert abs(grads[False][0] - L0) <= 1e-9 * L0
    assert grads[False][0] == grads[True][0]
    rng = np.random.default_rng(9)
    checked = 0
    for c in rng.choice(np.flatnonzero(W[3].ravel() > 5.0), 6, replace=False):
        j, i = divmod(int(c), prm.nx)
        eps = 1e-5 * W[3, j, i]
        Wp, Wm = W.copy(), W.copy()
        Wp[3, j, i] += eps
        Wm[3, j, i] -= eps
        lp = fd_loss(oracle, po, Wp, H, dts, gauges, obs)
        lm = fd_loss(oracle, po, Wm, H, dts, gauges, obs)
        fp, fm, fc = (lp - L0) / eps, (L0 - lm) / eps, (lp - lm) / (2 * eps)
        ad = grads[False][1][j, i]
        say(f"cell {c}: autograd {ad:.6e} central {fc:.6e} one-sided {fp:.6e} {fm:.6e}")
        if abs(fp - fm) <= 1e-3 * abs(fc) + 1e-12:
            checked += 1
            assert abs(ad - fc) <= 1e-4 * abs(fc) + 1e-12
    assert checked >= 3
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        res = sim.adjoint(N, 6, gauges, obs)
    wet = (W[0] + W[3] - W[3]) > 0.0
    expect = res["grad_W0"][3] - np.where(wet, res["grad_W0"][0], 0.0)
    np.testing.assert_allclose(grads[True][1], expect, rtol=1e-12, atol=1e-18)
    assert grads[True][2] == pytest.approx(res["grad_angle1"], rel=1e-12)


def test_checkpoint_schedule_is_binomial():
    """Re-advance counts of the checkpoint tree equal the binomial optimum."""
    sc, W, H, prm = case(24, 18, angle=2.0, semi_axes=(5.0, 4.0), precision=L.FP32C)
    gauges = gauge_cells(sc, 3, seed=1)
    for N, S in ((17, 3), (40, 5), (25, 1)):
        obs = np.zeros((N, len(gauges)))
        with L.Simulation(prm) as sim:
            sim.upload(W, H)
            res = sim.adjoint(N, S, gauges, obs)
        chain = res["optimal_advances"] - res["advances"]
        say(f"N={N} S={S}: re-advances {res['advances']} optimal total {res['optimal_advances']}")
        assert res["reverse_steps"] == N
        assert 0 <= chain < N


def test_all_states_kept_equals_checkpointing():
    """nsnap > nsteps keeps every state in HBM (no re-advance): same loss and
    bit-identical gradients as the binomial schedule."""
    sc, W, H, prm = case(40, 30, angle=2.0, semi_axes=(8.0, 6.0), precision=L.FP32C)
    gauges = gauge_cells(sc, 4, seed=3)
    N = 24
    obs = np.zeros((N, len(gauges)))
    out = {}
    for snaps in (5, N + 1):
        with L.Simulation(prm) as sim:
            sim.upload(W, H)
            out[snaps] = sim.adjoint(N, snaps, gauges, obs)
    a, b = out[5], out[N + 1]
    say("advances", a["advances"], b["advances"])
    assert b["advances"] == 0 and b["reverse_steps"] == N
    assert a["loss"] == b["loss"]
    assert np.array_equal(a["grad_W0"], b["grad_W0"]) and a["grad_mu"] == b["grad_mu"]


def test_split_pair_equals_misfit_call():
    """lhysea_adjoint_forward + seeds 2 (eta - obs) + lhysea_adjoint_backward is
    the one-call misfit adjoint, bit for bit; the forward's gauge series is the
    plain stepping loop's."""
    N = 30
    for precision in (L.FP64, L.FP32C):
        sc, W, H, prm = case(40, 30, angle=2.0, semi_axes=(8.0, 6.0), precision=precision)
        gauges = gauge_cells(sc, 5, seed=4)
        rng = np.random.default_rng(11)
        obs = 1e-3 * rng.standard_normal((N, len(gauges)))
        say(f"split adjoint precision={precision} 40x30 N={N}")
        with L.Simulation(prm) as sim:
            sim.upload(W, H)
            sim.initial_dt()
            sim.set_gauges(gauges, N)
            sim.step(N)
            plain = sim.gauges()
        with L.Simulation(prm) as sim:
            sim.upload(W, H)
            one = sim.adjoint(N, 7, gauges, obs)
        with L.Simulation(prm) as sim:
            sim.upload(W, H)
            eta = sim.adjoint_forward(N, 7, gauges)
            two = sim.adjoint_backward(2.0 * (eta - obs))
        assert np.array_equal(eta, plain)
        assert float(((eta - obs) ** 2).sum()) == pytest.approx(one["loss"], rel=1e-15)
        assert np.array_equal(one["grad_W0"], two["grad_W0"])
        assert one["grad_mu"] == two["grad_mu"] and one["grad_angle1"] == two["grad_angle1"]
        assert one["advances"] == two["advances"] and two["reverse_steps"] == N


def test_pending_forward_refuses_steps():
    """Between the pair's calls the handle refuses stepping, keeps the state
    after the last step for download, and a second forward replaces the first."""
    N = 12
    sc, W, H, prm = case(24, 18, angle=2.0, semi_axes=(5.0, 4.0), precision=L.FP32C)
    gauges = gauge_cells(sc, 3, seed=1)
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.initial_dt()
        sim.step(N)
        after = sim.download()
    with L.Simulation(prm) as sim:
        with pytest.raises(L.LhyseaError):
            sim.adjoint_backward(np.zeros((N, len(gauges))))
        sim.upload(W, H)
        sim.adjoint_forward(N, 4, gauges)
        for call in (lambda: sim.step(1), lambda: sim.run_until(1e9, 1), lambda: sim.upload(W, H),
                     lambda: sim.set_dt(0.1), lambda: sim.set_gauges(gauges, 4)):
            with pytest.raises(L.LhyseaError):
                call()
        assert np.array_equal(sim.download(), after)
        first = sim.adjoint_backward(np.ones((N, len(gauges))))
        with pytest.raises(L.LhyseaError):
            sim.adjoint_backward(np.ones((N, len(gauges))))
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.adjoint_forward(N, 4, gauges)
        # a second forward discards the pending one, then finds the state after
        # N steps, not a fresh upload; the handle is free again afterwards
        with pytest.raises(L.LhyseaError):
            sim.adjoint_forward(N, 4, gauges)
        sim.upload(W, H)
        sim.adjoint_forward(N, 4, gauges)
        again = sim.adjoint_backward(np.ones((N, len(gauges))))
        assert np.array_equal(again["grad_W0"], first["grad_W0"])
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        sim.adjoint_forward(N, 4, gauges)
    # destroy with a pending forward frees its workspace (no leak, no error)
    assert np.isfinite(first["grad_W0"]).all()


def test_series_autograd_any_loss_vs_oracle_fd(oracle):
    """GaugeSeries: the torch loss L = sum(w * eta) + 0.5 sum(eta_last^2), not the
    misfit, differentiated by the C-ABI pair vs central differences of the FP64
    oracle (dt frozen), bar 1e-4 relative on smooth cells and on angle1."""
    from landslidehysea_amd.autograd import GaugeSeries
    N = 30
    sc, W, H, prm, po, gauges, obs, dts = setup(oracle, L.FP64, N=N)
    w = np.random.default_rng(21).standard_normal((N, len(gauges)))

    def loss_np(series):
        return float((w * series).sum() + 0.5 * (series[-1] ** 2).sum())

    spec = MisfitSpec(prm, W, H, gauges, None, N, 6)
    h2 = torch.tensor(W[3], dtype=torch.float64, requires_grad=True)
    ang = torch.tensor(math.radians(2.0), dtype=torch.float64, requires_grad=True)
    eta = GaugeSeries.apply(h2, ang, spec)
    loss = (torch.from_numpy(w) * eta).sum() + 0.5 * (eta[-1] ** 2).sum()
    loss.backward()
    L0 = loss_np(oracle.run_fixed_dt(po, W.copy(), H, dts, gauges))
    say("series loss", loss.item(), "oracle", L0)
    assert abs(loss.item() - L0) <= 1e-9 * abs(L0)
    rng = np.random.default_rng(13)
    checked = 0
    for c in rng.choice(np.flatnonzero(W[3].ravel() > 5.0), 6, replace=False):
        j, i = divmod(int(c), prm.nx)
        eps = 1e-5 * W[3, j, i]
        Wp, Wm = W.copy(), W.copy()
        Wp[3, j, i] += eps
        Wm[3, j, i] -= eps
        lp = loss_np(oracle.run_fixed_dt(po, Wp, H, dts, gauges))
        lm = loss_np(oracle.run_fixed_dt(po, Wm, H, dts, gauges))
        fp, fm, fc = (lp - L0) / eps, (L0 - lm) / eps, (lp - lm) / (2 * eps)
        ad = float(h2.grad[j, i])
        say(f"cell {c}: autograd {ad:.6e} central {fc:.6e} one-sided {fp:.6e} {fm:.6e}")
        if abs(fp - fm) <= 1e-3 * abs(fc) + 1e-12:
            checked += 1
            assert abs(ad - fc) <= 1e-4 * abs(fc) + 1e-12
    assert checked >= 3
    da = 1e-6
    pp, pm = oracle_params(oracle, prm), oracle_params(oracle, prm)
    for p, s in ((pp, 1), (pm, -1)):
        p.angle1 += s * da
        p.angle2 += s * da
        p.angle3 += s * da
        p.angle4 += s * da
    fd = (loss_np(oracle.run_fixed_dt(pp, W.copy(), H, dts, gauges))
          - loss_np(oracle.run_fixed_dt(pm, W.copy(), H, dts, gauges))) / (2 * da)
    say(f"angle: autograd {float(ang.grad):.6e} fd {fd:.6e}")
    assert abs(float(ang.grad) - fd) <= 1e-4 * abs(fd)


@pytest.mark.parametrize("precision,tol", [(L.FP64, 1e-4), (L.FP32C, 2e-2)])
def test_pouliquen_adjoint_matches_oracle_fd(oracle, precision, tol):
    """The Pouliquen law (src/GPU/Arista_kernel.cu:20-58) in the adjoint:
    dL/dW0 at smooth slide cells and dL/dangle of each of the four angles
    against the oracle's central differences, at the bars of the Coulomb
    test; 8 snapshots equal every state kept."""
    N = 40
    sc, W, H, prm, po, gauges, obs, dts = setup(oracle, precision, N=N, pouliquen=True)
    say(f"Pouliquen adjoint precision={precision} 48x36 N={N}, angles {PQ_ANGLES}")
    out = []
    for nsnap in (8, N + 1):
        with L.Simulation(prm) as sim:
            sim.upload(W, H)
            out.append(sim.adjoint(N, nsnap, gauges, obs))
    res = out[0]
    assert res["grad_mu"] is None and res["reverse_steps"] == N
    assert np.array_equal(out[0]["grad_W0"], out[1]["grad_W0"])
    assert np.array_equal(out[0]["grad_angles"], out[1]["grad_angles"])
    L0 = fd_loss(oracle, po, W, H, dts, gauges, obs)
    assert abs(res["loss"] - L0) <= (1e-9 if precision == L.FP64 else 5e-2) * L0
    rng = np.random.default_rng(5)
    smooth = 0
    for c in rng.choice(np.flatnonzero(W[3].ravel() > 5.0), 8, replace=False):
        j, i = divmod(int(c), prm.nx)
        for k in (0, 3):
            eps = 1e-5 * max(1.0, W[k, j, i])
            Wp, Wm = W.copy(), W.copy()
            Wp[k, j, i] += eps
            Wm[k, j, i] -= eps
            lp = fd_loss(oracle, po, Wp, H, dts, gauges, obs)
            lm = fd_loss(oracle, po, Wm, H, dts, gauges, obs)
            fp, fm, fc = (lp - L0) / eps, (L0 - lm) / eps, (lp - lm) / (2 * eps)
            ad = res["grad_W0"][k, j, i]
            if abs(fp - fm) <= 1e-3 * abs(fc) + 1e-12:
                smooth += 1
                assert abs(ad - fc) <= tol * abs(fc) + 1e-12, (c, k, ad, fc)
            else:
                assert min(abs(ad - fp), abs(ad - fm)) <= max(tol, 1e-3) * max(abs(fp), abs(fm)) + 1e-12
    assert smooth >= 8
    # each angle alone; a step whose one-sided differences disagree straddles a
    # branch (a Coulomb stop / go flip somewhere in the 40 steps: at 1e-6 rad
    # for angle1 on this slide), so a smaller step is tried, and where both
    # straddle one the adjoint must equal the derivative of one side
    fds = []
    for a in ("angle1", "angle2", "angle3", "angle4"):
        for da in (1e-6, 1e-7):
            pp, pm = oracle_params(oracle, prm), oracle_params(oracle, prm)
            setattr(pp, a, getattr(pp, a) + da)
            setattr(pm, a, getattr(pm, a) - da)
            lp = fd_loss(oracle, pp, W, H, dts, gauges, obs)
            lm = fd_loss(oracle, pm, W, H, dts, gauges, obs)
            fp, fm, fc = (lp - L0) / da, (L0 - lm) / da, (lp - lm) / (2 * da)
            if abs(fp - fm) <= 1e-3 * abs(fc) + 1e-12:
                break
        fds.append((fc, fp, fm))
    say("dL/dangle adjoint", list(res["grad_angles"]), "fd (central, forward, backward)", fds)
    big = max(abs(f[0]) for f in fds)
    checked = 0
    for g, (fc, fp, fm) in zip(res["grad_angles"], fds):
        if max(abs(fc), abs(fp), abs(fm)) < 1e-5 * big:
            # at the differences' noise (~1e-7 here): the adjoint is as small
            assert abs(g) <= 1e-4 * big, (g, fc)
        elif abs(fp - fm) > 1e-3 * abs(fc) + 1e-12:
            assert min(abs(g - fp), abs(g - fm)) <= tol * max(abs(fp), abs(fm)), (g, fp, fm)
        else:
            checked += 1
            assert abs(g - fc) <= tol * abs(fc), (g, fc)
    assert checked >= 2
    assert res["grad_angle1"] == res["grad_angles"][0]


def test_pouliquen_adjoint_pair_and_torch_wrapper(oracle):
    """Under the Pouliquen law: lhysea_adjoint_backward refuses dL/dmu (the
    Coulomb coefficient's) and keeps the workspace, which
    lhysea_adjoint_backward_angles then reverses; the torch wrapper with a
    (4,) angle tensor returns the same four gradients."""
    N = 30
    sc, W, H, prm, po, gauges, obs, dts = setup(oracle, L.FP64, N=N, pouliquen=True)
    obs = obs[:N]
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        ref = sim.adjoint(N, 6, gauges, obs)
    with L.Simulation(prm) as sim:
        sim.upload(W, H)
        eta = sim.adjoint_forward(N, 6, gauges)
        seeds = np.ascontiguousarray(2.0 * (eta - obs))
        gmu = C.c_double()
        rc = sim.lib.lhysea_adjoint_backward(sim.h, seeds.ctypes.data_as(C.POINTER(C.c_double)), None,
                                             C.byref(gmu), None, None)
        assert rc == 3, rc     # LHYSEA_ERR_ARGUMENT
        assert "Pouliquen" in sim.lib.lhysea_last_error().decode()
        r = sim.adjoint_backward(seeds)     # the workspace is still there
    assert np.array_equal(r["grad_W0"], ref["grad_W0"]) and np.array_equal(r["grad_angles"], ref["grad_angles"])
    spec = MisfitSpec(prm, W, H, gauges, obs, N, 6)
    h2 = torch.tensor(W[3], dtype=torch.float64, requires_grad=True)
    ang = torch.tensor([math.radians(a) for a in PQ_ANGLES], dtype=torch.float64, requires_grad=True)
    loss = GaugeMisfit.apply(h2, ang, spec)
    loss.backward()
    say("torch dL/dangles", ang.grad.tolist())
    assert ang.grad.shape == (4,)
    assert np.array_equal(ang.grad.numpy(), ref["grad_angles"])
    assert np.array_equal(h2.grad.numpy(), ref["grad_W0"][3])


@pytest.mark.parametrize("pouliquen", [False, True])
def test_fp32c_fused_reverse_equals_fourpass(oracle, pouliquen):
    """The FP32C reverse step's fused tile kernel (one 38 x 20 tile and its
    2-cell cone per workgroup, everything in LDS) against its four-pass form
    (cell adjoint, then one launch per edge class with the partial-sum
    adjoints in HBM): the same generated edge and cell adjoints in the same
    class order with the same float and double accumulations.  The compiler
    fuses multiplies and adds across statements per kernel, so the two
    kernels round some FP32 products differently: dL/dW0 agrees to 1e-5 of
    its largest component (6.2e-7 measured on the MI355X; 1e-6 under the CPU
    emulation), dL/dangle to 1e-6.  A wrong index, class order or dropped
    term gives differences of order one."""
    N = 40
    sc, W, H, prm, po, gauges, obs, dts = setup(oracle, L.FP32C, N=N, pouliquen=pouliquen)
    res = {}
    for mode in (L.KERNEL_FUSED, L.KERNEL_FOURPASS):
        with L.Simulation(prm) as sim:
            sim.set_adjoint_kernel(mode)
            sim.upload(W, H)
            res[mode] = sim.adjoint(N, 8, gauges, obs)
    a, b = res[L.KERNEL_FUSED], res[L.KERNEL_FOURPASS]
    assert a["loss"] == b["loss"]
    gmax = np.abs(a["grad_W0"]).max()
    dw = float(np.abs(a["grad_W0"] - b["grad_W0"]).max() / gmax)
    da = float(np.abs(a["grad_angles"] - b["grad_angles"]).max() / np.abs(a["grad_angles"]).max())
    say(f"fused vs four-pass reverse ({'Pouliquen' if pouliquen else 'Coulomb'}): dW0 {dw:.3e}, dangles {da:.3e}")
    assert gmax > 0 and dw <= 1e-5 and da <= 1e-6
