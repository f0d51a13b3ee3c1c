"""torch.autograd wrappers of the C-ABI adjoint pair (gradient checks only).

``GaugeSeries.apply(h2_init, angles, spec)`` runs ``lhysea_adjoint_forward``
on the GPU and returns the gauge series eta (nsteps, ng); its backward feeds
the incoming gradient dL/deta, for any loss L the caller builds from the
series in torch, to the reverse sweep and returns dL/dh2^0 (per cell) and
the friction-angle gradient (radians) in the shape of ``angles``: a scalar
angle1, or under the Pouliquen law the four angles (shape (4,)).
``GaugeMisfit.apply`` is the one-call form for the gauge misfit
L = sum (eta - obs)^2.

With ``couple_water`` the initial water depth follows the sediment,
h1^0 = max(0, H - Hsurf - h2^0) (water at rest over the slide), so dL/dh2^0
also collects -dL/dh1^0 where the cell is wet.

Tensors live on the CPU (float64); the simulation itself runs in the HIP
library, so this module is plumbing for tests, not a compute path.
"""
from __future__ import annotations

from dataclasses import dataclass, replace
import math
from typing import Optional

import numpy as np
import torch

from . import solver as L


@dataclass
class MisfitSpec:
    params: L.Params
    W0: np.ndarray          # (6, ny, nx) initial state
    H: np.ndarray           # (ny, nx) depth
    gauges: np.ndarray      # flat cell indices
    obs: Optional[np.ndarray]  # (nsteps, ng); only the misfit form reads it
    nsteps: int
    nsnap: int
    couple_water: bool = False


def _start(spec: MisfitSpec, h2_init: torch.Tensor, angles: torch.Tensor):
    """Parameters and initial state for sediment depth h2_init and the friction
    angle(s): a scalar angle1, or the four angles of the Pouliquen law."""
    a = angles.detach().cpu().numpy().reshape(-1)
    prm = replace(spec.params, angle_deg=tuple(float(x) * 180.0 / math.pi for x in a))
    W = spec.W0.copy()
    h2 = h2_init.detach().cpu().numpy().reshape(W.shape[1:])
    if spec.couple_water:
        surf = spec.W0[0] + spec.W0[3]            # water surface depth below datum at rest
        W[0] = np.maximum(0.0, surf - h2)
    W[3] = h2
    return prm, W, h2


def _grad_angles(res: dict, angles: torch.Tensor) -> torch.Tensor:
    """dL/dangle in the shape of the angle argument (radians)."""
    if angles.numel() == 4:
        return torch.tensor(np.asarray(res["grad_angles"]), dtype=torch.float64).reshape(angles.shape)
    return torch.tensor(res["grad_angle1"], dtype=torch.float64).reshape(angles.shape)


def _grad_h2(spec: MisfitSpec, h2: np.ndarray, g: np.ndarray, shape) -> torch.Tensor:
    """dL/dh2^0 from dL/dW0, through h1^0(h2^0) when the water is coupled."""
    g_h2 = g[3].copy()
    if spec.couple_water:
        wet = (spec.W0[0] + spec.W0[3] - h2) > 0.0
        g_h2 = g_h2 - np.where(wet, g[0], 0.0)
    return torch.from_numpy(g_h2.reshape(shape))


class GaugeSeries(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h2_init: torch.Tensor, angles: torch.Tensor, spec: MisfitSpec):
        prm, W, h2 = _start(spec, h2_init, angles)
        sim = L.Simulation(prm)
        try:
            sim.upload(W, spec.H)
            eta = sim.adjoint_forward(spec.nsteps, spec.nsnap, spec.gauges)
        except Exception:
            sim.close()
            raise
        # the handle keeps the checkpoints until the backward
        ctx.sim, ctx.spec, ctx.h2, ctx.shape, ctx.angles = sim, spec, h2, h2_init.shape, angles.detach()
        return torch.from_numpy(eta)

    @staticmethod
    def backward(ctx, grad_eta):
        sim = ctx.sim
        try:
            res = sim.adjoint_backward(grad_eta.detach().cpu().numpy())
        finally:
            sim.close()
            ctx.sim = None
        ctx.stats = res
        g_h2 = _grad_h2(ctx.spec, ctx.h2, res["grad_W0"], ctx.shape)
        return g_h2, _grad_angles(res, ctx.angles), None


class GaugeMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h2_init: torch.Tensor, angles: torch.Tensor, spec: MisfitSpec):
        prm, W, h2 = _start(spec, h2_init, angles)
        with L.Simulation(prm) as sim:
            sim.upload(W, spec.H)
            res = sim.adjoint(spec.nsteps, spec.nsnap, spec.gauges, spec.obs)
        ctx.save_for_backward(_grad_h2(spec, h2, res["grad_W0"], h2_init.shape), _grad_angles(res, angles))
        ctx.stats = res
        return torch.tensor(res["loss"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        g_h2, g_ang = ctx.saved_tensors
        return gout * g_h2, gout * g_ang, None
