"""torch.autograd wrapper of the C-ABI forward/backward pair (gradient checks only).

``GaugeMisfit.apply(h2_init, angle1, sim_spec)`` runs the gauge-misfit loss
of ``lhysea_adjoint`` on the GPU: the forward returns the loss, the backward
returns dL/dh2^0 (per cell) and dL/dangle1 (radians) computed by the discrete
adjoint with binomial checkpointing.  With ``couple_water`` the initial water
depth follows the sediment, h1^0 = max(0, H - Hsurf - h2^0) (water at rest
over the slide), so dL/dh2^0 also collects -dL/dh1^0 where the cell is wet.

Tensors live on the CPU (float64); the simulation itself runs in the HIP
library, so this module is plumbing for tests, not a compute path.
"""
from __future__ import annotations

from dataclasses import dataclass, replace
import math

import numpy as np
import torch

from . import solver as L


@dataclass
class MisfitSpec:
    params: L.Params
    W0: np.ndarray          # (6, ny, nx) initial state
    H: np.ndarray           # (ny, nx) depth
    gauges: np.ndarray      # flat cell indices
    obs: np.ndarray         # (nsteps, ng)
    nsteps: int
    nsnap: int
    couple_water: bool = False


class GaugeMisfit(torch.autograd.Function):
    @staticmethod
    def forward(ctx, h2_init: torch.Tensor, angle1: torch.Tensor, spec: MisfitSpec):
        prm = replace(spec.params, angle_deg=(float(angle1) * 180.0 / math.pi,))
        W = spec.W0.copy()
        h2 = h2_init.detach().cpu().numpy().reshape(W.shape[1:])
        if spec.couple_water:
            surf = spec.W0[0] + spec.W0[3]            # water surface depth below datum at rest
            W[0] = np.maximum(0.0, surf - h2)
        W[3] = h2
        with L.Simulation(prm) as sim:
            sim.upload(W, spec.H)
            res = sim.adjoint(spec.nsteps, spec.nsnap, spec.gauges, spec.obs)
        g = res["grad_W0"]
        g_h2 = g[3].copy()
        if spec.couple_water:
            wet = (spec.W0[0] + spec.W0[3] - h2) > 0.0
            g_h2 = g_h2 - np.where(wet, g[0], 0.0)
        ctx.save_for_backward(torch.from_numpy(g_h2.reshape(h2_init.shape)),
                              torch.tensor(res["grad_angle1"], dtype=torch.float64))
        ctx.stats = res
        return torch.tensor(res["loss"], dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        g_h2, g_ang = ctx.saved_tensors
        return gout * g_h2, gout * g_ang, None
