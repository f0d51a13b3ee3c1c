"""Shared test cases: one spec builds both the oracle's and the library's parameters."""
from __future__ import annotations

import numpy as np

from landslidehysea_amd import scenario as S
from landslidehysea_amd.solver import Params, FP64


def case(nx, ny, seed=1, angle=10.0, borde=(-1.0, -1.0, -1.0, -1.0), n_slides=1, thickness=(40.0, 40.0),
         semi_axes=(60.0, 40.0), precision=FP64, **kw):
    sc = S.make_scenario(nx, ny, seed=seed, n_slides=n_slides, thickness=thickness, semi_axes=semi_axes)
    H, hmin = S.hmin_shift(sc.H)
    W = np.ascontiguousarray(sc.state)
    prm = Params(nx=nx, ny=ny, dx=sc.dx, dy=sc.dy, borde=borde, angle_deg=(angle,), precision=precision, **kw)
    return sc, W, np.ascontiguousarray(H), prm


def oracle_params(binding, prm: Params):
    a = tuple(prm.angle_deg)
    if len(a) == 1:
        a = a * 4
    return binding.make_params(prm.nx, prm.ny, prm.dx, prm.dy, borde=prm.borde, cfl=prm.cfl, r=prm.r,
                               mfc=prm.mfc, mf0=prm.mf0, mfs=prm.mfs, vmax1=prm.vmax1, vmax2=prm.vmax2,
                               g=prm.g, eps_h=prm.eps_h, L=prm.L, H=prm.H, friction_law=prm.friction_law,
                               angles_deg=a)


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))
