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


# Port Valdez data.dat (the reference's only sample case; tests/golden/pvaldez_data.dat):
# CFL 0.9, r 0.44, Coulomb 13 deg, mfc 8e-2, mf0 2e-3, mfs 5e-2, vmax 120 / 180 m/s,
# open boundaries, normalised units with L = 1000 m, H = 100 m.
PVALDEZ = dict(cfl=0.9, r=0.44, angle=13.0, mfc=8e-2, mf0=2e-3, mfs=5e-2, vmax=(120.0, 180.0),
               borde=(1.0, 1.0, 1.0, 1.0), L=1000.0, H=100.0)


def pvaldez_case(nx, ny, seed=3, precision=FP64, vmax=None, spacing_deg=1e-4, semi_axes=(14.0, 10.0),
                 thickness=(40.0, 40.0)):
    """A synthetic grid under Port Valdez's parameters, in the reference's
    normalised units (src/GPU/Problema.cxx:267-287: g = 1, Q = sqrt(9.81 H^3),
    mfc * L, mf0 and mfs * (Q/H) sqrt(L) / H^(7/6), vmax / (Q/H),
    eps_h = 5e-3 / H, lengths / L; :436-517: depths and thicknesses / H,
    discharges / Q).  The default spacing (about 11 m) gives a 21 deg shelf,
    so the slide flows at 13 deg.  vmax (m/s) replaces the sample's caps."""
    Lr, Hr = PVALDEZ["L"], PVALDEZ["H"]
    Q = float(np.sqrt(9.81 * Hr ** 3))
    fac = (Q / Hr) * np.sqrt(Lr) / Hr ** (7.0 / 6.0)
    sc = S.make_scenario(nx, ny, seed=seed, semi_axes=semi_axes, thickness=thickness, spacing_deg=spacing_deg)
    H, _ = S.hmin_shift(sc.H / Hr)
    W = np.ascontiguousarray(sc.state / np.array([Hr, Q, Q, Hr, Q, Q]).reshape(6, 1, 1))
    v1, v2 = vmax if vmax is not None else PVALDEZ["vmax"]
    prm = Params(nx=nx, ny=ny, dx=sc.dx / Lr, dy=sc.dy / Lr, borde=PVALDEZ["borde"], cfl=PVALDEZ["cfl"],
                 r=PVALDEZ["r"], angle_deg=(PVALDEZ["angle"],), mfc=PVALDEZ["mfc"] * Lr,
                 mf0=PVALDEZ["mf0"] * fac, mfs=PVALDEZ["mfs"] * fac, vmax1=v1 / (Q / Hr), vmax2=v2 / (Q / Hr),
                 g=1.0, eps_h=5e-3 / Hr, L=Lr, H=Hr, precision=precision)
    return sc, W, np.ascontiguousarray(H), prm


def thin_front_case(layer, precision=FP64, nx=160, ny=6, vmax=0.5, eps_h=0.2):
    """Dam break of one layer (1 water, 2 sediment) onto a dry flat bed, walls
    all round, with the dry threshold eps_h raised to 0.2 m (the reference sets
    eps_h = 5e-3 / H, so it is a case parameter) and that layer's speed cap at
    vmax: the front is a band of cells thinner than eps_h moving faster than
    the cap, where the cap's q = u h (src/GPU/Volumen_kernel.cu:150-179)
    differs from a plain rescaling of q by a factor h s(h) < 1.  Returns
    (W, H, params, steps)."""
    W = np.zeros((6, ny, nx))
    H = np.full((ny, nx), 1.0)
    k = 0 if layer == 1 else 3
    W[k][:, : nx // 3] = 1.0
    prm = Params(nx=nx, ny=ny, dx=0.5, dy=0.5, angle_deg=(1.0,), vmax1=vmax if layer == 1 else 0.0,
                 vmax2=vmax if layer == 2 else 0.0, mf0=0.0, mfs=0.0, mfc=0.0, eps_h=eps_h, precision=precision)
    return np.ascontiguousarray(W), H, prm, 120
