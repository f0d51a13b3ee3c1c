"""Writers for the reference's input files: data.dat, .bat and .ini.

Formats follow the reference loader (src/GPU/Problema.cxx:189-287 for the
positional data.dat grammar, :215-221 and :429-456 for the .bat text file
"xmin xmax ymin ymax nx ny" followed by nx*ny depths, row 0 first, and
:486-537 for the .ini text file of nx*ny records "h1 q1x q1y h2 q2x q2y").
"""
from __future__ import annotations

import os

import numpy as np

from .scenario import Scenario


def write_bat(path: str, sc: Scenario):
    with open(path, "w") as f:
        xmax = sc.xmin + sc.nx * sc.dx
        ymax = sc.ymin + sc.ny * sc.dy
        f.write(f"{sc.xmin!r} {xmax!r} {sc.ymin!r} {ymax!r} {sc.nx} {sc.ny}\n")
        np.savetxt(f, sc.H.reshape(-1, sc.nx), fmt="%.17g")


def write_ini(path: str, sc: Scenario):
    recs = np.stack([sc.h1.ravel(), sc.q1x.ravel(), sc.q1y.ravel(), sc.h2.ravel(), sc.q2x.ravel(),
                     sc.q2y.ravel()], axis=1)
    np.savetxt(path, recs, fmt="%.17g")


def write_case(directory: str, sc: Scenario, *, name="synthetic", prefix="out", borde=(-1, -1, -1, -1),
               t_end=10.0, t_save=-1.0, cfl=0.5, r=0.5, angle_deg=10.0, mfc=0.08, mf0=0.03, mfs=0.03,
               vmax=(0.0, 0.0), normalise=None, gauges_file=None) -> str:
    """Write data.dat + .bat + .ini for ``sc``; returns the data.dat path.

    normalise = (L, H) selects the reference's normalised units; angle_deg is
    one friction angle (Coulomb) or four (Pouliquen, read with --pouliquen)."""
    os.makedirs(directory, exist_ok=True)
    write_bat(os.path.join(directory, "bathy.bat"), sc)
    write_ini(os.path.join(directory, "state.ini"), sc)
    lines = [name, "1", "bathy.bat", "state.ini"]
    lines += [repr(float(b)) for b in borde]
    lines += [repr(float(t_end)), repr(float(t_save))]
    if gauges_file:
        lines += ["1", gauges_file]
    else:
        lines += ["0"]
    angles = angle_deg if isinstance(angle_deg, (tuple, list)) else (angle_deg,)
    lines += [repr(float(cfl)), repr(float(r))] + [repr(float(a)) for a in angles]
    lines += [repr(float(x)) for x in (mfc, mf0, mfs, vmax[0], vmax[1])]
    if normalise:
        lines += ["1", repr(float(normalise[0])), repr(float(normalise[1]))]
    else:
        lines += ["0"]
    lines += [prefix]
    path = os.path.join(directory, "data.dat")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path
