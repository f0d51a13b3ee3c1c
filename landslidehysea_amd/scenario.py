"""Seeded synthetic landslide scenarios.

The reference ships no runnable sample (its Port Valdez inputs are missing), so
every workload here is generated from a documented seed:

* bathymetry: a plane shelf whose elevation falls linearly from +20 m to
  -400 m across x, plus ``n_bumps`` Gaussian bumps (amplitude U[-30, 30] m,
  sigma U[5, 15] cells, centre uniform over the grid);
* sediment: one or more paraboloid caps ``h2 = hmax * max(0, 1 - (di/a)^2 - (dj/b)^2)``;
* water at rest: ``h1 = max(0, H - h2)``, all discharges zero.

Depth ``H`` is positive down (``H = -z``), row 0 is the first file row and the
"superior" boundary, as in the reference loader (src/GPU/Problema.cxx:425-546).
The random stream is splitmix64 so that the same seed gives the same case on
every machine and in every language that implements it.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

DEG_TO_M = 111195.0          # metres per degree used for lon/lat spacing
DEFAULT_SPACING_DEG = 0.0005  # 0.0005 deg, about 55.6 m

_MASK = (1 << 64) - 1


class SplitMix64:
    """splitmix64 generator (Steele, Lea, Flood 2014); doubles from the top 53 bits."""

    def __init__(self, seed: int):
        self.state = seed & _MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & _MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self, lo: float = 0.0, hi: float = 1.0) -> float:
        return lo + (hi - lo) * ((self.next_u64() >> 11) * (1.0 / (1 << 53)))

    def randint(self, lo: int, hi: int) -> int:
        """Integer in [lo, hi] inclusive."""
        return lo + int(self.uniform() * (hi - lo + 1)) if hi > lo else lo


@dataclass
class Slide:
    i0: float          # centre column (cell units)
    j0: float          # centre row
    hmax: float        # maximum thickness, m
    a: float = 60.0    # semi-axis in x, cells
    b: float = 40.0    # semi-axis in y, cells

    def volume(self, dx: float, dy: float) -> float:
        return 0.5 * math.pi * self.hmax * self.a * self.b * dx * dy


@dataclass
class Scenario:
    nx: int
    ny: int
    dx: float
    dy: float
    H: np.ndarray          # (ny, nx) depth, positive down, metres (not yet Hmin-shifted)
    h1: np.ndarray
    h2: np.ndarray
    q1x: np.ndarray
    q1y: np.ndarray
    q2x: np.ndarray
    q2y: np.ndarray
    slides: list = field(default_factory=list)
    seed: int = 0
    xmin: float = 0.0
    ymin: float = 0.0

    @property
    def state(self) -> np.ndarray:
        """(6, ny, nx) planes h1, q1x, q1y, h2, q2x, q2y."""
        return np.stack([self.h1, self.q1x, self.q1y, self.h2, self.q2x, self.q2y])

    def wet_cells(self) -> np.ndarray:
        return np.flatnonzero((self.h1 + self.h2).ravel() > 1.0)


def bathymetry(nx: int, ny: int, rng: SplitMix64, n_bumps: int = 8,
               z_top: float = 20.0, z_drop: float = 420.0) -> np.ndarray:
    """Depth H (ny, nx), positive down: plane shelf plus Gaussian bumps."""
    i = np.arange(nx, dtype=np.float64)[None, :]
    j = np.arange(ny, dtype=np.float64)[:, None]
    z = z_top - z_drop * (i + 0.5) / nx + 0.0 * j
    for _ in range(n_bumps):
        amp = rng.uniform(-30.0, 30.0)
        sig = rng.uniform(5.0, 15.0)
        ci = rng.uniform(0.0, nx)
        cj = rng.uniform(0.0, ny)
        z = z + amp * np.exp(-((i - ci) ** 2 + (j - cj) ** 2) / (2.0 * sig * sig))
    return -z


def paraboloid(nx: int, ny: int, s: Slide) -> np.ndarray:
    i = np.arange(nx, dtype=np.float64)[None, :]
    j = np.arange(ny, dtype=np.float64)[:, None]
    f = 1.0 - ((i - s.i0) / s.a) ** 2 - ((j - s.j0) / s.b) ** 2
    return s.hmax * np.maximum(f, 0.0)


def make_scenario(nx: int, ny: int, seed: int, slides: list | None = None, n_slides: int = 1,
                  thickness: tuple = (40.0, 40.0), semi_axes: tuple = (60.0, 40.0),
                  spacing_deg: float = DEFAULT_SPACING_DEG, n_bumps: int = 8) -> Scenario:
    """Build a seeded scenario.  With ``slides`` None, ``n_slides`` caps are
    placed on the slope: the first at (0.45 nx, ny/2) when n_slides == 1,
    otherwise at seeded positions, with thickness U[thickness]."""
    rng = SplitMix64(seed)
    H = bathymetry(nx, ny, rng, n_bumps=n_bumps)
    if slides is None:
        slides = []
        a, b = semi_axes
        a = min(a, 0.25 * nx)
        b = min(b, 0.25 * ny)
        for k in range(n_slides):
            hmax = rng.uniform(*thickness) if thickness[1] > thickness[0] else thickness[0]
            if n_slides == 1:
                i0, j0 = 0.45 * nx, 0.5 * ny
            else:
                i0 = rng.uniform(0.25 * nx, 0.75 * nx)
                j0 = rng.uniform(0.15 * ny, 0.85 * ny)
            slides.append(Slide(i0, j0, hmax, a, b))
    h2 = np.zeros((ny, nx))
    for s in slides:
        h2 = np.maximum(h2, paraboloid(nx, ny, s))
    h1 = np.maximum(0.0, H - h2)
    z = np.zeros((ny, nx))
    dx = dy = spacing_deg * DEG_TO_M
    return Scenario(nx, ny, dx, dy, H, h1, h2, z.copy(), z.copy(), z.copy(), z.copy(), slides, seed)


def gauge_cells(sc: Scenario, n: int, seed: int) -> np.ndarray:
    """``n`` distinct seeded wet cells (flat row-major indices)."""
    rng = SplitMix64(seed ^ 0x5EED5EED)
    wet = sc.wet_cells()
    out = []
    while len(out) < n and len(out) < len(wet):
        c = int(wet[rng.randint(0, len(wet) - 1)])
        if c not in out:
            out.append(c)
    return np.array(out, dtype=np.int32)


def hmin_shift(H: np.ndarray) -> tuple[np.ndarray, float]:
    """Reference loader's depth shift (src/GPU/Problema.cxx:459-481): if the
    minimum depth is negative, every depth is raised by -Hmin."""
    hmin = float(H.min())
    if hmin >= 0.0:
        return H.copy(), 0.0
    return H - hmin, hmin


def make_ensemble(n: int = 64, seed: int = 64, side: tuple = (256, 1536), volume: tuple = (1e6, 5e7),
                  angle_deg: tuple = (6.0, 16.0), r: tuple = (0.45, 0.6), semi_axes: tuple = (10.0, 30.0)):
    """A seeded ensemble of slide scenarios for batched runs.

    Scenario k: grid sides nx, ny ~ U{side}, its own seeded bathymetry, one
    paraboloid slide with semi-axes ~ U[semi_axes] cells centred at
    (U[0.3, 0.6] nx, U[0.3, 0.7] ny), thickness set by a volume
    V = (pi/2) hmax a b dx dy with log V ~ U[log volume], friction angle
    ~ U[angle_deg] and density ratio r ~ U[r].  Returns (specs, build): the
    drawn parameters of every scenario, and build(spec) -> (Scenario,
    {"angle_deg": ..., "r": ...}) that makes one scenario's arrays at a time."""
    rng = SplitMix64(seed)
    specs = []
    for k in range(n):
        nx = rng.randint(*side)
        ny = rng.randint(*side)
        a = rng.uniform(*semi_axes)
        b = rng.uniform(*semi_axes)
        i0 = rng.uniform(0.3, 0.6) * nx
        j0 = rng.uniform(0.3, 0.7) * ny
        vol = math.exp(rng.uniform(math.log(volume[0]), math.log(volume[1])))
        ang = rng.uniform(*angle_deg)
        rr = rng.uniform(*r)
        specs.append(dict(nx=nx, ny=ny, seed=(seed << 16) + k, a=a, b=b, i0=i0, j0=j0, volume=vol,
                          angle_deg=ang, r=rr))

    def build(spec):
        dx = dy = DEFAULT_SPACING_DEG * DEG_TO_M
        hmax = spec["volume"] / (0.5 * math.pi * spec["a"] * spec["b"] * dx * dy)
        slide = Slide(spec["i0"], spec["j0"], hmax, spec["a"], spec["b"])
        sc = make_scenario(spec["nx"], spec["ny"], spec["seed"], slides=[slide])
        return sc, {"angle_deg": spec["angle_deg"], "r": spec["r"]}

    return specs, build
